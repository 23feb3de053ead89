"""Seeded inputs of the golden fixtures (shared by make_golden.py and the tests)."""
import numpy as np

from cps_veloc_amd import lattice as lat

CASES = {
    # the CPU-runnable plain DWF configuration: 8^3x8, Ls 8, M5 1.8, mf 0.04,
    # links reunitarize(exp(i 0.3 H)), Gaussian source, stop 1e-10
    "dwf_8x8x8x8_ls8_cg": dict(action="dwf", dims=(8, 8, 8, 8), ls=8, mass=0.04, eps=0.3,
                               gauge_seed=1234, src_seed=5678, cg=True),
    "dwf_4x4x4x4_ls4_mdagm": dict(action="dwf", dims=(4, 4, 4, 4), ls=4, mass=0.1, eps=0.3,
                                  gauge_seed=11, src_seed=12, cg=False),
    "mobius_4x4x4x4_ls4_mdagm": dict(action="mobius", dims=(4, 4, 4, 4), ls=4, mass=0.1,
                                     eps=0.3, gauge_seed=13, src_seed=14, cg=False),
}


def inputs(c):
    U = lat.random_gauge(c["dims"], c["eps"], c["gauge_seed"])
    vh = int(np.prod(c["dims"])) // 2
    src = lat.to_reals(lat.gaussian_spinor((c["ls"], vh), c["src_seed"]))
    return U, src
