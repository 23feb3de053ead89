#!/usr/bin/env python3
"""Benchmark of the even-odd preconditioned Mobius MatPcDagMatPc and CG.

Default workload: 24^3x64, Ls=12 Mobius (b=1.5, c=0.5, M5=1.8, mf=0.005),
links reunitarize(exp(i*0.4*H)), seeded unit Gaussian source, FP64.
One step = one MatPcDagMatPc application on HBM-resident vectors; the value
is its algorithmic HBM throughput (include/roofline.h byte model).  The line
also carries the time to a 1e-10 CG solve, the dominant kernel's roofline and
the CPU oracle timed on the host cores.

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Time-to-1e-10 solve vs plain FP64 CG (same GPU); MatPcDagMatPc HBM GB/s"
HBM_PEAK_GBS = 8000.0
SLOPPY = ("f32", "bf16", "i16")
KERNELS = {
    "dwf": ["dwf_hop_cb1", "dwf_hop_cb0_xpay_norm", "dwf_hop_cb1_dag", "dwf_hop_cb0_xpay_dag"],
    "mobius": ["mobius_hop_m5inv", "mobius_hop_schur_norm", "mobius_hop_m5inv_dag",
               "mobius_hop_schur_dag"],
}

WORKLOADS = {
    "mobius_24x24x24x64_ls12": dict(action="mobius", dims=(24, 24, 24, 64), ls=12, mass=0.005,
                                    eps=0.4, b=1.5, c=0.5),
    "mobius_16x16x16x32_ls16": dict(action="mobius", dims=(16, 16, 16, 32), ls=16, mass=0.01,
                                    eps=0.3, b=1.5, c=0.5),
    "dwf_16x16x16x32_ls16": dict(action="dwf", dims=(16, 16, 16, 32), ls=16, mass=0.01, eps=0.3),
    "dwf_8x8x8x8_ls8": dict(action="dwf", dims=(8, 8, 8, 8), ls=8, mass=0.04, eps=0.3),
    # 24 RHS: 12 spin-colour point sources at the origin (P+ on s = 0, P- on
    # s = Ls-1) and 12 complex Z2 noise sources, masses 0.005 / 0.01 / 0.04
    "dwf_16x16x16x32_ls16_mrhs24": dict(action="dwf", dims=(16, 16, 16, 32), ls=16, mass=0.01,
                                        eps=0.3, nrhs=24, masses=(0.005, 0.01, 0.04)),
    "mobius_16x16x16x32_ls16_mrhs24": dict(action="mobius", dims=(16, 16, 16, 32), ls=16,
                                           mass=0.01, eps=0.3, b=1.5, c=0.5, nrhs=24,
                                           masses=(0.005, 0.01, 0.04)),
}
for _name, _wl in WORKLOADS.items():
    _wl["name"] = _name


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profile(workload: str, precision: str):
    """Memory traffic per launch of the hop kernels from the committed rocprof
    PMC summary (profiles/<workload>_<precision>_traffic.json), if present."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def site_errors(got, want, dgot=None, dwant=None):
    g, w = np.asarray(got).reshape(-1, 24), np.asarray(want).reshape(-1, 24)
    nw = np.maximum(np.linalg.norm(w, axis=1), 1e-300)
    out = {"max_site_rel_err": float(np.max(np.linalg.norm(g - w, axis=1) / nw))}
    if dgot is not None:
        out["dot_rel_err"] = float(abs(dgot - dwant) / abs(dwant))
    return out


def cpu_baseline(wl, U, src):
    """CPS's CPU path as the oracle restates it (oracle/cps_oracle.c), timed on
    the host cores: one MatPcDagMatPc of the bench workload with the 4D hop in
    one thread as in CPS's noarch wilson_dslash.C:280-477 (the 5D terms and
    vector arithmetic on OpenMP threads, as CPS's *_omp.C / vector_util.C), and
    one with the hop spread over OpenMP threads.  Returns the timings and the
    output of the serial run (the same bits as the parallel one)."""
    from oracle import oracle as orc
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = orc.Oracle(wl["dims"], wl["ls"], wl["action"], m5=1.8, b=wl.get("b", 1.5),
                   c=wl.get("c", 0.5), gauge_canonical=U)
    o.matpcdagmatpc(src, wl["mass"])  # warm (page faults, OpenMP start)
    t0 = time.perf_counter()
    o.matpcdagmatpc(src, wl["mass"])
    t_par = time.perf_counter() - t0
    orc.set_serial_hop(True)
    try:
        t0 = time.perf_counter()
        want, dwant = o.matpcdagmatpc(src, wl["mass"])
        t_ser = time.perf_counter() - t0
    finally:
        orc.set_serial_hop(False)
    return {"serial_hop_s": t_ser, "omp_s": t_par, "cores": cores, "out": want, "dot": dwant}


def c1_solves():
    """The 8^3x8, Ls 8 DWF InvCg (mf 0.04, stop 1e-10): the CPS-faithful CPU
    solve (serial 4D hop) and the GPU's plain FP64 CG on the same system."""
    from cps_veloc_amd import lattice as lat
    from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu
    from oracle import oracle as orc
    wl = WORKLOADS["dwf_8x8x8x8_ls8"]
    dims, ls, mass = wl["dims"], wl["ls"], wl["mass"]
    U = lat.random_gauge(dims, wl["eps"], 1234)
    src = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 5678))
    o = orc.Oracle(dims, ls, "dwf", m5=1.8, gauge_canonical=U)
    orc.set_serial_hop(True)
    try:
        t0 = time.perf_counter()
        xo, oit, otr, _ = o.invcg(src, mass, stop_rsd=1e-10, max_iter=5000)
        t_cpu = time.perf_counter() - t0
    finally:
        orc.set_serial_hop(False)
    op = DiracOpDwfGpu(U, dims, ls, m5=1.8)
    vin, vsol = op.vector(src), op.vector()
    op.InvCg_dev(vsol, vin, CgArg(mass, 3, 1e-10))  # first-use allocations
    vsol.upload(np.zeros(op.n))
    t0 = time.perf_counter()
    git, gtr, gst = op.InvCg_dev(vsol, vin, CgArg(mass, 5000, 1e-10))
    t_gpu = time.perf_counter() - t0
    err = float(np.linalg.norm(vsol.download() - xo) / np.linalg.norm(xo))
    vin.free()
    vsol.free()
    op.close()
    return {"workload": "dwf_8x8x8x8_ls8", "mass": mass, "stop_rsd": 1e-10,
            "cpu_seconds": round(t_cpu, 3), "cpu_iters": oit, "cpu_true_res": otr,
            "gpu_seconds": round(t_gpu, 4), "gpu_iters": git, "gpu_true_res": gtr,
            "gpu_status": gst, "rel_err_gpu_vs_cpu_solution": err}


def mrhs_full_sources(wl, seed=99):
    """The full 5D sources of the batched workload in CPS's DWF order
    [5D-odd | 5D-even]: 12 spin-colour point sources at the origin (spins 0, 1
    on s = 0 and spins 2, 3 on s = Ls-1, as Ffour2five(src, 0, Ls-1) places
    them, QPropW.C:1204-1205) and complex Z2 noise on every site; masses
    cycle through wl["masses"]."""
    rng = np.random.default_rng(seed)
    ls = wl["ls"]
    vh = int(np.prod(wl["dims"])) // 2
    firsts, seconds, masses = [], [], []
    for j in range(wl["nrhs"]):
        ev = np.zeros((ls, vh, 4, 3, 2))
        od = np.zeros((ls, vh, 4, 3, 2))
        if j < 12:
            spin, col = divmod(j, 3)
            if spin < 2:
                ev[0, 0, spin, col, 0] = 1.0
            else:
                od[ls - 1, 0, spin, col, 0] = 1.0
        else:
            ev = rng.choice([-1.0, 1.0], size=ev.shape) / np.sqrt(2.0)
            od = rng.choice([-1.0, 1.0], size=od.shape) / np.sqrt(2.0)
        firsts.append(od.reshape(-1))
        seconds.append(ev.reshape(-1))
        masses.append(wl["masses"][j % len(wl["masses"])])
    return firsts, seconds, np.array(masses)


def bits_digest(x):
    """XOR and wrapping sum of the 64-bit words of an FP64 array: equal for
    bit-identical arrays (a cheap check, not a hash)"""
    w = np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)
    return int(np.bitwise_xor.reduce(w)), int(np.sum(w, dtype=np.uint64))


def run_batched(wl, sloppy, delta, max_iter, trace=False):
    """DiracOp*::MatInv over all RHS of the workload: one batched solve
    (dwfgpu_matinv_mrhs_dev: sources prepared and solutions completed on the
    device, one graph-replayed CG for all RHS) against the same entry point
    called one RHS at a time, and the batched solve again with plain kernel
    launches instead of graph replays (what a single persistent launch could
    at most save is the difference between launch-bound and GPU-bound).
    trace: every call with plain launches (a kernel trace then covers them)
    and 1 s with the GPU idle on either side of the timed batched call, so
    that its kernels stand apart in a trace.  Returns the record of the
    bench line."""
    from cps_veloc_amd import lattice as lat
    from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu, DiracOpMobiusGpu
    dims, ls = wl["dims"], wl["ls"]
    U = lat.random_gauge(dims, wl["eps"], 1234)
    if wl["action"] == "dwf":
        op = DiracOpDwfGpu(U, dims, ls, m5=1.8)
    else:
        op = DiracOpMobiusGpu(U, dims, ls, m5=1.8, b=wl["b"], c=wl["c"])
    firsts, seconds, masses = mrhs_full_sources(wl)
    nrhs = len(masses)
    inf, ins = op.multi_vector(nrhs, firsts), op.multi_vector(nrhs, seconds)
    outf, outs = op.multi_vector(nrhs), op.multi_vector(nrhs)
    arg = CgArg(0.0, max_iter, 1e-10)
    o1 = [op.multi_vector(1) for _ in range(4)]  # out_first, out_second, in_first, in_second
    # warm-up: autotuner, graph set-up and first-use allocations of both calls
    op.MatInvMrhs_dev(outf, outs, inf, ins, CgArg(0.0, 3, 1e-10), masses, sloppy, delta,
                      eager=trace)
    o1[2].upload(0, firsts[0])
    o1[3].upload(0, seconds[0])
    op.MatInvMrhs_dev(o1[0], o1[1], o1[2], o1[3], CgArg(0.0, 3, 1e-10), masses[:1], sloppy,
                      delta, eager=trace)
    for j in range(nrhs):
        outf.upload(j, np.zeros(op.n))
        outs.upload(j, np.zeros(op.n))
    log(f"[bench] batched MatInv of {nrhs} RHS")
    if trace:
        time.sleep(1.0)
    t0 = time.perf_counter()
    it, tr, st, nrel = op.MatInvMrhs_dev(outf, outs, inf, ins, arg, masses, sloppy, delta,
                                         eager=trace)
    t_batch = time.perf_counter() - t0
    if trace:
        time.sleep(1.0)
    ns, nr = op.mrhs_shape()
    log(f"[bench] batched: {t_batch:.2f}s, iters {it.tolist()}, shape ns={ns} nr={nr}")

    t_seq, it_seq, err, digests = 0.0, [], [], []
    for j in range(nrhs):
        xb = np.concatenate([outf.download(j), outs.download(j)])
        digests.append(bits_digest(xb))
        o1[2].upload(0, firsts[j])
        o1[3].upload(0, seconds[j])
        o1[0].upload(0, np.zeros(op.n))
        o1[1].upload(0, np.zeros(op.n))
        t0 = time.perf_counter()
        i1, _, _, _ = op.MatInvMrhs_dev(o1[0], o1[1], o1[2], o1[3], arg, masses[j:j + 1],
                                        sloppy, delta, eager=trace)
        t_seq += time.perf_counter() - t0
        it_seq.append(int(i1[0]))
        x1 = np.concatenate([o1[0].download(0), o1[1].download(0)])
        err.append(float(np.linalg.norm(xb - x1) / np.linalg.norm(x1)))
    log(f"[bench] one at a time: {t_seq:.2f}s, iters {it_seq}")
    for j in range(nrhs):
        outf.upload(j, np.zeros(op.n))
        outs.upload(j, np.zeros(op.n))
    t0 = time.perf_counter()
    it_e, _, _, _ = op.MatInvMrhs_dev(outf, outs, inf, ins, arg, masses, sloppy, delta, eager=True)
    t_eager = time.perf_counter() - t0
    same = bool(np.array_equal(it_e, it)) and all(
        bits_digest(np.concatenate([outf.download(j), outs.download(j)])) == digests[j]
        for j in range(nrhs))
    log(f"[bench] batched with plain launches: {t_eager:.2f}s (same solutions: {same})")
    for v in [inf, ins, outf, outs] + o1:
        v.free()
    op.close()
    return {"workload": wl["name"], "action": wl["action"], "lattice": list(dims), "ls": ls,
            "nrhs": nrhs, "masses": masses.tolist(), "stop_rsd": 1e-10, "sloppy": sloppy,
            "delta": delta, "shape": {"ns": ns, "nr": nr},
            "batched": {"seconds": round(t_batch, 3), "iters": it.tolist(),
                        "true_res_max": float(np.max(tr)), "status": st.tolist(),
                        "reliable_updates": nrel.tolist()},
            "batched_plain_launches": {"seconds": round(t_eager, 3),
                                       "same_iters_and_solutions": same},
            "one_at_a_time": {"seconds": round(t_seq, 3), "iters": it_seq},
            "rhs_per_second_batched": round(nrhs / t_batch, 3),
            "speedup_batched_vs_one_at_a_time": round(t_seq / t_batch, 3),
            "max_rel_diff_batched_vs_single": max(err)}


def bench_mrhs(args, wl):
    """Bench line of the batched workload alone (--workload ..._mrhs24)."""
    rec = run_batched(wl, args.mrhs_sloppy, args.delta, args.max_iter, args.mrhs_trace)
    out = {
        "metric": "batched 24-RHS MatInv to 1e-10: right-hand sides solved per second",
        "value": rec["rhs_per_second_batched"], "unit": "RHS/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": round(rec["batched"]["seconds"] * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.mrhs_sloppy + "+f64", "data": "synthetic",
        "config": {"workload": args.workload, "lattice": rec["lattice"], "ls": rec["ls"],
                   "masses": rec["masses"], "stop_rsd": 1e-10, "delta": args.delta,
                   "sloppy": args.mrhs_sloppy},
        "batched_matinv": rec,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mobius_24x24x24x64_ls12", choices=sorted(WORKLOADS))
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-iter", type=int, default=20000)
    ap.add_argument("--mixed", nargs="*", default=["f32", "bf16", "i16"],
                    choices=["f32", "bf16", "i16"],
                    help="sloppy precisions of the reliable-update CG solves")
    ap.add_argument("--delta", type=float, default=0.1,
                    help="reliable update when |r| drops by this factor")
    ap.add_argument("--mrhs-sloppy", default="f32", choices=["f32", "bf16", "i16"])
    ap.add_argument("--mrhs-trace", action="store_true",
                    help="batched leg with plain launches only, the timed batched call "
                         "between 1 s idle gaps (for a kernel trace)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 24-RHS batched MatInv leg (DWF 16^3x32, Ls 16)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip training the network preconditioner and its solve")
    ap.add_argument("--learned-sloppy", default="i16", choices=["f32", "bf16", "i16"])
    # training recipe of the network: `solves` Gaussian sources, each giving
    # `pairs-per-solve` pairs from its own CG iterates every `stride` steps
    # (e = x* - x_k, r = A e) plus `random` pairs with Gaussian e.  The
    # default is the quick high-mode recipe; 20 x 100 solver pairs, random 0,
    # is the 2000-pair recipe of the learned-preconditioner workload.
    ap.add_argument("--learned-solves", type=int, default=6)
    ap.add_argument("--learned-pairs-per-solve", type=int, default=0)
    ap.add_argument("--learned-random", type=int, default=32)
    ap.add_argument("--learned-stride", type=int, default=3)
    ap.add_argument("--learned-epochs", type=int, default=3)
    ap.add_argument("--gcr-krylov", type=int, default=8,
                    help="restart length of the flexible GCR solve with the network (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU "
              "sharding of this operator is a design note, not a built path.", file=sys.stderr)
        sys.exit(2)

    from cps_veloc_amd import lattice as lat
    from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu, DiracOpMobiusGpu

    wl = WORKLOADS[args.workload]
    if "nrhs" in wl:
        bench_mrhs(args, wl)
        return
    dims, ls, mass = wl["dims"], wl["ls"], wl["mass"]
    vh = int(np.prod(dims)) // 2
    log(f"[bench] workload {args.workload}: generating inputs")
    t0 = time.perf_counter()
    U = lat.random_gauge(dims, wl["eps"], 1234)
    src = lat.to_reals(lat.gaussian_spinor((ls, vh), 5678))
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f}s; uploading")
    if wl["action"] == "dwf":
        op = DiracOpDwfGpu(U, dims, ls, m5=1.8)
    else:
        op = DiracOpMobiusGpu(U, dims, ls, m5=1.8, b=wl["b"], c=wl["c"])
    vin, vout = op.vector(src), op.vector()
    bytes_app, flops_app = op.cost()

    for _ in range(args.warmup):
        op.MatPcDagMatPc_dev(vout, vin, mass)
    log(f"[bench] warmup done ({args.warmup}); timing {args.steps} applications")
    tw = time.perf_counter()
    ms_app, ms_k, bytes_k = op.time_MatPcDagMatPc(vout, vin, mass, args.steps)
    wall = time.perf_counter() - tw
    gbs = bytes_app / (ms_app * 1e-3) / 1e9
    gflops = flops_app / (ms_app * 1e-3) / 1e9
    log(f"[bench] MatPcDagMatPc {ms_app:.3f} ms  {gbs:.0f} GB/s  {gflops:.0f} GFLOP/s "
        f"(wall {wall:.2f}s); kernels ms {np.round(ms_k, 4).tolist()}")
    v_gpu = vout.download()  # MatPcDagMatPc(src), checked against the oracle below
    d_gpu = op.MatPcDagMatPc(np.zeros(op.n), src, mass)
    hop = int(np.argmax(ms_k[:4]))
    k_gbs = bytes_k[hop] / (ms_k[hop] * 1e-3) / 1e9
    copy_gbs = op.time_copy(2 * 1024 ** 3, 10)
    bw_copy, bw_read, bw_r3w1 = op.time_bandwidth(2 * 1024 ** 3, 10)
    log(f"[bench] dominant kernel #{hop}: {k_gbs:.0f} GB/s; copy kernel {copy_gbs:.0f} GB/s; "
        f"probes: copy {bw_copy:.0f}, read-only {bw_read:.0f}, 3 reads + 1 write "
        f"{bw_r3w1:.0f} GB/s")
    prof = traffic_from_profile(args.workload, "f64")
    kprof = {} if prof is None else prof.get("kernels", {}).get(KERNELS[wl["action"]][hop], {})

    sloppy = {}
    for p in SLOPPY:
        ms_p, msk_p, byk_p = op.time_MatPcDagMatPc(vout, vin, mass, args.steps, precision=p)
        b_p, _ = op.cost(p)
        sloppy[p] = {"ms": round(ms_p, 5), "gbs": round(b_p / (ms_p * 1e-3) / 1e9, 2),
                     "bytes": b_p, "kernels_ms": [round(float(x), 5) for x in msk_p]}
        log(f"[bench] {p} MatPcDagMatPc {ms_p:.3f} ms  {sloppy[p]['gbs']:.0f} GB/s "
            f"(algorithmic bytes in {p})")

    solve = None
    if not args.no_solve:
        arg = CgArg(mass, args.max_iter, 1e-10)
        # warm-up: each solver once for 3 iterations, so that first-use work
        # (work-vector allocation, kernel code loading) stays out of the timings
        warm = CgArg(mass, 3, 1e-10)
        vsol = op.vector()
        op.InvCg_dev(vsol, vin, warm)
        for p in args.mixed:
            vsol.upload(np.zeros(op.n))
            op.InvCgMixed_dev(vsol, vin, warm, p, args.delta)
        vsol.upload(np.zeros(op.n))
        t1 = time.perf_counter()
        it, tr, st = op.InvCg_dev(vsol, vin, arg)
        t_plain = time.perf_counter() - t1
        x_plain = vsol.download()
        nx = float(np.linalg.norm(x_plain))
        solve = {"stop_rsd": 1e-10,
                 "plain_fp64_cg": {"seconds": round(t_plain, 4), "iters": it, "true_res": tr,
                                   "status": st}}
        log(f"[bench] plain CG: {it} iters, true res {tr:.3e}, {t_plain:.2f}s")
        best = None
        for p in args.mixed:
            vsol.upload(np.zeros(op.n))
            t1 = time.perf_counter()
            it, tr, st, nrel = op.InvCgMixed_dev(vsol, vin, arg, p, args.delta)
            ts = time.perf_counter() - t1
            err = float(np.linalg.norm(vsol.download() - x_plain)) / nx
            solve[f"mixed_{p}_cg"] = {"seconds": round(ts, 4), "iters": it, "true_res": tr,
                                      "status": st, "reliable_updates": nrel,
                                      "delta": args.delta, "rel_err_vs_plain_fp64": err,
                                      "speedup_vs_plain_fp64": round(t_plain / ts, 3)}
            log(f"[bench] mixed {p} CG: {it} iters, {nrel} reliable updates, true res "
                f"{tr:.3e}, |x - x_fp64|/|x_fp64| {err:.2e}, {ts:.2f}s "
                f"({t_plain / ts:.2f}x)")
            if st == 0 and tr <= 1e-10 and (best is None or ts < best[1]):
                best = (p, ts)
        if not args.no_learned:
            # the learned preconditioner: train on the GPU, then flexible CG
            log("[bench] training the network preconditioner")
            op.ptc_init(5, 0.02)
            t1 = time.perf_counter()
            l0, l1, npairs = op.ptc_train(mass, nsolves=args.learned_solves,
                                          pairs_per_solve=args.learned_pairs_per_solve,
                                          random_pairs=args.learned_random,
                                          stride=args.learned_stride,
                                          epochs=args.learned_epochs, batch=4, lr=2e-3, seed=17)
            t_train = time.perf_counter() - t1
            vsol.upload(np.zeros(op.n))
            op.InvCgPtc_dev(vsol, vin, warm, args.learned_sloppy, args.delta)
            vsol.upload(np.zeros(op.n))
            t1 = time.perf_counter()
            it, tr, st, nrel, nnet = op.InvCgPtc_dev(vsol, vin, arg, args.learned_sloppy,
                                                     args.delta)
            ts = time.perf_counter() - t1
            err = float(np.linalg.norm(vsol.download() - x_plain)) / nx
            solve["learned_fcg"] = {"seconds": round(ts, 4), "iters": it, "net_steps": nnet,
                                    "true_res": tr, "status": st, "reliable_updates": nrel,
                                    "sloppy": args.learned_sloppy,
                                    "rel_err_vs_plain_fp64": err,
                                    "speedup_vs_plain_fp64": round(t_plain / ts, 3),
                                    "training": {"seconds": round(t_train, 3), "pairs": npairs,
                                                 "loss_first": l0, "loss_last": l1}}
            log(f"[bench] network: trained on {npairs} pairs in {t_train:.1f}s (loss "
                f"{l0:.3f} -> {l1:.3f}); flexible CG {it} iters ({nnet} with the net), "
                f"true res {tr:.3e}, {ts:.2f}s ({t_plain / ts:.2f}x)")
            if st == 0 and tr <= 1e-10 and (best is None or ts < best[1]):
                best = ("learned", ts)
            solve["learned_fcg"]["training"]["recipe"] = {
                "solves": args.learned_solves, "pairs_per_solve": args.learned_pairs_per_solve,
                "random_pairs": args.learned_random, "stride": args.learned_stride,
                "epochs": args.learned_epochs}
            if args.gcr_krylov > 0:
                # flexible GCR(m) with the same network, FP64 vectors and operator
                vsol.upload(np.zeros(op.n))
                op.InvGcrPtc_dev(vsol, vin, warm, args.gcr_krylov)
                vsol.upload(np.zeros(op.n))
                t1 = time.perf_counter()
                it, tr, st, nnet, ncyc = op.InvGcrPtc_dev(vsol, vin, arg, args.gcr_krylov)
                ts = time.perf_counter() - t1
                err = float(np.linalg.norm(vsol.download() - x_plain)) / nx
                solve["learned_fgcr"] = {"seconds": round(ts, 4), "iters": it, "net_steps": nnet,
                                         "restarts": ncyc, "krylov": args.gcr_krylov,
                                         "true_res": tr, "status": st,
                                         "rel_err_vs_plain_fp64": err,
                                         "speedup_vs_plain_fp64": round(t_plain / ts, 3)}
                log(f"[bench] network: flexible GCR({args.gcr_krylov}) {it} iters ({nnet} with "
                    f"the net, {ncyc} restarts), true res {tr:.3e}, {ts:.2f}s "
                    f"({t_plain / ts:.2f}x)")
                if st == 0 and tr <= 1e-10 and (best is None or ts < best[1]):
                    best = ("learned_fgcr", ts)
        if best:
            name = ("learned_fcg" if best[0] == "learned" else
                    best[0] if best[0] == "learned_fgcr" else f"mixed_{best[0]}_cg")
            solve["best"] = {"method": name,
                             "speedup_vs_plain_fp64": round(t_plain / best[1], 3)}
        vsol.free()

    batched = None
    if not args.no_batched:
        log("[bench] 24-RHS batched MatInv, DWF 16^3x32 Ls 16")
        batched = run_batched(WORKLOADS["dwf_16x16x16x32_ls16_mrhs24"], args.mrhs_sloppy,
                              args.delta, args.max_iter, args.mrhs_trace)

    cpu, parity = None, None
    if not args.no_cpu:
        log("[bench] CPU baseline: one MatPcDagMatPc application, serial 4D hop and OpenMP hop")
        cb = cpu_baseline(wl, U, src)
        parity = site_errors(v_gpu, cb["out"], d_gpu, cb["dot"])
        parity["against"] = "oracle/cps_oracle.c MatPcDagMatPc of the same source (FP64)"
        log(f"[bench] GPU vs oracle: max site rel err {parity['max_site_rel_err']:.2e}, "
            f"dot rel err {parity['dot_rel_err']:.2e}")
        log("[bench] CPU baseline: 8^3x8 Ls 8 full solve (serial 4D hop) and the GPU's")
        c1 = c1_solves()
        t_ser, t_par, cores = cb["serial_hop_s"], cb["omp_s"], cb["cores"]
        cpu = {"value": round(bytes_app / t_ser / 1e9, 3), "unit": "GB/s", "cores": cores,
               "kind": "port",
               "label": "CPS-faithful (serial 4D hop, as wilson_dslash.C:280-477)",
               "sample": f"1 MatPcDagMatPc application of the same workload in {t_ser:.2f}s "
                         f"(oracle/cps_oracle.c, FP64; the 4D hop in one thread as CPS's "
                         f"noarch build, the 5D terms and vector arithmetic on {cores} OpenMP "
                         f"threads as CPS's *_omp.C)",
               "ms_per_application": round(t_ser * 1e3, 1),
               "gflops": round(flops_app / t_ser / 1e9, 3),
               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "omp_parallel_hop": {"label": "OMP-parallel restatement (4D hop over sites)",
                                    "value": round(bytes_app / t_par / 1e9, 3), "unit": "GB/s",
                                    "ms_per_application": round(t_par * 1e3, 1),
                                    "cores": cores},
               "dwf_8x8x8x8_ls8_solve": c1}
        log(f"[bench] CPU {t_ser:.2f}s (serial hop) / {t_par:.2f}s (OpenMP hop) on {cores} "
            f"threads; 8^3x8 solve CPU {c1['cpu_seconds']}s ({c1['cpu_iters']} it), GPU "
            f"{c1['gpu_seconds']}s ({c1['gpu_iters']} it)")

    out = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_app, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "lattice": list(dims), "ls": ls,
                   "action": wl["action"], "mass": mass, "m5": 1.8,
                   "mobius_b": wl.get("b"), "mobius_c": wl.get("c"), "gauge_eps": wl["eps"]},
        "gflops": round(gflops, 2),
        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        "copy_kernel_gbs": round(copy_gbs, 1),
        "bandwidth_probes_gbs": {"copy": round(bw_copy, 1), "read_only": round(bw_read, 1),
                                 "three_reads_one_write": round(bw_r3w1, 1),
                                 "peak": HBM_PEAK_GBS},
        "hbm_frac_of_best_probe": round(gbs / max(bw_copy, bw_read, bw_r3w1), 4),
        "parity_vs_oracle": parity,
        "kernels_ms": [round(float(x), 5) for x in ms_k],
        "sloppy_matpcdagmatpc": sloppy,
        "roofline": {"bound": "hbm", "achieved": round(k_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(k_gbs / HBM_PEAK_GBS, 4),
                     "traffic": kprof.get("traffic_bytes"),
                     "traffic_source": None if prof is None else prof.get("traffic_source"),
                     "l2_fill_bytes": kprof.get("l2_fill_bytes"),
                     "kernel": KERNELS[wl["action"]][hop],
                     "algorithmic_bytes": bytes_k[hop]},
        "cpu_baseline": cpu,
        "solve": solve,
        "batched_24rhs": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
