"""Test helper: write a NERSC gauge file the way CPS's writer lays it out
(LatticeHeader::write, cps_pp/src/util/qcdio/comsrc/latheader.C:213-258;
data in CANONICAL site order t,z,y,x with mu and the row-major matrix inside,
iostyle.C:786-812), with the FPConv checksum (fpconv.C:320-360): the sum
mod 2^32 of the data's 32-bit words read in the file's byte order."""
from __future__ import annotations

import numpy as np

FORMATS = {  # FLOATING_POINT name -> numpy dtype of the stored reals
    "IEEE64BIG": ">f8", "IEEE64": ">f8", "IEEE64LITTLE": "<f8",
    "IEEE32BIG": ">f4", "IEEE32": ">f4", "IEEE32LITTLE": "<f4",
}


def checksum(data: bytes, big: bool) -> int:
    w = np.frombuffer(data, dtype=">u4" if big else "<u4")
    return int(w.astype(np.uint64).sum() % (1 << 32))


def write(path, links, dims, fp="IEEE64BIG", two_rows=True, plaquette=1.0, link_trace=1.0,
          checksum_override=None, truncate=0, end_header=True):
    """links: CANONICAL reals [V4*4*18].  Returns the checksum written."""
    dt = FORMATS[fp]
    u = np.asarray(links, np.float64).reshape(-1, 4, 18)
    if two_rows:
        u = u[:, :, :12]
    data = np.ascontiguousarray(u).astype(dt).tobytes()
    cs = checksum(data, dt.startswith(">")) if checksum_override is None else checksum_override
    lines = ["BEGIN_HEADER", "HDR_VERSION = 1.0",
             "DATATYPE = " + ("4D_SU3_GAUGE" if two_rows else "4D_SU3_GAUGE_3x3"),
             "STORAGE_FORMAT = 1.0"]
    lines += [f"DIMENSION_{i + 1} = {d}" for i, d in enumerate(dims)]
    lines += [f"LINK_TRACE = {link_trace:.10g}", f"PLAQUETTE  = {plaquette:.10g}"]
    lines += [f"BOUNDARY_{i + 1} = PERIODIC" for i in range(4)]
    lines += [f"CHECKSUM = {cs:8x}", "ENSEMBLE_ID = test", "ENSEMBLE_LABEL = test",
              "SEQUENCE_NUMBER = 1", "CREATOR = test", "CREATOR_HARDWARE = test",
              "CREATION_DATE = today", "ARCHIVE_DATE = today", f"FLOATING_POINT = {fp}"]
    if end_header:
        lines.append("END_HEADER")
    head = ("\n".join(lines) + "\n").encode()
    with open(path, "wb") as f:
        f.write(head)
        f.write(data[:len(data) - truncate] if truncate else data)
    return cs
