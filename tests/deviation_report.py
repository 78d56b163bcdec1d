#!/usr/bin/env python3
"""How far this build's results lie from the reference's own algorithm and from binary64,
computed with the oracle on the CPU. Test infrastructure: prints one JSON line per
comparison.

    python tests/deviation_report.py [--threads N] [--only NAME]

Comparisons:
* mask_geometry: the 512 x 512 workload (dry start, left z inflow, 3600 s). The reference's
  masked algorithm (rmg_kernel<50,false> with ENABLE_MASK, src/kernels.cu:13,18-20,129-171,
  283-299,441-452: 32-lane warps keeping 28 columns, 50 rows per block, its eval_mask /
  store_mask dilation, the literal warm-up store of Q3) against this build's masked default
  (60-column x 16-row tiles, the fip CLI's default) and against dense mode (every cell
  every step, ENABLE_MASK 0).
* precision_cfg0: the same workload in binary32 and in binary64 (this build's masked
  default geometry).
* precision_cfg1: 200 dense steps of the 5667 x 6500 workload from its 30 %-wet state in
  binary32 and in binary64.
* open_sign: quirk Q8 (src/bc_kernel_gen.cpp:132): the reference's top/bottom open boundary
  takes sign = (float)(idx - idx_neighbor) = +-pitch instead of +-1; a lake with open top
  and bottom sides, 200 steps, sign +-1 (this build) against +-256 (the row pitch in
  elements of a 200-column float field allocated with a 1024-byte pitch).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from fip_amd.workloads import BCSpec, Case, SplitMix64, config0, config1, level_state, terrain_plane_hills  # noqa: E402
from oracle_lib import DENSE, TILED, WARP_REF, Oracle  # noqa: E402

FIELDS = ("h", "qx", "qy")


def diff(a: Oracle, b: Oracle, valid=None) -> dict:
    out = {}
    for k in FIELDS + ("hmax",):
        x = a.field(k, f64=True)
        y = b.field(k, f64=True)
        d = np.abs(x - y)
        if valid is not None:
            d = d[valid]
        out[f"max_abs_d{k}"] = float(d.max())
        if k in ("h", "hmax"):
            out[f"cells_d{k}_gt_1mm"] = int((d > 1e-3).sum())
            out[f"cells_d{k}_nonzero"] = int((d > 0).sum())
    return out


def volume(o: Oracle, dx: float) -> float:
    return float(o.field("h", f64=True)[1:-1, 1:-1].sum() * dx * dx)


def lockstep(a: Oracle, b: Oracle, duration: float) -> dict:
    """Steps a and b until both reach `duration`; the first iteration whose dt differs and
    the step counts."""
    first_dt = None
    while True:
        ta, tb = a.time()[0], b.time()[0]
        if ta >= duration and tb >= duration:
            break
        if ta < duration:
            a.step()
        if tb < duration:
            b.step()
        da, db = a.time(), b.time()
        if first_dt is None and da[2] == db[2] and np.float32(da[1]) != np.float32(db[1]):
            first_dt = {"iteration": da[2], "dt_a": da[1], "dt_b": db[1]}
    return {"steps_a": a.time()[2], "steps_b": b.time()[2], "t_a": a.time()[0], "t_b": b.time()[0],
            "first_dt_difference": first_dt}


def mask_geometry(threads: int) -> list[dict]:
    case = config0()
    out = []
    ref = Oracle(case, WARP_REF, tile_w=32, tile_rows=50, warp_mask=1, quirks=1, threads=threads)
    for name, kw in (("masked 60x16 (fip CLI default)", dict(mode=TILED, tile_w=60, tile_rows=16)),
                     ("dense", dict(mode=DENSE))):
        t0 = time.perf_counter()
        r = Oracle(case, WARP_REF, tile_w=32, tile_rows=50, warp_mask=1, quirks=1, threads=threads)
        o = Oracle(case, threads=threads, **kw)
        ls = lockstep(r, o, case.duration)
        res = {"comparison": "mask_geometry", "workload": "512x512 dx=3, dry start, left z inflow, 3600 s, binary32",
               "a": "reference masked algorithm (28x50 warp tiles, eval_mask dilation, literal Q3)", "b": name,
               **ls, **diff(r, o), "volume_a_m3": volume(r, case.dx), "volume_b_m3": volume(o, case.dx),
               "seconds": round(time.perf_counter() - t0, 1)}
        out.append(res)
        print(json.dumps(res), flush=True)
    del ref
    return out


def precision_cfg0(threads: int) -> dict:
    case = config0()
    t0 = time.perf_counter()
    a = Oracle(case, TILED, tile_w=60, tile_rows=16, threads=threads)
    b = Oracle(case, TILED, tile_w=60, tile_rows=16, precision=64, threads=threads)
    ls = lockstep(a, b, case.duration)
    res = {"comparison": "precision_cfg0", "workload": "512x512 dx=3, dry start, left z inflow, 3600 s, masked 60x16",
           "a": "binary32", "b": "binary64", **ls, **diff(a, b), "volume_a_m3": volume(a, case.dx),
           "volume_b_m3": volume(b, case.dx), "seconds": round(time.perf_counter() - t0, 1)}
    print(json.dumps(res), flush=True)
    return res


def precision_cfg1(threads: int, steps: int = 200) -> dict:
    case = level_state(config1(), 0.30)
    valid = case.bed != np.float32(case.no_data)
    t0 = time.perf_counter()
    a = Oracle(case, DENSE, threads=threads)
    b = Oracle(case, DENSE, precision=64, threads=threads)
    rel = 0.0
    for _ in range(steps):
        a.step()
        b.step()
        da, db = a.time()[1], b.time()[1]
        rel = max(rel, abs(da - db) / db)
    res = {"comparison": "precision_cfg1",
           "workload": f"5667x6500 dx=3 kSt_var no_data, left q hydrograph, right open, 30% wet, {steps} dense steps",
           "a": "binary32", "b": "binary64", "steps": steps, "max_rel_dt_difference": rel,
           "t_a": a.time()[0], "t_b": b.time()[0], **diff(a, b, valid),
           "seconds": round(time.perf_counter() - t0, 1)}
    print(json.dumps(res), flush=True)
    return res


def open_sign(threads: int, steps: int = 200, pitch: int = 256) -> dict:
    W, H = 200, 150
    bed = terrain_plane_hills(W, H, 3.0, 0.0, 12, SplitMix64(77), amp=(0.0, 2.0), sigma=(4.0, 15.0))
    bed = bed.astype(np.float32)
    level = np.float32(bed.max() + 1.0)
    case = Case(name="open_tb", W=W, H=H, dx=3.0, dt=0.1, duration=1e9, bed=bed,
                h=(level - bed).astype(np.float32),
                bcs=[BCSpec("close", "left"), BCSpec("close", "right"), BCSpec("open", "top"),
                     BCSpec("open", "bottom")])
    a = Oracle(case, DENSE, threads=threads)
    b = Oracle(case, DENSE, threads=threads, open_pitch=pitch)
    v0 = volume(a, case.dx)
    for _ in range(steps):
        a.step()
        b.step()
    res = {"comparison": "open_sign", "workload": f"{W}x{H} lake at rest, open top and bottom, closed left/right, "
                                                  f"{steps} steps, binary32",
           "a": "sign +-1 (this build)", "b": f"sign +-{pitch} (reference, pitch {pitch} elements)",
           "outflow_a_m3": v0 - volume(a, case.dx), "outflow_b_m3": v0 - volume(b, case.dx),
           "t_a": a.time()[0], "t_b": b.time()[0], "steps": steps, **diff(a, b)}
    print(json.dumps(res), flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--only", choices=["mask_geometry", "precision_cfg0", "precision_cfg1", "open_sign"])
    args = ap.parse_args()
    for name, fn in (("open_sign", open_sign), ("precision_cfg0", precision_cfg0),
                     ("mask_geometry", mask_geometry), ("precision_cfg1", precision_cfg1)):
        if args.only in (None, name):
            fn(args.threads)


if __name__ == "__main__":
    main()
