"""End to end: `fip parameters.yml` on the GPU against the oracle's simulate() loop with
sample() and dumpDeviceBuffers() (src/main.cpp:45-202)."""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import small_case, write_case
from oracle_lib import TILED, Oracle

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fip_amd", "bin", "fip")


def fmt(x):
    """Default iostream formatting (precision 6), as the reference prints samples."""
    s = "%.6g" % x
    return s


def oracle_run(case, points, interval, saves, rows):
    o = Oracle(case, TILED, tile_w=60, tile_rows=rows)
    timer, pending, lines, dumps = 0.0, sorted(saves), [], []
    while o.time()[0] < case.duration:
        if case.variable_dt:
            o.control()
        t, _, it = o.time()
        if t >= timer and interval > 0:
            timer += interval
            vals = []
            for x, y, ty in points:
                if ty == 0:
                    vals.append(np.float32(o.field("h")[y, x] + case.bed[y, x]))
                else:
                    vals.append(o.field(("h", "h", "qx", "qy")[ty])[y, x])
            lines.append(fmt(t) + ";" + "".join(fmt(float(v)) + ";" for v in vals))
        if it == 0:
            o.apply_bcs()
        o.stencil()
        o.apply_bcs()
        if o.time()[2] % 100 == 0:
            o.floodplain()
        if pending and o.time()[0] >= pending[0]:
            pending.pop(0)
            dumps.append((int(o.time()[0]), {k: o.field(k, 1) for k in ("h", "qx", "qy")},
                          o.field("hmax")))
    return lines, dumps, o.time()[2]


def test_cli_matches_oracle(tmp_path):
    case = small_case(97, 83, wet=True, bc_kinds=("q", "open", "close", "close"))
    case.duration = 45.0
    pts = [{"x": 10, "y": 40, "label": "zl", "type": "z"}, {"x": 50, "y": 41, "label": "h", "type": "h"},
           {"x": 60, "y": 40, "label": "qx", "type": "qx"}, {"x": 30, "y": 30, "label": "qy", "type": "qy"}]
    p = write_case(case, str(tmp_path / "run"), sampling=pts, sampling_interval=4,
                   save_state=["12", "3*10.5"])
    print("cli case", p, flush=True)
    env = dict(os.environ, FIP_MODE="masked", FIP_VARIANT="exact", FIP_TILE_ROWS="32")
    r = subprocess.run([CLI, p], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    out = r.stdout.strip().splitlines()
    assert out[0] == "t;zl;h;qx;qy;"
    tp = [(q["x"], q["y"], {"z": 0, "h": 1, "qx": 2, "qy": 3}[q["type"]]) for q in pts]
    lines, dumps, iters = oracle_run(case, tp, 4, [12.0, 31.5], 32)
    assert out[1:] == lines
    assert f"Finished {iters} iterations after" in r.stderr
    name = case.name
    for t, st, hmax in dumps:
        for tag, key in (("ah", "h"), ("qx", "qx"), ("qy", "qy"), ("flood_plains", None)):
            files = glob.glob(str(tmp_path / "run" / f"{name}_{tag}_{t}_*"))
            assert len(files) == 1, (tag, t, files)
            raw = open(files[0], "rb").read()
            W, H = np.frombuffer(raw[:16], np.uint64)
            assert (W, H) == (case.W, case.H)
            a = np.frombuffer(raw[16:], np.float32).reshape(case.H, case.W)
            b = hmax if key is None else st[key]
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), tag
