"""The plain single-domain workload at its full size and duration, end to end: 512 x 512
cells at dx = 3 m, seeded plane + 64 hills + a 20-cell 3 m channel, dry start, 41 left
type-z inflow rows (bed + 2 m), right open, top and bottom closed, uniform kSt 30,
variable dt from 0.1 s, 3600 s simulated.

The GPU runs it (masked, 3600 s; dense, the first 900 s) through the device-resident loop (hipGraph chunks, dt on the device) in
the exact variant; the binary32 oracle runs it at the same tile geometry. The step count,
the final time and dt, both ping-pong sides of h, qx and qy, hmax and the tile masks
must agree bit for bit: any difference in one step's dt would change every later step."""
import os

import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import config0
from oracle_lib import DENSE, TILED, Oracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

THREADS = min(16, os.cpu_count() or 1)


def run_both(mode, rows, duration=3600.0):
    from fip_amd.sim import Sim
    case = config0(duration=duration)
    g = Sim(case, mode=mode, variant="exact", tile_rows=rows)
    g.set_run(case.duration)
    ev, done = g.simulate(10 ** 7)
    assert ev == 1, (ev, done)
    if mode == "masked":
        o = Oracle(case, TILED, tile_w=60, tile_rows=rows, threads=THREADS)
    else:
        o = Oracle(case, DENSE, threads=THREADS)
    n = 0
    while o.time()[0] < case.duration:
        o.step()
        n += 1
    print(f"config0 {mode} R{rows}: {done} GPU steps, {n} oracle steps, t = {o.time()[0]:.4f} s", flush=True)
    return case, g, o, done, n


def check(g, o, done, n, masks):
    assert done == n
    tg, dtg, itg = g.time()
    to, dto, ito = o.time()
    assert (tg, np.float32(dtg), itg) == (to, np.float32(dto), ito)
    for side in (0, 1):
        for k in ("h", "qx", "qy"):
            a, b = g.field(k, side), o.field(k, side)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (k, side)
    assert np.array_equal(g.field("hmax").view(np.uint32), o.field("hmax").view(np.uint32))
    if masks:
        assert np.array_equal(g.masks(), o.masks())
    assert g.field("h").max() > 1.0  # the inflow filled the channel


def test_config0_full_run_masked_exact():
    case, g, o, done, n = run_both("masked", 64)
    check(g, o, done, n, masks=True)
    g.close()


def test_config0_dense_exact_first_900s():
    """Dense mode (every tile computed, as the reference's ENABLE_MASK 0 build) over the
    first 900 s, against the oracle's dense restatement."""
    case, g, o, done, n = run_both("dense", 64, duration=900.0)
    check(g, o, done, n, masks=False)
    g.close()
