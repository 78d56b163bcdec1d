"""Launch shapes and the autotuner: results do not depend on the launch shape (waves per
work-group, persistent work-groups per CU), and in dense mode not on the tile rows; the
tuners time every candidate and return one of them."""
import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import config2, small_case

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def run(case, mode, rows, wpb, bpc, steps=60):
    from fip_amd.sim import Sim
    s = Sim(case, mode=mode, variant="exact", tile_rows=rows)
    s.set_launch(wpb, bpc)
    s.set_run(1e9)
    s.simulate(steps)
    out = {f: s.field(f) for f in ("h", "qx", "qy")}
    out["t"] = s.time()
    s.close()
    return out


def same(a, b):
    return all(np.array_equal(bits(a[f]), bits(b[f])) for f in ("h", "qx", "qy")) and a["t"] == b["t"]


def test_dense_independent_of_shape():
    case = small_case(250, 140, wet=True)
    ref = run(case, "dense", 32, 4, 0)
    for rows, wpb in ((16, 1), (64, 2), (32, 8), (16, 16)):
        print(f"dense rows={rows} wpb={wpb}", flush=True)
        assert same(ref, run(case, "dense", rows, wpb, 0)), (rows, wpb)


def test_masked_independent_of_launch():
    case = small_case(250, 140, wet=False)
    ref = run(case, "masked", 32, 4, 8)
    for wpb, bpc in ((1, 4), (8, 16), (16, 1), (4, 0), (1, 0)):
        print(f"masked wpb={wpb} bpc={bpc}", flush=True)
        assert same(ref, run(case, "masked", 32, wpb, bpc)), (wpb, bpc)


def test_tune_single_domain():
    from fip_amd.sim import tune
    case = small_case(300, 200, wet=True)
    r = tune(case, mode="dense", variant="fast", steps=5)
    assert r["tried"] == 5 and r["blocks_per_cu"] == -1
    assert r["tile_rows"] in (16, 32, 64) and r["waves_per_block"] in (1, 2, 4, 8, 16)
    assert 0 < r["stencil_ms"] < 100
    r = tune(case, mode="masked", variant="fast", steps=5)
    assert r["tried"] == 60 and r["blocks_per_cu"] in (0, 4, 8, 16)


def test_tune_ensemble():
    from fip_amd.sim import tune_ensemble
    cases = [config2(s, sizes=((96, 64), (130, 70))) for s in range(3)]
    r = tune_ensemble(cases, mode="masked", variant="fast", steps=4)
    assert r["tried"] == 48
    assert r["tile_rows"] in (16, 32, 64)
