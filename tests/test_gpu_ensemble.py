"""Ensembles on the GPU: every scenario of an ensemble run equals a single-domain run of
the same scenario through the same backend (bit for bit with FIP_EXACT), and the oracle.

Scenarios differ in size, terrain, roughness, hydrograph, BC list and target duration,
so they take different dt sequences and finish at different iterations.
"""
import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import BCSpec, config2, small_case
from oracle_lib import TILED, Oracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

SMALL_SIZES = ((96, 64), (130, 70), (64, 128), (181, 97))


def same_bits(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def scenarios():
    cases = [config2(s, sizes=SMALL_SIZES, duration=40.0 + 7.0 * s) for s in range(4)]
    # a z-inflow case with single-cell BCs and a lake, and an all-open lake
    z = small_case(97, 83, scenario=3, wet=True)
    z.duration = 33.0
    o = small_case(91, 77, scenario=4, wet=True, bc_kinds=("open", "open", "open", "open"))
    o.duration = 25.0
    return cases + [z, o]


def run_single(case, mode, variant, rows):
    from fip_amd.sim import Sim
    s = Sim(case, mode=mode, variant=variant, tile_rows=rows)
    s.set_run(case.duration)
    ev, _ = s.simulate(1_000_000)
    assert ev == 1
    return s


@pytest.mark.parametrize("mode,rows", [("masked", 32), ("dense", 16)])
def test_ensemble_equals_single_domain_exact(mode, rows):
    from fip_amd.sim import Ensemble
    cases = scenarios()
    print(f"ensemble {mode} R={rows}: " + ", ".join(f"{c.W}x{c.H}" for c in cases), flush=True)
    e = Ensemble(cases, mode=mode, variant="exact", tile_rows=rows)
    e.set_run([c.duration for c in cases])
    active, steps = e.simulate(1_000_000)
    assert active == 0
    its = set()
    for k, c in enumerate(cases):
        s = run_single(c, mode, "exact", rows)
        t1, dt1, it1 = s.time()
        t2, dt2, it2 = e.time(k)
        assert (it1, t1, np.float32(dt1)) == (it2, t2, np.float32(dt2)), (k, c.name)
        its.add(it2)
        for side in (0, 1):
            for f in ("h", "qx", "qy"):
                assert same_bits(s.field(f, side), e.field(k, f, side)), (k, c.name, f, side)
        assert same_bits(s.field("hmax"), e.field(k, "hmax")), (k, c.name)
        s.close()
    assert len(its) > 1  # scenarios stopped at different iterations
    assert steps >= max(its)
    e.close()


def test_ensemble_scenario_equals_oracle():
    from fip_amd.sim import Ensemble
    cases = scenarios()[:3]
    e = Ensemble(cases, mode="masked", variant="exact", tile_rows=32)
    e.set_run([c.duration for c in cases])
    active, _ = e.simulate(1_000_000)
    assert active == 0
    c = cases[1]
    o = Oracle(c, TILED, tile_w=60, tile_rows=32)
    while o.time()[0] < c.duration:
        o.step()
    assert e.time(1)[2] == o.time()[2]
    for f in ("h", "qx", "qy"):
        assert same_bits(e.field(1, f), o.field(f)), f
    assert same_bits(e.field(1, "hmax"), o.field("hmax"))
    e.close()


def test_ensemble_fast_within_tolerance():
    from fip_amd.sim import Ensemble
    cases = scenarios()[:4]
    e = Ensemble(cases, mode="masked", variant="fast", tile_rows=32)
    e.set_run([c.duration for c in cases])
    e.simulate(1_000_000)
    for k, c in enumerate(cases):
        s = run_single(c, "masked", "fast", 32)
        assert s.time()[2] == e.time(k)[2]
        for f in ("h", "qx", "qy"):
            d = np.abs(s.field(f).astype(np.float64) - e.field(k, f)).max()
            assert d <= 1e-5, (k, f, d)
        s.close()
    e.close()


def test_ensemble_launch_shapes_agree():
    """Grid and persistent list kernels, any waves per work-group: same bits."""
    from fip_amd.sim import Ensemble
    cases = scenarios()[:4]
    ref = None
    for wpb, bpc in ((4, 8), (1, 0), (8, 0), (2, 16)):
        e = Ensemble(cases, mode="masked", variant="exact", tile_rows=32)
        e.set_launch(wpb, bpc)
        e.set_run([c.duration for c in cases])
        e.simulate(1_000_000)
        got = [(e.time(k), e.field(k, "h"), e.field(k, "qx")) for k in range(len(cases))]
        e.close()
        if ref is None:
            ref = got
            continue
        for (t1, h1, q1), (t2, h2, q2) in zip(ref, got):
            assert t1 == t2 and same_bits(h1, h2) and same_bits(q1, q2), (wpb, bpc)


@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_ensemble_dense_shapes_agree(variant):
    """Dense march: any waves per work-group (so any span length) gives the same bits,
    and each scenario matches its single-domain dense run (bits for the exact variant,
    1e-5 for the fast one, whose single-domain kernel is compiled separately)."""
    from fip_amd.sim import Ensemble
    cases = scenarios()
    ref = None
    for wpb in (8, 1, 2, 16):
        print(f"ensemble dense {variant} wpb={wpb}", flush=True)
        e = Ensemble(cases, mode="dense", variant=variant, tile_rows=16)
        e.set_launch(wpb, 0)
        e.set_run([c.duration for c in cases])
        active, _ = e.simulate(1_000_000)
        assert active == 0
        st = e.stats()
        assert st["dense_span"] >= 16 and (st["dense_span"] + 2) % 6 == 0
        got = [(e.time(k), [e.field(k, f) for f in ("h", "qx", "qy", "hmax")]) for k in range(len(cases))]
        e.close()
        if ref is None:
            ref = got
            continue
        for k, ((t1, f1), (t2, f2)) in enumerate(zip(ref, got)):
            assert t1 == t2, (wpb, k)
            assert all(same_bits(a, b) for a, b in zip(f1, f2)), (wpb, k)
    for k, c in enumerate(cases):
        s = run_single(c, "dense", variant, 16)
        assert s.time()[2] == ref[k][0][2], k
        for f, got in zip(("h", "qx", "qy"), ref[k][1]):
            if variant == "exact":
                assert same_bits(s.field(f), got), (k, f)
            else:
                d = np.abs(s.field(f).astype(np.float64) - got).max()
                assert d <= 1e-5, (k, f, d)
        s.close()


@pytest.mark.parametrize("mode,rows", [("masked", 32), ("dense", 16)])
def test_ensemble_kst_var_equals_single_domain(mode, rows):
    """Scenarios with per-cell kSt (the ensemble reads it as floats; a single domain with
    at most 64 distinct values reads a one-byte class): same bits, exact variant."""
    from fip_amd.sim import Ensemble
    cases = []
    for k, (W, H) in enumerate(((97, 83), (130, 70), (64, 128))):
        c = small_case(W, H, scenario=k, wet=True, kst_var=True, no_data=(k == 1))
        c.duration = 20.0 + 6.0 * k
        cases.append(c)
    print(f"ensemble kst_var {mode}", flush=True)
    e = Ensemble(cases, mode=mode, variant="exact", tile_rows=rows)
    e.set_run([c.duration for c in cases])
    active, _ = e.simulate(1_000_000)
    assert active == 0
    for k, c in enumerate(cases):
        s = run_single(c, mode, "exact", rows)
        assert s.time() == e.time(k), k
        for f in ("h", "qx", "qy", "hmax"):
            assert same_bits(s.field(f), e.field(k, f)), (k, f)
        s.close()
    e.close()


def test_ensemble_chunked_runs_match_one_run():
    """Stopping and resuming (non-graph chunks, then graph chunks) changes nothing."""
    from fip_amd.sim import Ensemble
    cases = scenarios()[:3]
    a = Ensemble(cases, mode="masked", variant="exact", tile_rows=32)
    a.set_run([c.duration for c in cases])
    a.simulate(1_000_000)
    b = Ensemble(cases, mode="masked", variant="exact", tile_rows=32)
    b.set_run([c.duration for c in cases])
    for n in (7, 93, 100, 1, 250):
        b.simulate(n)
    b.simulate(1_000_000)
    for k in range(len(cases)):
        assert a.time(k) == b.time(k)
        for f in ("h", "qx", "qy", "hmax"):
            assert same_bits(a.field(k, f), b.field(k, f)), (k, f)
    a.close()
    b.close()
