"""The configurations at their real sizes on the GPU, against the oracle.

* 5667 x 6500 (kSt_var Voronoi, no_data polygon, left q hydrograph, right open) from its
  30 %-wet state: exact dense steps bit-exact with the DENSE oracle, exact masked steps
  bit-exact with the TILED oracle at the GPU's tile geometry (masks included), and the fast
  variant within its tolerance of the exact one, for one step and over a run of steps.
* The 64-scenario ensemble at its real sizes (256^2 .. 2048 x 1536): every one of four
  scenarios of all five sizes equals its single-domain run bit for bit, and the 1024^2 one
  the TILED oracle.
* The closure at its real shape: restriction of a 2048^2 fine run onto 512^2 bit-exact with
  the oracle, and the network on the 512^2 coarse features within its tolerance of the
  oracle's.
"""
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import config1, config2, level_state
from oracle_lib import DENSE, TILED, Oracle
from test_gpu_parity import assert_state_equal, assert_time_equal, same_bits

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import closure_oracle as co  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

THREADS = min(16, os.cpu_count() or 1)

# The fast variant against the exact one (hardware rcp/log2/exp2, FMA contraction): one
# step from the same state within 1e-5 m / 1e-5 m^2/s and 1e-5 relative on the CFL
# minimum; over 20 steps from the same state the fields drift apart by at most FAST_RUN_TOL
# and every dt of the sequence agrees to FAST_DT_REL.
FAST_STEP_TOL = 1e-5
FAST_RUN_TOL = 1e-4
FAST_DT_REL = 1e-6


@pytest.fixture(scope="module")
def cfg1():
    return level_state(config1(), 0.30)


def log(msg):
    print(msg, flush=True)


def test_config1_dense_exact_bit_exact(cfg1):
    """105 steps: past the flood plains of iteration 100, which dense mode takes inside the
    stencil step that follows."""
    from fip_amd.sim import Sim
    case = cfg1
    steps = 105
    log(f"config1 5667x6500 dense exact: {steps} steps on the GPU")
    g = Sim(case, mode="dense", variant="exact")
    g.set_run(case.duration)
    ev, done = g.simulate(steps)
    assert done == steps
    log(f"config1 dense: {steps} oracle steps on {THREADS} threads")
    o = Oracle(case, DENSE, threads=THREADS)
    o.step(steps)
    o.control()  # the device loop has already taken the next step's dt (end of its post kernel)
    assert_state_equal(g, o, "config1 dense exact")
    assert_time_equal(g, o)
    assert np.float32(g.cfl()) == np.float32(o.cfl_min())
    assert same_bits(g.field("hmax"), o.field("hmax"))
    assert not np.array_equal(o.field("hmax"), case.h)  # the flood plains changed it
    g.close()


def test_config1_masked_exact_bit_exact(cfg1):
    from fip_amd.sim import Sim
    case = cfg1
    log("config1 5667x6500 masked exact, 16-row tiles: 20 steps on the GPU")
    g = Sim(case, mode="masked", variant="exact", tile_rows=16)
    g.set_run(case.duration)
    ev, done = g.simulate(20)
    assert done == 20
    st = g.stats()
    assert 0 < st["tiles_computed"] < st["tiles_total"]  # some tiles skipped, some run
    log(f"config1 masked: 20 TILED oracle steps on {THREADS} threads")
    o = Oracle(case, TILED, tile_w=60, tile_rows=16, threads=THREADS)
    o.step(20)
    o.control()
    assert_state_equal(g, o, "config1 masked exact")
    assert_time_equal(g, o)
    assert np.array_equal(g.masks(), o.masks())
    g.close()


def test_config1_fast_within_tolerance(cfg1):
    from fip_amd.sim import Sim
    case = cfg1
    log("config1 5667x6500: fast vs exact, one step from the same state")
    e = Sim(case, mode="dense", variant="exact")
    e.set_run(case.duration)
    e.simulate(20)
    f = Sim(case, mode="dense", variant="fast")
    for k in ("h", "qx", "qy"):
        for side in (0, 1):
            f.set_field(k, side, e.field(k, side))
    f.set_dt(e.time()[1])
    e.compute_step()
    f.compute_step()
    for k in ("h", "qx", "qy"):
        d = np.abs(e.field(k).astype(np.float64) - f.field(k).astype(np.float64))
        log(f"  one step: max |d{k}| = {d.max():.3e}")
        assert d.max() <= FAST_STEP_TOL, (k, d.max())
    assert abs(f.cfl() - e.cfl()) <= FAST_STEP_TOL * e.cfl(), (f.cfl(), e.cfl())
    e.close()
    f.close()

    log("config1 5667x6500: fast vs exact, 20 steps each from the 30 %-wet state")
    runs = {}
    for variant in ("exact", "fast"):
        s = Sim(case, mode="dense", variant=variant)
        s.set_run(case.duration)
        dts = []
        for _ in range(20):
            s.simulate(1)
            dts.append(s.time()[1])
        runs[variant] = (s.state(), np.array(dts, np.float64))
        s.close()
    (se, de), (sf, df) = runs["exact"], runs["fast"]
    for k in ("h", "qx", "qy"):
        d = np.abs(se[k].astype(np.float64) - sf[k].astype(np.float64)).max()
        log(f"  20 steps: max |d{k}| = {d:.3e}")
        assert d <= FAST_RUN_TOL, (k, d)
    rel = np.abs(de - df) / de
    log(f"  20 steps: max dt relative difference {rel.max():.3e}")
    assert rel.max() <= FAST_DT_REL


def config2_all():
    return [level_state(config2(s), 0.30) for s in range(64)]


def test_config2_ensemble_at_real_sizes():
    from fip_amd.sim import Ensemble, Sim
    cases = config2_all()
    sizes = {(c.W, c.H) for c in cases}
    assert len(sizes) == 5
    # one scenario of each of four sizes, the largest among them
    pick = []
    for want in ((2048, 1536), (1024, 1024), (512, 384), (256, 256)):
        pick.append(next(s for s, c in enumerate(cases) if (c.W, c.H) == want))
    steps, rows = 200, 16
    log(f"config2: 64 scenarios ({sum(c.W * c.H for c in cases)} cells), masked exact, "
        f"{steps} ensemble steps")
    e = Ensemble(cases, mode="masked", variant="exact", tile_rows=rows)
    e.set_run([c.duration for c in cases])
    act, done = e.simulate(steps)
    assert done == steps and act == 64
    got = {s: {(k, side): e.field(s, k, side) for k in ("h", "qx", "qy") for side in (0, 1)} for s in pick}
    times = {s: e.time(s) for s in pick}
    e.close()
    for s in pick:
        c = cases[s]
        log(f"config2 scenario {s} ({c.W}x{c.H}): single-domain run")
        g = Sim(c, mode="masked", variant="exact", tile_rows=rows)
        g.set_run(c.duration)
        g.simulate(steps)
        for (k, side), a in got[s].items():
            assert same_bits(a, g.field(k, side)), (s, k, side)
        assert times[s] == g.time(), (s, times[s], g.time())
        g.close()
    s = pick[1]
    log(f"config2 scenario {s} (1024x1024): {steps} TILED oracle steps on {THREADS} threads")
    o = Oracle(cases[s], TILED, tile_w=60, tile_rows=rows, threads=THREADS)
    o.step(steps)
    o.control()
    for (k, side), a in got[s].items():
        assert same_bits(a, o.field(k, side)), (s, k, side)
    t, dt, it = times[s]
    to, dto, ito = o.time()
    assert it == ito and t == to and np.float32(dt) == np.float32(dto)


def test_closure_at_real_shape():
    from fip_amd import closure_run
    from fip_amd.sim import Closure, Sim, restrict
    fine = level_state(closure_run.fine_case(5, W=2048, H=2048), 0.30)
    fine.h[:, :1024] += np.float32(0.5)  # a raised left half: the water moves
    coarse, bstd, zref = closure_run.coarse_of(fine)
    assert (coarse.W, coarse.H) == (512, 512)
    log("closure: 2048^2 fine run (400 steps, 20 s), restriction onto 512^2")
    f = Sim(fine, mode="dense", variant="fast")
    f.set_run(1e12)
    f.simulate(400)
    c = Sim(coarse, mode="dense")
    restrict(f, c, 4)
    hc, qxc, qyc = co.restrict(f.field("h"), f.field("qx"), f.field("qy"), 4)
    assert (hc > 0.01).sum() > 10000 and np.abs(qxc).max() > 0.01  # a wet, moving state
    fh = f.field("h")
    for side in (0, 1):
        got = c.field("h", side)
        if not same_bits(got, hc):
            bad = np.argwhere(got.view(np.uint32) != hc.view(np.uint32))
            J, I = bad[0]
            blk = fh[4 * J:4 * J + 4, 4 * I:4 * I + 4]
            raise AssertionError(f"restricted h side {side}: {len(bad)} cells differ; first (I={I}, J={J}) "
                                 f"gpu {got[J, I]!r} oracle {hc[J, I]!r} fine block {blk.ravel().tolist()!r}")
        assert same_bits(c.field("qx", side), qxc)
        assert same_bits(c.field("qy", side), qyc)
    assert c.time()[0] == f.time()[0]
    log("closure: network on the 512^2 coarse features against the oracle")
    rng = np.random.default_rng(17)
    cl = Closure(512, 512, 0, seed=3)
    p = np.where(co.live_mask(), rng.normal(0.0, 0.2, co.NPARAMS), 0.0).astype(np.float32)
    cl.params = p
    feat = co.features(hc, qxc, qyc, coarse.bed, bstd, coarse.kst, zref)[None]
    got = cl.forward(feat)
    ref, _ = co.forward(feat, p, emulate=True)
    err = np.abs(got - ref).max()
    assert err <= 2e-3 * max(1.0, np.abs(ref).max()), err
    # the fused inference kernel on the coarse simulation itself, against the same network
    of = cl.fused_outputs(c, bstd, zref)
    o2 = cl.forward(co.features(c.field("h"), c.field("qx"), c.field("qy"), coarse.bed, bstd,
                                coarse.kst, zref)[None])[0]
    assert np.abs(of - o2).max() <= 2e-2 * max(1.0, np.abs(o2).max())
    for x in (f, c, cl):
        x.close()
