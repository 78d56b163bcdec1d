"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: bit-exact for the FIP_EXACT variant (depth, discharges, both ping-pong sides, hmax,
tile masks, dt and duration); within 1e-5 m / 1e-5 m^2/s per step for FIP_FAST.
"""
import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import BCSpec, small_case
from oracle_lib import DENSE, TILED, Oracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def sim(case, **kw):
    from fip_amd.sim import Sim
    return Sim(case, **kw)


def same_bits(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a.view(np.uint32)[~na], b.view(np.uint32)[~nb])


def assert_state_equal(g, o, what=""):
    for side in (0, 1):
        for k in ("h", "qx", "qy"):
            a, b = g.field(k, side), o.field(k, side)
            if not same_bits(a, b):
                bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
                j, i = bad[0]
                raise AssertionError(f"{what} field {k} side {side}: {len(bad)} cells differ, "
                                     f"first (x={i}, y={j}) gpu={a[j, i]!r} oracle={b[j, i]!r}")


def assert_time_equal(g, o):
    tg, dtg, itg = g.time()
    to, dto, ito = o.time()
    assert itg == ito
    assert tg == to, (tg, to)
    assert np.float32(dtg) == np.float32(dto), (dtg, dto)


CASES = {
    "wet_97x83": dict(W=97, H=83, wet=True),
    "dry_z_inflow_130x70": dict(W=130, H=70, wet=False),
    "kvar_nodata_121x90": dict(W=121, H=90, wet=True, kst_var=True, no_data=True),
    "kvar_many_110x75": dict(W=110, H=75, wet=True, kst_var="many"),
    "q_series_100x64": dict(W=100, H=64, wet=False, bc_kinds=("q", "open", "close", "close")),
    "all_open_91x77": dict(W=91, H=77, wet=True, bc_kinds=("open", "open", "open", "open")),
    # right-hand edges: W mod 60 = 59, 0 and 58 (the last tile column full, empty, or one
    # column short of full; lanes from W + 2 on load nothing)
    "open_119x70": dict(W=119, H=70, wet=True, bc_kinds=("open", "open", "open", "open")),
    "wet_120x66": dict(W=120, H=66, wet=True),
    "kvar_178x61": dict(W=178, H=61, wet=True, kst_var=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_dense_exact_fine_grained(name):
    case = small_case(**CASES[name])
    g = sim(case)
    o = Oracle(case, DENSE)
    for step in range(40):
        g.step_reference_order()
        o.step()
        if step in (0, 1, 9, 39):
            print(f"case {name} step {step}", flush=True)
            assert_state_equal(g, o, f"{name} step {step}")
            assert_time_equal(g, o)
            assert np.float32(g.cfl()) == np.float32(o.cfl_min())


@pytest.mark.parametrize("name", ["wet_97x83", "q_series_100x64", "kvar_nodata_121x90",
                                  "kvar_many_110x75"])
def test_dense_exact_device_resident(name):
    case = small_case(**CASES[name])
    case.duration = 25.0
    g = sim(case)
    g.set_run(case.duration)
    ev, done = g.simulate(10_000)
    assert ev == 1
    o = Oracle(case, DENSE)
    n = 0
    while o.time()[0] < case.duration:
        o.step()
        n += 1
    assert done == n
    assert_state_equal(g, o, name)
    assert_time_equal(g, o)
    assert same_bits(g.field("hmax"), o.field("hmax"))


@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_dense_flood_plains_across_stops(variant):
    """The flood plains of the steps that reach iteration % 100 == 0 (src/main.cpp:125-128)
    across device-resident runs that stop before, at and after multiples of 100 (and
    graph-sized chunks between them): the same hmax as one uninterrupted run, and for the
    exact variant the oracle's hmax and state at every stop."""
    case = small_case(**CASES["wet_97x83"])
    case.duration = 1e9
    a = sim(case, variant=variant)
    a.set_run(case.duration)
    b = sim(case, variant=variant)
    b.set_run(case.duration)
    o = Oracle(case, DENSE) if variant == "exact" else None
    total = 0
    for n in (37, 63, 100, 100, 1, 99, 200, 150, 50):
        print(f"flood {variant}: {n} steps from {total}", flush=True)
        ev, done = b.simulate(n)
        assert done == n
        total += n
        if o is not None:
            for _ in range(n):
                o.step()
            assert same_bits(b.field("hmax"), o.field("hmax")), total
            assert_state_equal(b, o, f"flood stop {total}")
    a.simulate(total)
    assert a.time() == b.time()
    assert same_bits(a.field("hmax"), b.field("hmax"))
    assert same_bits(a.field("h"), b.field("h"))


@pytest.mark.parametrize("rows", [16, 32, 64])
@pytest.mark.parametrize("name", ["wet_97x83", "dry_z_inflow_130x70", "kvar_nodata_121x90",
                                  "kvar_many_110x75"])
def test_masked_exact_vs_tiled_oracle(name, rows):
    case = small_case(**CASES[name])
    g = sim(case, mode="masked", tile_rows=rows)
    o = Oracle(case, TILED, tile_w=60, tile_rows=rows)
    for step in range(30):
        g.step_reference_order()
        o.step()
        if step in (0, 3, 29):
            print(f"case {name} rows {rows} step {step}", flush=True)
            assert_state_equal(g, o, f"{name}/R{rows} step {step}")
            assert_time_equal(g, o)
            assert np.array_equal(g.masks(), o.masks())


@pytest.mark.parametrize("rows", [16, 64])
def test_masked_device_resident(rows):
    case = small_case(**CASES["dry_z_inflow_130x70"])
    case.duration = 30.0
    g = sim(case, mode="masked", tile_rows=rows)
    g.set_run(case.duration)
    ev, done = g.simulate(10_000)
    assert ev == 1
    o = Oracle(case, TILED, tile_w=60, tile_rows=rows)
    while o.time()[0] < case.duration:
        o.step()
    assert_state_equal(g, o, f"masked R{rows}")
    assert_time_equal(g, o)
    assert np.array_equal(g.masks(), o.masks())
    assert same_bits(g.field("hmax"), o.field("hmax"))


def test_cn_solver_boundaries():
    case = small_case(W=80, H=72, wet=True, bc_kinds=("z", "open", "close", "open"))
    g = sim(case, staggered=False)
    o = Oracle(case, DENSE, staggered=0)
    for _ in range(15):
        g.step_reference_order()
        o.step()
    assert_state_equal(g, o, "cn")


def test_bc_params_and_host_launch():
    case = small_case(**CASES["q_series_100x64"])
    g = sim(case)
    o = Oracle(case, DENSE)
    for _ in range(12):
        g.step_reference_order()
        o.step()
    gp = g.bc_params()
    for i in range(len(case.bcs)):
        q, z = o.bc_params(i)
        assert np.float32(gp[i, 0]) == np.float32(q) and np.float32(gp[i, 1]) == np.float32(z)
    # the reference seam: host-evaluated parameters through BoundaryConditionKernel::launch
    g.control()
    o.control()
    g.compute_step()
    o.stencil()
    params = np.array([o.bc_params(i) for i in range(len(case.bcs))], dtype=np.float32).ravel()
    g.bc_launch(params)
    o.apply_bcs()
    assert_state_equal(g, o, "host-launched BC")
    assert_time_equal(g, o)


@pytest.mark.parametrize("kst_var", [False, True, "many"])
def test_fast_variant_within_tolerance(kst_var):
    """One step of the fast variant from the exact run's state (uniform kSt, kSt read by
    class, kSt read as floats): fields within 1e-5, CFL minimum within 1e-5 relative."""
    case = small_case(W=150, H=120, wet=True, kst_var=kst_var)
    e = sim(case)
    for _ in range(20):
        e.step_reference_order()
    st = e.state()
    f = sim(case, variant="fast")
    for k, v in st.items():
        f.set_field(k, 0, v)
        f.set_field(k, 1, e.field(k, 1))
    f.set_dt(e.time()[1])
    e.compute_step()
    f.compute_step()
    for k in ("h", "qx", "qy"):
        d = np.abs(e.field(k).astype(np.float64) - f.field(k).astype(np.float64))
        assert d.max() <= 1e-5, (k, d.max())
    assert abs(f.cfl() - e.cfl()) <= 1e-5 * e.cfl(), (f.cfl(), e.cfl())
    classes = f.stats()["kst_classes"]
    assert classes == (5 if kst_var is True else 0), classes


def test_sampling_and_dumps():
    case = small_case(W=97, H=83, wet=True)
    case.duration = 40.0
    pts = [(10, 20, 0), (50, 41, 1), (60, 40, 2), (30, 30, 3)]
    g = sim(case)
    g.set_run(case.duration, sampling_interval=5, points=pts, save_times=[12.0, 33.3])
    o = Oracle(case, DENSE)
    # oracle emulation of simulate() with sample() and dumpDeviceBuffers()
    timer, saves, rows, dumps = 0.0, [12.0, 33.3], [], []
    bed = case.bed
    while o.time()[0] < case.duration:
        t, dt, it = o.time()
        if case.variable_dt:
            o.control()
        t = o.time()[0]
        if t >= timer:
            timer += 5
            cur = {k: o.field(k) for k in ("h", "qx", "qy")}
            vals = []
            for x, y, ty in pts:
                if ty == 0:
                    vals.append(np.float32(cur["h"][y, x] + bed[y, x]))
                else:
                    vals.append(cur[("h", "h", "qx", "qy")[ty]][y, x])
            rows.append((t, vals))
        if it == 0:
            o.apply_bcs()
        o.stencil()
        o.apply_bcs()
        if o.time()[2] % 100 == 0:
            o.floodplain()
        if saves and o.time()[0] >= saves[0]:
            saves.pop(0)
            dumps.append((o.time()[2], {k: o.field(k, 1) for k in ("h", "qx", "qy")}))
    got_rows, got_dumps = [], []
    while True:
        ev, _ = g.simulate(1_000_000)
        t, v = g.read_samples()
        got_rows += list(zip(t, v))
        if ev == 2:
            got_dumps.append((g.time()[2], {k: g.field(k, 1) for k in ("h", "qx", "qy")}))
            continue
        assert ev == 1
        break
    assert len(got_rows) == len(rows)
    for (tg, vg), (to, vo) in zip(got_rows, rows):
        assert tg == to
        assert same_bits(vg, np.array(vo, np.float32))
    assert [d[0] for d in got_dumps] == [d[0] for d in dumps]
    for (_, a), (_, b) in zip(got_dumps, dumps):
        for k in a:
            assert same_bits(a[k], b[k])


def test_face_depth_reciprocal_is_correctly_rounded():
    # the exact variant's reciprocal of a face depth (hardware reciprocal and two Newton
    # steps) equals the IEEE division on every binary32 b in [DH, 2^64]: all face depths
    # of a state without infinities (1/max(h, DH), src/kernels.cu:89,116)
    from fip_amd.sim import check_rcp
    nbad, first = check_rcp(np.float32(0.01), np.float32(2.0 ** 64))
    assert nbad == 0, (nbad, first)
