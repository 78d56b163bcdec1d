"""Pins the oracle's stencil, boundary bodies, dt control and interp with checks that do not
restate its formulas: volume budgets through the boundary faces, symmetries of the scheme,
closed forms of the dt sequence, and the reference's own TimeSeries::interp. CPU only.

Faces: qx(i, j) is the east face of cell i, qy(i, j) the north face of row j. The boundary
faces the stencil holds are qx at columns 0 and W-2 and qy at rows 0 and H-2
(src/kernels.cu:373-376); the side templates of src/bc_kernel_gen.cpp:11-73 write the
staggered face at column 0 (left), W-2 (right), row H-2 (top) and row 0 (bottom).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from fip_amd.workloads import BCSpec, Case, SplitMix64, terrain_plane_hills
from oracle_lib import DENSE, Oracle, interp

F32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = 9.81


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def hilly(W, H, seed, sigma=(2.0, 6.0), n=8):
    return terrain_plane_hills(W, H, 3.0, 0.001, n, SplitMix64(seed), amp=(0.0, 1.5),
                               sigma=sigma).astype(np.float32)


def interior_volume(o, dx):
    """dx^2 times the sum of h over the stored cells [1, W-2] x [1, H-2], binary64 oracle."""
    return o.field("h", f64=True)[1:-1, 1:-1].sum() * dx * dx


def held_face_outflow(o):
    """Net discharge out of the stored box through the held faces of the current state:
    sum_j qx(W-2, j) - qx(0, j) over rows 1..H-2, plus sum_i qy(i, H-2) - qy(i, 0) over
    columns 1..W-2 (m^3/s per m of face width, before the factor dx)."""
    qx = o.field("qx", f64=True)
    qy = o.field("qy", f64=True)
    return (qx[1:-1, -2] - qx[1:-1, 0]).sum() + (qy[-2, 1:-1] - qy[0, 1:-1]).sum()


def budget_case(W, H, seed, bcs, q0=0.3):
    """Water above the bed everywhere and nonzero initial discharges on every face, so a
    boundary body that wrote the wrong face would leave a nonzero held face."""
    bed = hilly(W, H, seed)
    level = F32(bed.max() + 0.8)
    h = (level - bed).astype(np.float32)
    rng = np.random.default_rng(seed)
    qx = rng.uniform(-q0, q0, (H, W)).astype(np.float32)
    qy = rng.uniform(-q0, q0, (H, W)).astype(np.float32)
    return Case(name="budget", W=W, H=H, dx=3.0, dt=0.05, duration=1e9, bed=bed, h=h, qx=qx,
                qy=qy, variable_dt=False, bcs=bcs)


def stencil_budget_residual(o, dx, dt=None):
    """One stencil step (no BCs): |dV - (-dt dx outflow)| relative to V, with the step's
    own dt (the binary32 value of the case's dt, widened)."""
    dt = o.time()[1]
    v0 = interior_volume(o, dx)
    out = held_face_outflow(o)
    o.stencil()
    v1 = interior_volume(o, dx)
    return abs((v1 - v0) + dt * dx * out) / v0, v1 - v0


# ---------------------------------------------------------------------------------------
# continuity: the interior volume changes by the flux through the held faces


def test_stencil_volume_budget_through_held_faces():
    W, H, dx, dt = 47, 39, 3.0, 0.05
    case = budget_case(W, H, 11, bcs=[])
    o = Oracle(case, DENSE, precision=64, with_bcs=False)
    worst = 0.0
    moved = 0.0
    for _ in range(20):
        r, dv = stencil_budget_residual(o, dx, dt)
        worst = max(worst, r)
        moved = max(moved, abs(dv) / interior_volume(o, dx))
    assert worst < 1e-13, worst
    assert moved > 1e-6  # the held faces carry flux: the check is not vacuous


# ---------------------------------------------------------------------------------------
# close: the staggered face is 0, the lateral face and the depth are copied from the
# neighbour; a box closed on all four sides keeps its volume from a moving start


SIDES = ("left", "right", "top", "bottom")


def staggered_face(side, W, H):
    """(field, index) of the face each side's boundary body writes."""
    return {"left": ("qx", (slice(None), 0)), "right": ("qx", (slice(None), W - 2)),
            "top": ("qy", (H - 2, slice(None))), "bottom": ("qy", (0, slice(None)))}[side]


def cell_and_neighbour(side, W, H):
    return {"left": ((slice(None), 0), (slice(None), 1)),
            "right": ((slice(None), W - 1), (slice(None), W - 2)),
            "top": ((H - 1, slice(None)), (H - 2, slice(None))),
            "bottom": ((0, slice(None)), (1, slice(None)))}[side]


@pytest.mark.parametrize("side", SIDES)
def test_close_writes_its_own_faces(side):
    W, H = 31, 27
    rng_lo, rng_hi = 3, (H - 4 if side in ("left", "right") else W - 4)
    case = budget_case(W, H, 5, bcs=[BCSpec("close", side, rng_lo, rng_hi)])
    o = Oracle(case, DENSE, precision=64)
    before = {k: o.field(k, f64=True) for k in ("h", "qx", "qy")}
    o.apply_bcs()
    after = {k: o.field(k, f64=True) for k in ("h", "qx", "qy")}
    fname, fidx = staggered_face(side, W, H)
    lat = "qy" if fname == "qx" else "qx"
    ci, ni = cell_and_neighbour(side, W, H)
    along = np.arange(W if side in ("top", "bottom") else H)
    inr = (along >= rng_lo) & (along <= rng_hi)
    assert (after[fname][fidx][inr] == 0).all()
    assert (after[fname][fidx][~inr] == before[fname][fidx][~inr]).all()
    assert (after[lat][ci][inr] == before[lat][ni][inr]).all()
    assert (after["h"][ci][inr] == before["h"][ni][inr]).all()
    # nothing else changed
    for k in ("h", "qx", "qy"):
        changed = np.argwhere(after[k] != before[k])
        for j, i in changed:
            on_cell = (side == "left" and i == 0) or (side == "right" and i == W - 1) or \
                      (side == "top" and j == H - 1) or (side == "bottom" and j == 0)
            on_face = (k == fname) and ((side == "left" and i == 0) or (side == "right" and i == W - 2) or
                                        (side == "top" and j == H - 2) or (side == "bottom" and j == 0))
            assert on_cell or on_face, (k, j, i)


def test_closed_box_from_moving_start_keeps_its_volume():
    W, H, dx, dt = 45, 37, 3.0, 0.05
    case = budget_case(W, H, 7, bcs=[BCSpec("close", s) for s in SIDES], q0=0.5)
    o = Oracle(case, DENSE, precision=64)
    o.apply_bcs()
    v0 = interior_volume(o, dx)
    for _ in range(60):
        o.stencil()
        o.apply_bcs()
    v1 = interior_volume(o, dx)
    assert abs(v1 - v0) / v0 < 1e-12, (v1 - v0) / v0
    # and the water moved
    assert np.abs(o.field("qx", f64=True)[2:-3, 1:-3]).max() > 1e-3


# ---------------------------------------------------------------------------------------
# open: water only leaves, at the critical discharge sqrt(g h) h of the neighbour cell


@pytest.mark.parametrize("side", SIDES)
def test_open_outflow_direction_and_budget(side):
    W, H, dx, dt = 33, 29, 3.0, 0.05
    case = budget_case(W, H, 9, bcs=[BCSpec("close", s) for s in SIDES if s != side] +
                       [BCSpec("open", side)], q0=0.0)
    o = Oracle(case, DENSE, precision=64)
    before = o.field("h", f64=True)
    o.apply_bcs()
    fname, fidx = staggered_face(side, W, H)
    _, ni = cell_and_neighbour(side, W, H)
    face = o.field(fname, f64=True)[fidx]
    hn = before[ni]
    outward = face if side in ("right", "top") else -face
    # the corners belong to the closing sides applied before this one; the open face of
    # every other cell carries the critical discharge, directed out of the domain
    inner = slice(2, -2)
    assert (outward[inner] > 0).all()
    np.testing.assert_allclose(outward[inner], np.sqrt(G * hn[inner]) * hn[inner], rtol=1e-6)
    # the volume then falls by exactly the flux through the held faces
    total = 0.0
    for _ in range(5):
        r, dv = stencil_budget_residual(o, dx, dt)
        assert r < 1e-13, r
        assert dv < 0
        total += dv
        o.apply_bcs()
    assert total < 0


# ---------------------------------------------------------------------------------------
# z: depth set to the level minus the bed in the cell and its neighbour, face -0; a lake at
# the boundary level stays at rest


@pytest.mark.parametrize("side", SIDES)
def test_z_sets_level_and_zero_face(side):
    W, H = 29, 25
    level = 3.0
    case = budget_case(W, H, 13, bcs=[BCSpec("z", side, 2, (H if side in ("left", "right") else W) - 3,
                                             z=level)])
    o = Oracle(case, DENSE)
    o.apply_bcs()
    h = o.field("h")
    fname, fidx = staggered_face(side, W, H)
    ci, ni = cell_and_neighbour(side, W, H)
    face = o.field(fname)[fidx][2:-2]
    assert (bits(face) == bits(F32(-0.0))).all()  # the sign of the zero too
    for idx in (ci, ni):
        want = np.maximum(F32(0), F32(level) - case.bed[idx]).astype(np.float32)
        assert (bits(h[idx][2:-2]) == bits(want[2:-2])).all()


def test_z_on_all_sides_holds_a_lake_at_rest():
    W, H = 36, 30
    bed = hilly(W, H, 17)
    level = float(bed.max() + 1.0)
    case = Case(name="zlake", W=W, H=H, dx=3.0, dt=0.1, duration=1e9, bed=bed,
                h=(F32(level) - bed).astype(np.float32),
                bcs=[BCSpec("z", s, z=level) for s in SIDES])
    o = Oracle(case, DENSE, precision=64)
    for _ in range(40):
        o.step()
    z = o.field("h", f64=True) + bed
    assert np.abs(z - level).max() < 1e-5
    assert np.abs(o.field("qx", f64=True)).max() < 1e-5
    assert np.abs(o.field("qy", f64=True)).max() < 1e-5


def test_z_inflow_budget():
    # a z boundary below-level lake: the stencil moves nothing through the -0 face, so each
    # step's interior volume change is the held-face flux of the other sides (0, closed)
    # and all inflow comes from the neighbour depths the z body resets
    W, H, dx, dt = 40, 31, 3.0, 0.05
    bed = hilly(W, H, 19)
    level = float(bed.max() + 0.5)
    h = np.maximum(F32(0), F32(level - 0.4) - bed).astype(np.float32)
    case = Case(name="zin", W=W, H=H, dx=dx, dt=dt, duration=1e9, bed=bed, h=h, variable_dt=False,
                bcs=[BCSpec("close", "top"), BCSpec("close", "bottom"), BCSpec("close", "right"),
                     BCSpec("z", "left", 1, H - 2, z=level)])
    o = Oracle(case, DENSE, precision=64)
    o.apply_bcs()
    grow = 0.0
    for _ in range(30):
        r, _ = stencil_budget_residual(o, dx, dt)
        assert r < 1e-13
        hb = o.field("h", f64=True)[1:-1, 1].copy()
        v_before_bc = interior_volume(o, dx)
        o.apply_bcs()
        # the z body writes the neighbour column 1, inside the stored box
        reset = (o.field("h", f64=True)[1:-1, 1] - hb).sum() * dx * dx
        assert abs(interior_volume(o, dx) - v_before_bc - reset) < 1e-9
        grow += reset
    assert grow > 0


# ---------------------------------------------------------------------------------------
# q: the inflow of the cells from..to-1 sums to Q on every side


@pytest.mark.parametrize("side", SIDES)
def test_q_spreads_discharge_on_every_side(side):
    W, H = 24, 22
    n = H if side in ("left", "right") else W
    bed = np.zeros((H, W), np.float32)
    frm, to = 3, 12
    # cell `to` dry: the sum of sumWaterDepths excludes it, the face of it then carries 0
    ci, _ = cell_and_neighbour(side, W, H)
    edge = bed[ci]
    edge[to] = 100.0
    bed[ci] = edge
    case = Case(name="q", W=W, H=H, dx=3.0, dt=0.1, duration=1e9, bed=bed,
                bcs=[BCSpec("q", side, frm, to, z_series=([0.0, 1.0], [1.5, 1.5]),
                            q_series=([0.0, 1.0], [7.0, 7.0]))])
    o = Oracle(case, DENSE, precision=64)
    o.apply_bcs()
    fname, fidx = staggered_face(side, W, H)
    face = o.field(fname, f64=True)[fidx]
    assert abs(face[frm:to + 1].sum() * 3.0 - 7.0) < 1e-5
    assert face[to] == 0.0
    assert (face[:frm] == 0).all() and (face[to + 1:] == 0).all()
    assert (face[frm:to] > 0).all()
    _ = n


# ---------------------------------------------------------------------------------------
# symmetries of the scheme


def sym_case(W, H, seed, kst_var=False):
    bed = hilly(W, H, seed, sigma=(1.5, 4.0), n=10)
    rng = np.random.default_rng(seed)
    level = F32(np.percentile(bed, 60))
    h = np.maximum(F32(0), level - bed).astype(np.float32)
    h += rng.uniform(0.0, 0.3, (H, W)).astype(np.float32) * (h > 0)
    qx = rng.uniform(-0.4, 0.4, (H, W)).astype(np.float32) * (h > 0)
    qy = rng.uniform(-0.4, 0.4, (H, W)).astype(np.float32) * (h > 0)
    kv = rng.choice(np.array([15, 25, 35, 45, 60], np.float32), (H, W)) if kst_var else None
    return Case(name="sym", W=W, H=H, dx=3.0, dt=0.05, duration=1e9, bed=bed, h=h.astype(np.float32),
                qx=qx.astype(np.float32), qy=qy.astype(np.float32), kst_var=kv, variable_dt=False)


def run_stencil(case, steps, precision):
    o = Oracle(case, DENSE, precision=precision, with_bcs=False)
    for _ in range(steps):
        o.stencil()
    return {k: o.field(k, f64=(precision == 64)) for k in ("h", "qx", "qy")}, o.cfl_min()


@pytest.mark.parametrize("kst_var", [False, True])
def test_transpose_symmetry_bit_exact(kst_var):
    # (i, j) -> (j, i) maps qx to qy; the reference's qx and qy expression trees are
    # mirror images term by term, so in binary32 the transposed run equals the transposed
    # result bit for bit. A wrong neighbour in one of the two updates breaks it.
    N = 41
    case = sym_case(N, N, 23, kst_var)
    t = Case(**{**case.__dict__})
    t.bed = case.bed.T.copy()
    t.h, t.qx, t.qy = case.h.T.copy(), case.qy.T.copy(), case.qx.T.copy()
    if kst_var:
        t.kst_var = case.kst_var.T.copy()
    a, ca = run_stencil(case, 12, 32)
    b, cb = run_stencil(t, 12, 32)
    assert np.array_equal(bits(a["h"].T), bits(b["h"]))
    assert np.array_equal(bits(a["qx"].T), bits(b["qy"]))
    assert np.array_equal(bits(a["qy"].T), bits(b["qx"]))
    assert ca == cb


def mirror_x(case):
    m = Case(**{**case.__dict__})
    m.bed = case.bed[:, ::-1].copy()
    m.h = case.h[:, ::-1].copy()
    m.qy = case.qy[:, ::-1].copy()
    # the east face of column i becomes the east face of column W-2-i, reversed
    qx = np.zeros_like(case.qx)
    qx[:, :-1] = -case.qx[:, -2::-1]
    m.qx = qx
    return m


def mirror_y(case):
    m = Case(**{**case.__dict__})
    m.bed = case.bed[::-1, :].copy()
    m.h = case.h[::-1, :].copy()
    m.qx = case.qx[::-1, :].copy()
    qy = np.zeros_like(case.qy)
    qy[:-1, :] = -case.qy[-2::-1, :]
    m.qy = qy
    return m


@pytest.mark.parametrize("axis", ["x", "y"])
def test_mirror_symmetry_binary64(axis):
    # x -> W-1-x (or y -> H-1-y) with the normal discharge negated: the upwinded fluxes,
    # the pressure and bed terms, the friction and the held faces are all mirror
    # symmetric, so in binary64 the mirrored run equals the mirrored result up to the
    # rounding of reordered sums. Two conditions: kSt uniform (the reference takes the kSt
    # of the west or south cell for a face, which is not mirror symmetric), and depths that
    # stay non-negative. The face limiter applies min(q, h_i dx/4dt) before
    # max(q, -h_n dx/4dt), which is the mirror image of itself only while h_i + h_n >= 0;
    # held faces are not limited, so a held face that drains a cell can make its depth
    # negative. The held faces therefore start at 0 (the limited interior faces alone
    # cannot empty a cell: each takes at most a quarter of it per step).
    W, H = 43, 37
    case = sym_case(W, H, 29)
    case.qx[:, -1] = 0  # the face beyond the grid mirrors to one that reads 0
    case.qy[-1, :] = 0
    case.qx[:, 0] = case.qx[:, W - 2] = 0
    case.qy[0, :] = case.qy[H - 2, :] = 0
    m = mirror_x(case) if axis == "x" else mirror_y(case)
    a, _ = run_stencil(case, 10, 64)
    b, _ = run_stencil(m, 10, 64)
    scale = max(np.abs(a["qx"]).max(), np.abs(a["qy"]).max())
    if axis == "x":
        assert np.abs(a["h"][:, ::-1] - b["h"]).max() < 1e-12
        assert np.abs(a["qy"][:, ::-1] - b["qy"]).max() < 1e-12 * scale
        assert np.abs(-a["qx"][:, -2::-1] - b["qx"][:, :-1]).max() < 1e-12 * scale
    else:
        assert np.abs(a["h"][::-1, :] - b["h"]).max() < 1e-12
        assert np.abs(a["qx"][::-1, :] - b["qx"]).max() < 1e-12 * scale
        assert np.abs(-a["qy"][-2::-1, :] - b["qy"][:-1, :]).max() < 1e-12 * scale
    # a real flow: the check would not pass for a state at rest by accident
    assert scale > 0.1
    assert a["h"].min() >= -1e-12


def test_limiter_order_breaks_mirror_symmetry_on_negative_depth():
    # the converse of the condition above: from a held face that drains a cell below zero,
    # the mirrored run differs (the reference's min-then-max order, kernels.cu:97-98)
    W, H = 43, 37
    case = sym_case(W, H, 29)
    case.qx[:, -1] = 0
    case.qy[-1, :] = 0
    a, _ = run_stencil(case, 2, 64)
    b, _ = run_stencil(mirror_x(case), 2, 64)
    assert a["h"].min() < 0
    assert np.abs(-a["qx"][:, -2::-1] - b["qx"][:, :-1]).max() > 1e-3


# ---------------------------------------------------------------------------------------
# dt control: the x 0.99 branch (src/main.cpp:36-42)


def test_dt_decreases_by_099_towards_half_the_cfl():
    # still water of depth 8 on a flat closed box: every step's CFL is dx / sqrt(G h)
    # exactly, so from dt = 0.5 the sequence is dt *= 0.99f while dt - target > 0.001f
    W, H, dx, d = 24, 20, 3.0, 8.0
    case = Case(name="deep", W=W, H=H, dx=dx, dt=0.5, duration=1e9, bed=np.zeros((H, W), np.float32),
                h=np.full((H, W), d, np.float32), bcs=[BCSpec("close", s) for s in SIDES])
    o = Oracle(case, DENSE)
    cfl = F32(dx) / np.sqrt(F32(9.81) * F32(d), dtype=np.float32)
    target = cfl * F32(0.5)
    dt = F32(0.5)
    seen_099 = seen_hold = 0
    for n in range(200):
        o.step()
        # step 0 takes the initial CFL of 1.0 (target 0.5: no change), later steps the
        # still water's
        tg = F32(0.5) if n == 0 else target
        if dt < tg and tg - dt > F32(0.001):
            dt = dt * F32(1.001)
        elif dt > tg and dt - tg > F32(0.001):
            dt = dt * F32(0.99)
            seen_099 += 1
        else:
            seen_hold += 1
        assert F32(o.time()[1]) == dt, n
    assert seen_099 > 50 and seen_hold > 50
    assert abs(float(dt) - float(target)) <= 0.001


# ---------------------------------------------------------------------------------------
# interp against the reference's own TimeSeries (src/timeseries.cpp:14-26), compiled from
# its sources into oracle/_ref


def _ref_lib():
    p = os.path.join(ROOT, "oracle", "_ref", "libref_host.so")
    if not os.path.isfile(p):
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    L = C.CDLL(p)
    fp = C.POINTER(C.c_float)
    L.ref_series_interp.argtypes = [C.c_char_p, C.c_int, C.c_int, fp, C.c_int, fp, C.c_char_p, C.c_int]
    return L


def test_interp_matches_reference_timeseries(tmp_path):
    L = _ref_lib()
    rng = np.random.default_rng(31)
    t = np.cumsum(rng.uniform(0.5, 90.0, 40)).astype(np.float32) - F32(120.0)
    v = rng.normal(0.0, 200.0, 40).astype(np.float32)
    path = tmp_path / "s.csv"
    with open(path, "w") as f:
        f.write("time;value\n")
        for a, b in zip(t, v):
            f.write(f"{float(a)!r};{float(b)!r}\n")
    q = np.concatenate([t, t - F32(1e-3), (t[:-1] + t[1:]) * F32(0.5),
                        rng.uniform(float(t[0]) - 50, float(t[-1]) + 50, 500).astype(np.float32),
                        [t[0] - 1000, t[-1] + 1000]]).astype(np.float32)
    out = np.zeros(len(q), np.float32)
    err = C.create_string_buffer(256)
    rc = L.ref_series_interp(str(path).encode(), 0, 1, q.ctypes.data_as(C.POINTER(C.c_float)), len(q),
                             out.ctypes.data_as(C.POINTER(C.c_float)), err, 256)
    assert rc == 0, err.value
    mine = np.array([interp(t, v, float(x)) for x in q], np.float32)
    assert np.array_equal(bits(mine), bits(out))
