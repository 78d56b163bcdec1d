"""Pins the CPU oracle: against the reference's literal kernel structure, closed forms,
invariants and an independent binary64 evaluation. CPU only."""
import math
import os
import struct

import numpy as np
import pytest

from fip_amd.workloads import BCSpec, Case, config0, small_case, terrain_plane_hills, SplitMix64
from oracle_lib import DENSE, TILED, WARP_REF, Oracle, interp, pow233

F32 = np.float32


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


# ---------------------------------------------------------------------------------------
# arithmetic primitives


def test_pow233_correctly_rounded_on_sweep():
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        np.exp(rng.uniform(math.log(1e-3), math.log(100.0), 30000)),
        np.linspace(0.01, 100.0, 3001),
        [1.0, 2.0, 0.5, 100.0, 1e-3, 37.5],
    ]).astype(np.float32)
    p = np.longdouble(np.float32(2.33))
    worst = 0.0
    for x in xs:
        got = np.float32(pow233(float(x)))
        ref = np.power(np.longdouble(x), p)
        ulp = np.longdouble(np.spacing(np.float32(got)))
        worst = max(worst, float(abs(np.longdouble(got) - ref) / ulp))
    assert worst <= 0.5 + 1e-6, worst


def test_pow233_exhaustive_on_the_friction_domain():
    # every binary32 x in [0.01, 100] (the reciprocal of a face depth of 0.01 m to 100 m):
    # correctly rounded except where x^2.33f lies within 2^-50 relative of a rounding
    # midpoint (the error bound of the binary64 evaluation), checked for each miss with
    # 60-digit arithmetic
    from decimal import Decimal, getcontext
    from oracle_lib import pow233_sweep
    getcontext().prec = 60
    n, nbad, bad = pow233_sweep(0.01, 100.0)
    assert n > 1.1e8
    assert nbad <= 4, nbad
    p = Decimal(float(np.float32(2.33)))
    for x in bad:
        exact = (Decimal(float(x)).ln() * p).exp()
        got = np.float32(pow233(float(x)))
        other = np.nextafter(got, np.float32(np.inf) if Decimal(float(got)) < exact else np.float32(0))
        mid = (Decimal(float(got)) + Decimal(float(other))) / 2
        assert abs(exact - mid) / exact < Decimal(2) ** -50, (x, exact, got)


def test_pow233_tables_are_rounded_values():
    # c_i = 128 / (129 + 2 i) to 28 significant bits (nearest); M_i = c_i^-p, E[e] = 2^(e p)
    # for e = -65..54 and the series coefficients b_k = binomial(p, k) correctly rounded to
    # binary64, p = 2.33f, as the header states
    import re
    from decimal import Decimal, getcontext
    from fractions import Fraction
    getcontext().prec = 60
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "fip_amd", "csrc", "fip_detmath.h")).read()

    def table(name):
        body = re.search(name + r"\[[^]]*\] = \{(.*?)\};", src, re.S).group(1)
        return [float.fromhex(v) for v in re.findall(r"-?0x[0-9a-fp.+-]+", body)]

    def rn(x, bits):
        f = Fraction(x)
        sign = 1
        if f < 0:
            f, sign = -f, -1
        e = 0
        while f >= 2:
            f /= 2
            e += 1
        while f < 1:
            f *= 2
            e -= 1
        s = f * (1 << (bits - 1))
        k = int(s)
        if s - k > Fraction(1, 2) or (s - k == Fraction(1, 2) and k % 2):
            k += 1
        return sign * float(Fraction(k, 1 << (bits - 1)) * Fraction(2) ** e)

    p = Decimal(float(np.float32(2.33)))
    c, m, et = table("kFipPowC"), table("kFipPowM"), table("kFipPowE")
    assert len(c) == len(m) == 64 and len(et) == 120
    for i in range(64):
        assert c[i] == rn(Fraction(128, 129 + 2 * i), 28), i
        assert m[i] == rn((-(p * Decimal(c[i]).ln())).exp(), 53), i
    for k, e in enumerate(range(-65, 55)):
        assert et[k] == rn((Decimal(e) * p * Decimal(2).ln()).exp(), 53), e
    # the series coefficients b6..b1 in the Horner steps of fip_pow233_tab
    body = src[src.index("FIP_HD float fip_pow233_tab"):]
    coef = {int(k): float.fromhex(v.replace(" ", "")) * (-1 if sgn == "-" else 1)
            for v, sgn, k in [(m_.group(2), m_.group(1), m_.group(3)) for m_ in
                              re.finditer(r"([-+]) (0x[0-9a-fp.+-]+);\s*// b(\d)", body)]}
    coef[6] = float.fromhex(re.search(r"double s = (-?0x[0-9a-fp.+-]+);\s*// b6", body).group(1))
    pf = Fraction(float(np.float32(2.33)))
    for k in range(1, 7):
        num, den = Fraction(1), 1
        for j in range(k):
            num *= pf - j
            den *= j + 1
        assert coef[k] == rn(num / den, 53), k


def test_pow233_special_values():
    assert pow233(0.0) == 0.0
    assert math.isnan(pow233(float("nan")))
    assert math.isnan(pow233(-1.0))
    assert pow233(float("inf")) == float("inf")
    assert pow233(1.0) == 1.0
    # subnormal input and an input whose result overflows binary32
    assert pow233(1e-40) == 0.0
    assert pow233(1e17) == float("inf")


def test_interp_edges():
    t = [0.0, 10.0, 20.0]
    v = [1.0, 3.0, -1.0]
    assert interp(t, v, -5.0) == 1.0           # before the first key: first value
    assert interp(t, v, 0.0) == 1.0
    assert interp(t, v, 10.0) == 3.0           # exact key
    assert interp(t, v, 25.0) == -1.0          # after the last key: last value
    f = F32(5.0) / F32(10.0)
    assert interp(t, v, 5.0) == F32(f * F32(3.0) + (F32(1.0) - f) * F32(1.0))


# ---------------------------------------------------------------------------------------
# three restatements of the stencil agree


@pytest.mark.parametrize("W,H,wet", [(97, 83, True), (64, 61, True), (130, 70, False), (62, 120, True)])
def test_warp_transcription_equals_field_pass_dense(W, H, wet):
    # W-2 mod 28 not in {0,1,3} and H mod 50 not in {2,3}: the reference's store races
    # (duplicate writers of column W-2, warm-up stores of row H-2) do not occur.
    assert (W - 2) % 28 not in (0, 1, 3) and H % 50 not in (2, 3)
    case = small_case(W, H, wet=wet)
    a = Oracle(case, DENSE)
    lit = Oracle(case, WARP_REF, tile_w=32, tile_rows=50, warp_mask=0, quirks=1)
    for _ in range(25):
        a.step()
        lit.step()
    for side in (0, 1):
        for k in ("h", "qx", "qy"):
            assert same(a.field(k, side), lit.field(k, side)), (k, side)
    assert a.time() == lit.time()


def test_quirk_rows_only_affect_predicted_cells():
    # H = 102: H mod 50 = 2, so row H-2 = 100 is also stored by block-row 2 on its first
    # warm-up iteration (garbage). The clean transcription equals the field pass; the
    # literal one differs only in row H-2 (and what later flows from it).
    case = small_case(97, 102, wet=True)
    a = Oracle(case, DENSE)
    clean = Oracle(case, WARP_REF, tile_w=32, tile_rows=50, warp_mask=0, quirks=0)
    lit = Oracle(case, WARP_REF, tile_w=32, tile_rows=50, warp_mask=0, quirks=1)
    a.stencil()
    clean.stencil()
    lit.stencil()
    for k in ("h", "qx", "qy"):
        assert same(a.field(k), clean.field(k))
    dh = np.argwhere(bits(a.field("h")) != bits(lit.field("h")))
    assert len(dh) > 0 and set(dh[:, 0]) == {100}


def test_warp_mask_on_all_wet_domain_equals_dense():
    # nothing ever dries, so the mask never skips a computed row
    case = small_case(97, 83, wet=False)
    case.h = np.full((83, 97), 2.0, np.float32)
    case.bcs = [BCSpec("close", s) for s in ("left", "right", "top", "bottom")]
    a = Oracle(case, DENSE)
    m = Oracle(case, WARP_REF, tile_w=32, tile_rows=50, warp_mask=1, quirks=0)
    for _ in range(10):
        a.step()
        m.step()
    for k in ("h", "qx", "qy"):
        assert same(a.field(k), m.field(k))


@pytest.mark.parametrize("rows", [16, 32, 64])
def test_tiled_geometry_independent_when_all_wet(rows):
    case = small_case(150, 97, wet=False)
    case.h = (F32(3.0) + np.zeros((97, 150), np.float32))
    case.bcs = [BCSpec("close", s) for s in ("left", "right", "top", "bottom")]
    a = Oracle(case, DENSE)
    t = Oracle(case, TILED, tile_w=60, tile_rows=rows)
    for _ in range(12):
        a.step()
        t.step()
    for side in (0, 1):
        for k in ("h", "qx", "qy"):
            assert same(a.field(k, side), t.field(k, side))
    assert a.cfl_min() == t.cfl_min()


def test_tiled_skips_dry_tiles_and_masks_track_wet_rows():
    case = small_case(130, 70, wet=False)
    t = Oracle(case, TILED, tile_w=60, tile_rows=16)
    for _ in range(30):
        t.step()
    m = t.masks().reshape(-1, 3)
    h = t.field("h")
    for ty in range(m.shape[0]):
        for tx in range(3):
            rows = [r for r in range(16) if (int(m[ty, tx]) >> r) & 1]
            for r in range(16):
                j = ty * 16 + r
                if j >= 70:
                    continue
                cols = range(tx * 60, min(130, tx * 60 + 60))
                wet = any(h[j, i] > 1e-6 for i in cols if 1 <= i <= 128 and 1 <= j <= 68)
                assert (r in rows) == wet, (ty, tx, r)
    assert t.tiles_computed() < 15  # dry tiles far from the inflow are never computed


# ---------------------------------------------------------------------------------------
# invariants and closed forms


def closed_box(W, H, h0, seed=3):
    rng = SplitMix64(seed)
    bed = terrain_plane_hills(W, H, 3.0, 0.0, 6, rng, sigma=(3.0, 8.0)).astype(np.float32)
    return Case(name="box", W=W, H=H, dx=3.0, dt=0.1, duration=1e9, bed=bed, h=h0,
                bcs=[BCSpec("close", s) for s in ("left", "right", "top", "bottom")])


def test_lake_at_rest():
    W, H = 64, 48
    case = closed_box(W, H, None)
    level = F32(case.bed.max() + 1.0)
    case.h = (level - case.bed).astype(np.float32)
    o = Oracle(case, DENSE)
    for _ in range(50):
        o.step()
    assert np.abs(o.field("qx")).max() < 1e-4
    assert np.abs(o.field("qy")).max() < 1e-4
    z = o.field("h")[1:-1, 1:-1] + case.bed[1:-1, 1:-1]
    assert np.abs(z - level).max() < 1e-4


def test_mass_conservation_closed_box():
    W, H = 80, 60
    case = closed_box(W, H, None, seed=5)
    h0 = np.maximum(F32(0), F32(case.bed.mean() + 1.5) - case.bed).astype(np.float32)
    h0[:, : W // 3] += F32(1.0)  # a dam break inside the box
    case.h = h0
    o = Oracle(case, DENSE, precision=64)
    v0 = o.field("h", f64=True)[1:-1, 1:-1].sum()
    for _ in range(100):
        o.step()
    v1 = o.field("h", f64=True)[1:-1, 1:-1].sum()
    assert abs(v1 - v0) / v0 < 1e-12
    o32 = Oracle(case, DENSE)
    for _ in range(100):
        o32.step()
    v32 = o32.field("h", f64=True)[1:-1, 1:-1].sum()
    assert abs(v32 - v0) / v0 < 1e-5


def test_normal_flow_is_steady():
    # Uniform depth h and discharge q on a plane of slope S: the friction balance of
    # compute_rmg_qx gives q = kSt sqrt(S) h^(3.33/2); interior faces keep that q.
    W, H, dx, S, k, hd = 40, 12, 3.0, 1e-3, 30.0, 1.2
    x = np.arange(W, dtype=np.float64)
    bed = np.broadcast_to(S * dx * (W - 1 - x), (H, W)).astype(np.float32)
    q = k * math.sqrt(S) * hd ** (3.33 / 2)
    case = Case(name="slope", W=W, H=H, dx=dx, dt=0.1, duration=1e9, bed=bed, kst=k,
                h=np.full((H, W), hd, np.float32), qx=np.full((H, W), q, np.float32),
                variable_dt=False)
    o = Oracle(case, DENSE, precision=64, with_bcs=False)
    o.stencil()
    qx = o.field("qx", side=0, f64=True)
    assert abs(qx[2:-3, 2:-3] - q).max() / q < 2e-4


def test_dt_closed_form_from_dry_state():
    case = small_case(64, 48, wet=False, bc_kinds=("", "close", "close", "close"))
    o = Oracle(case, DENSE)
    dt = F32(0.1)
    for n in range(40):
        o.step()
        dt = dt * F32(1.001) if (dt < F32(0.5) and F32(0.5) - dt > F32(0.001)) else dt
        assert F32(o.time()[1]) == dt, n


def test_cfl_hand_computed():
    # still water of depth 1 on a flat bed: the CFL minimum is dx / sqrt(G h)
    W, H = 32, 20
    case = Case(name="flat", W=W, H=H, dx=3.0, dt=0.1, duration=1e9,
                bed=np.zeros((H, W), np.float32), h=np.ones((H, W), np.float32))
    o = Oracle(case, DENSE, with_bcs=False)
    o.stencil()
    assert F32(o.cfl_min()) == F32(3.0) / np.sqrt(F32(9.81) * F32(1.0), dtype=np.float32)


def test_q_boundary_spreads_discharge():
    # type q: qval = Q / (dx sum sqrt(G ah) ah) over cells from..to-1, so the inflow
    # through the wet cells (the last one dry) sums to Q
    W, H = 20, 16
    bed = np.zeros((H, W), np.float32)
    bed[10:, 0] = 100.0  # cell `to` = 10 and above are dry
    case = Case(name="q", W=W, H=H, dx=3.0, dt=0.1, duration=1e9, bed=bed,
                bcs=[BCSpec("q", "left", 2, 10, z_series=([0.0, 1.0], [1.5, 1.5]),
                            q_series=([0.0, 1.0], [7.0, 7.0]))])
    o = Oracle(case, DENSE)
    o.apply_bcs()
    q, z = o.bc_params(0)
    assert z == F32(1.5)
    qx = o.field("qx")
    assert abs(qx[2:11, 0].astype(np.float64).sum() * 3.0 - 7.0) < 1e-5
    assert qx[10, 0] == 0.0


def test_fp32_tracks_fp64():
    case = small_case(97, 83, wet=True)
    a = Oracle(case, DENSE)
    b = Oracle(case, DENSE, precision=64)
    for _ in range(20):
        a.step()
        b.step()
    for k in ("h", "qx", "qy"):
        d = np.abs(a.field(k, f64=True) - b.field(k, f64=True)).max()
        assert d < 1e-3, (k, d)


def test_floodplain_is_running_max():
    case = small_case(64, 48, wet=True)
    o = Oracle(case, DENSE)
    hmax = case.h.copy()
    for _ in range(200):
        o.step()
        if o.time()[2] % 100 == 0:
            hmax = np.maximum(hmax, o.field("h"))
    assert same(o.field("hmax"), hmax)
