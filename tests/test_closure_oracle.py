"""Pins the closure oracle (oracle/closure_oracle.py) on the CPU: the restriction by
closed forms and conservation, the network gradient by central finite differences in
binary64, Adam by its first-step closed form, the metrics by hand-computed cases."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import closure_oracle as co  # noqa: E402


def rand_params(rng, scale=0.3):
    p = rng.normal(0.0, scale, co.NPARAMS)
    return np.where(co.live_mask(), p, 0.0)


def test_restrict_conserves_volume_and_discharge():
    rng = np.random.default_rng(3)
    F, Hc, Wc = 4, 6, 7
    h = rng.uniform(0, 2, (Hc * F, Wc * F)).astype(np.float32)
    qx = rng.normal(0, 1, h.shape).astype(np.float32)
    qy = rng.normal(0, 1, h.shape).astype(np.float32)
    hc, qxc, qyc = co.restrict(h, qx, qy, F)
    # volume: coarse cell area F^2
    assert abs(hc.astype(np.float64).sum() * F * F - h.astype(np.float64).sum()) < 1e-4
    # total discharge through each coarse east face = sum over its F fine faces (width 1)
    for J in range(Hc):
        for I in range(Wc):
            fine = qx[F * J:F * J + F, F * I + F - 1].astype(np.float64).sum()
            assert abs(qxc[J, I] * F - fine) < 1e-5
            fine = qy[F * J + F - 1, F * I:F * I + F].astype(np.float64).sum()
            assert abs(qyc[J, I] * F - fine) < 1e-5


def test_restrict_rounds_a_wide_block_in_sequential_order():
    # a block holding nearly dry cells (1e-35 m) beside deep ones: the binary64 sum is not
    # exact, and the sequential order of the definition gives the correctly rounded mean
    # here where summing row sums first does not
    from fractions import Fraction
    v = np.array([0.0, 5.430896994236511e-17, 0.6958233714103699, 0.7572036385536194,
                  3.762364820236067e-35, 5.900398158651158e-16, 0.6636788845062256, 0.6971497535705566,
                  1.932195512948111e-33, 0.8240683674812317, 0.8419486880302429, 0.7811002135276794,
                  4.914276678898646e-32, 0.7862984538078308, 0.8122485280036926, 0.875024139881134],
                 np.float32).reshape(4, 4)
    hc, _, _ = co.restrict(v, v, v, 4)
    exact = sum(Fraction(float(x)) for x in v.ravel()) / 16
    assert hc[0, 0] == np.float32(float(exact))
    rows_first = np.float32(v.astype(np.float64).sum(axis=1).sum() / 16)
    assert rows_first != hc[0, 0]


def test_restrict_brute_force_and_constants():
    F = 4
    h = np.arange(8 * 12, dtype=np.float32).reshape(8, 12) * np.float32(0.25)
    hc, _, _ = co.restrict(h, h, h, F)
    for J in range(2):
        for I in range(3):
            blk = [float(h[F * J + r, F * I + q]) for r in range(F) for q in range(F)]
            assert hc[J, I] == np.float32(sum(blk) / 16.0)
    c = np.full((8, 8), 1.25, np.float32)
    a, b, d = co.restrict(c, c, c, F)
    assert (a == 1.25).all() and (b == 1.25).all() and (d == 1.25).all()
    bed = np.zeros((8, 8), np.float32)
    bed[::2, :] = 1.0
    assert np.allclose(co.bed_std(bed, F), 0.5)
    assert np.allclose(co.coarse_bed(bed, F), 0.5)


@pytest.mark.parametrize("act", ["gelu", "hsilu"])
def test_backward_matches_finite_differences(act):
    rng = np.random.default_rng(5)
    N, H, W = 2, 5, 6
    feat = rng.normal(0, 1, (N, H, W, co.FEAT))
    qp = rng.normal(0, 1, (N, H, W, 2))
    qt = qp * (1 + rng.normal(0, 0.05, qp.shape))
    mask = (rng.uniform(0, 1, qp.shape) > 0.2).astype(np.uint8)
    p = rand_params(rng)
    loss, g = co.loss_and_grad(feat, p, qp, qt, mask, emulate=False, act=act)
    live = np.nonzero(co.live_mask())[0]
    pick = np.concatenate([rng.choice(live, 40, replace=False),
                           [co.LP - 1 - co.C + 5, 3 * co.LP + co.TAPS * co.C * co.C + 1]])  # two biases
    eps = 1e-6
    for i in pick:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = co.loss_and_grad(feat, pp, qp, qt, mask, emulate=False, act=act)
        lm, _ = co.loss_and_grad(feat, pm, qp, qt, mask, emulate=False, act=act)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
    # padded parameters get no gradient
    assert np.all(g[~co.live_mask()] == 0)


def test_conv_tap_order_against_correlate2d():
    # tap t = (dy + 1) * 3 + (dx + 1) reads the input at (y + dy, x + dx): for one input and
    # one output channel that is scipy's 2-D cross-correlation with zero fill, kernel
    # k[dy + 1, dx + 1] = w[t]; summed over channels for the multi-channel case
    from scipy.signal import correlate2d
    rng = np.random.default_rng(8)
    N, H, W, Ci, Co = 2, 7, 9, 3, 4
    x = rng.normal(0, 1, (N, H, W, Ci))
    w = rng.normal(0, 1, (co.TAPS, Ci, Co))
    b = rng.normal(0, 1, Co)
    got = co.conv(x, w, b)
    want = np.zeros((N, H, W, Co))
    for n in range(N):
        for o in range(Co):
            want[n, :, :, o] = b[o]
            for c in range(Ci):
                k = w[:, c, o].reshape(3, 3)
                want[n, :, :, o] += correlate2d(x[n, :, :, c], k, mode="same", boundary="fill")
    assert np.abs(got - want).max() < 1e-12


def test_conv_tap_order_brute_force_delta():
    # a unit input at (y0, x0) and a single nonzero tap (dy, dx) give a single output at
    # (y0 - dy, x0 - dx); at the image edge a tap that would read outside gives nothing
    H, W = 6, 8
    for t in range(co.TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        w = np.zeros((co.TAPS, 1, 1))
        w[t, 0, 0] = 2.0
        for (y0, x0) in ((3, 4), (0, 0), (H - 1, W - 1)):
            x = np.zeros((1, H, W, 1))
            x[0, y0, x0, 0] = 1.0
            out = co.conv(x, w)[0, :, :, 0]
            nz = np.argwhere(out != 0)
            y, xx = y0 - dy, x0 - dx
            if 0 <= y < H and 0 <= xx < W:
                assert nz.tolist() == [[y, xx]] and out[y, xx] == 2.0, (t, y0, x0)
            else:
                assert len(nz) == 0, (t, y0, x0)


def test_conv_input_and_weight_grads_are_adjoints():
    # <conv(x, w), dz> = <x, conv_input_grad(dz, w)> = <w, conv_weight_grad(x, dz)>: the
    # two backward convolutions are the adjoints of the forward one, tap order included
    rng = np.random.default_rng(9)
    x = rng.normal(0, 1, (2, 5, 7, 3))
    w = rng.normal(0, 1, (co.TAPS, 3, 4))
    dz = rng.normal(0, 1, (2, 5, 7, 4))
    lhs = (co.conv(x, w) * dz).sum()
    assert abs(lhs - (x * co.conv_input_grad(dz, w)).sum()) < 1e-10
    dw, db = co.conv_weight_grad(x, dz)
    assert abs(lhs - (w * dw).sum()) < 1e-10
    assert np.allclose(db, dz.reshape(-1, 4).sum(axis=0))


# hard SiLU's derivative jumps at -2 and 2, so rounding a pre-activation to bf16 can move
# it across a jump: its emulated gradient follows the exact one less closely
@pytest.mark.parametrize("act,min_cos", [("gelu", 0.995), ("hsilu", 0.99)])
def test_emulated_gradient_tracks_exact(act, min_cos):
    rng = np.random.default_rng(6)
    N, H, W = 2, 12, 10
    feat = rng.normal(0, 1, (N, H, W, co.FEAT))
    qp = rng.normal(0, 1, (N, H, W, 2))
    qt = qp * 1.03
    mask = np.ones(qp.shape, np.uint8)
    p = rand_params(rng, 0.2)
    _, g64 = co.loss_and_grad(feat, p, qp, qt, mask, emulate=False, act=act)
    _, g16 = co.loss_and_grad(feat, p, qp, qt, mask, emulate=True, act=act)
    cos = float(g64 @ g16 / (np.linalg.norm(g64) * np.linalg.norm(g16)))
    assert cos > min_cos, cos


def test_gelu_derivative():
    x = np.linspace(-4, 4, 101)
    fd = (co.gelu(x + 1e-6) - co.gelu(x - 1e-6)) / 2e-6
    assert np.abs(fd - co.dgelu(x)).max() < 1e-7
    assert abs(co.gelu(0.0)) == 0.0 and abs(co.gelu(10.0) - 10.0) < 1e-6


def test_hsilu_closed_forms_and_derivative():
    # x for x >= 2, -0 for x <= -2, x^2 / 4 + x / 2 between
    x = np.array([-5.0, -2.0, -1.0, 0.0, 0.5, 1.0, 2.0, 7.0])
    assert co.hsilu(x).tolist() == [0.0, 0.0, -0.25, 0.0, 0.3125, 0.75, 2.0, 7.0]
    assert np.signbit(co.hsilu(x)[:2]).all()
    # derivative by central differences away from the jumps at -2 and 2 (exact there: the
    # function is quadratic or linear on each side)
    x = np.linspace(-4, 4, 161)
    x = x[np.abs(np.abs(x) - 2) > 1e-3]
    fd = (co.hsilu(x + 1e-6) - co.hsilu(x - 1e-6)) / 2e-6
    assert np.abs(fd - co.dhsilu(x)).max() < 1e-8
    assert co.dhsilu(np.array([-3.0, -2.0, 0.0, 1.0, 2.0, 3.0])).tolist() == [0.0, 0.0, 0.5, 1.0, 1.0, 1.0]
    # in float32 the value is the correctly rounded x * RN32(x / 4 + 1 / 2): the quarter and
    # the product of two 24-bit significands are exact in binary64
    rng = np.random.default_rng(3)
    x32 = (rng.normal(0, 3, 100000)).astype(np.float32)
    t = np.float32(np.clip(x32.astype(np.float64) / 4 + 0.5, 0, 1))
    want = np.float32(x32.astype(np.float64) * t.astype(np.float64))
    got = co.hsilu(x32)
    assert got.dtype == np.float32 and (got.view(np.uint32) == want.view(np.uint32)).all()


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(1)
    w = np.where(co.live_mask(), rng.normal(0, 1, co.NPARAMS), 0).astype(np.float32)
    g = np.where(co.live_mask(), rng.normal(0, 1, co.NPARAMS), 0).astype(np.float32)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    w1, m1, v1 = co.adam(w, m, v, g, 1, 1e-3)
    live = co.live_mask() & (np.abs(g) > 1e-3)
    step = (w - w1)[live]
    assert np.allclose(step, 1e-3 * np.sign(g[live]), rtol=1e-4)
    assert np.all(w1[~co.live_mask()] == 0)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 3.14159, -0.0], np.float32)
    r = co.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0 + 2 ** -6)  # tie to even (up)
    assert abs(r[3] - 3.140625) < 1e-7


def test_metrics():
    a = np.zeros((4, 4))
    b = np.zeros((4, 4))
    b[0, 0] = 0.2
    assert abs(co.depth_rmse(a, b) - np.sqrt(0.04 / 16)) < 1e-12
    assert co.max_depth_error(a, b) == pytest.approx(0.2)
    a[0, 0] = 0.5
    a[1, 1] = 0.5
    assert co.wet_iou(a, b) == pytest.approx(0.5)
    assert co.interior_face_mask(8, 8)[:, :, 0].sum() == 4 * 5


def test_product_metrics_equal_oracle():
    from fip_amd import metrics
    rng = np.random.default_rng(4)
    a = np.maximum(0, rng.normal(0, 0.3, (20, 30)))
    b = np.maximum(0, rng.normal(0, 0.3, (20, 30)))
    assert metrics.depth_rmse(a, b) == co.depth_rmse(a, b)
    assert metrics.max_depth_error(a, b) == co.max_depth_error(a, b)
    assert metrics.wet_iou(a, b) == co.wet_iou(a, b)
