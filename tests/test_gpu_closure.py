"""The learned closure on the GPU against its oracle: the matrix-core operand layout, the
convolution forward and backward passes, the restriction, Adam training, evaluation and
the inference hook (corrections, saturation fallback, mass balance)."""
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available
from fip_amd.workloads import BCSpec, Case, SplitMix64, seed_for, terrain_plane_hills

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import closure_oracle as co  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def bf16_bits(x):
    return (co.bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def rand_params(rng, scale=0.2):
    return np.where(co.live_mask(), rng.normal(0.0, scale, co.NPARAMS), 0.0).astype(np.float32)


def test_mfma_operand_layout():
    from fip_amd.sim import mfma_probe
    rng = np.random.default_rng(0)
    A = co.bf16(rng.normal(0, 1, (16, 32)).astype(np.float32))
    B = co.bf16(rng.normal(0, 1, (32, 16)).astype(np.float32))
    D = mfma_probe(bf16_bits(A), bf16_bits(B))
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.abs(D - ref).max() < 1e-4 * np.abs(ref).max(), np.abs(D - ref).max()


@pytest.mark.parametrize("act", ["gelu", "hsilu"])
@pytest.mark.parametrize("N,H,W", [(1, 9, 20), (2, 20, 70), (1, 33, 128)])
def test_forward_matches_oracle(N, H, W, act):
    from fip_amd.sim import Closure
    rng = np.random.default_rng(N * 100 + H)
    c = Closure(16, 16, 0, seed=3, activation=act)
    assert c.activation == act
    p = rand_params(rng)
    c.params = p
    feat = rng.normal(0, 1, (N, H, W, co.FEAT)).astype(np.float32)
    got = c.forward(feat)
    ref, _ = co.forward(feat, p, emulate=True, act=act)
    err = np.abs(got - ref).max()
    assert err <= 2e-3 * max(1.0, np.abs(ref).max()), err
    c.close()


@pytest.mark.parametrize("act", ["gelu", "hsilu"])
def test_backward_matches_oracle(act):
    from fip_amd.sim import Closure
    rng = np.random.default_rng(11)
    N, H, W = 2, 24, 70
    c = Closure(16, 16, 0, seed=3, activation=act)
    p = rand_params(rng)
    c.params = p
    feat = rng.normal(0, 1, (N, H, W, co.FEAT)).astype(np.float32)
    qp = rng.normal(0, 1, (N, H, W, 2)).astype(np.float32)
    qt = (qp * (1 + rng.normal(0, 0.05, qp.shape))).astype(np.float32)
    mask = (rng.uniform(0, 1, qp.shape) > 0.1).astype(np.uint8)
    loss, g = c.backward(feat, qp, qt, mask)
    rloss, rg = co.loss_and_grad(feat, p, qp, qt, mask, emulate=True, act=act)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss), (loss, rloss)
    for l in range(co.LAYERS):
        a = g[l * co.LP:(l + 1) * co.LP].astype(np.float64)
        b = rg[l * co.LP:(l + 1) * co.LP]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 2e-2, (l, rel)
    assert np.all(g[~co.live_mask()] == 0)
    c.close()


def fine_coarse_case(W, H, dx, F, scenario=0, wet=0.35):
    rng = SplitMix64(seed_for(3, scenario))
    bed = terrain_plane_hills(W, H, dx, 0.002, max(2, W * H // 16384), rng, sigma=(3.0, 12.0))
    bed = bed.astype(np.float32)
    level = np.float32(np.quantile(bed, wet))
    h = np.maximum(np.float32(0), level - bed).astype(np.float32)
    fine = Case(name="fine", W=W, H=H, dx=dx, dt=0.05, duration=1e9, bed=bed, h=h, kst=30.0,
                variable_dt=False,
                bcs=[BCSpec("close", s) for s in ("left", "right", "top", "bottom")])
    cb = co.coarse_bed(bed, F)
    hc, _, _ = co.restrict(h, np.zeros_like(h), np.zeros_like(h), F)
    coarse = Case(name="coarse", W=W // F, H=H // F, dx=dx * F, dt=0.05 * F, duration=1e9, bed=cb, h=hc,
                  kst=30.0, variable_dt=False,
                  bcs=[BCSpec("close", s) for s in ("left", "right", "top", "bottom")])
    return fine, coarse, co.bed_std(bed, F), float(cb.mean())


def test_restriction_matches_oracle():
    from fip_amd.sim import Sim, restrict
    fine, coarse, _, _ = fine_coarse_case(256, 192, 1.0, 4)
    f = Sim(fine, mode="dense")
    f.set_run(1e9)
    f.simulate(40)
    c = Sim(coarse, mode="dense")
    restrict(f, c, 4)
    hc, qxc, qyc = co.restrict(f.field("h"), f.field("qx"), f.field("qy"), 4)
    for side in (0, 1):
        assert np.array_equal(c.field("h", side).view(np.uint32), hc.view(np.uint32))
        assert np.array_equal(c.field("qx", side).view(np.uint32), qxc.view(np.uint32))
        assert np.array_equal(c.field("qy", side).view(np.uint32), qyc.view(np.uint32))
    assert c.time()[0] == f.time()[0]
    f.close()
    c.close()


def make_samples(cl, nsamp, F=4, W=256, H=256, scenario=0, fine_steps=4):
    """Samples from one fine run: restrict, one plain coarse step, 4 fine steps, store."""
    from fip_amd.sim import Sim, restrict
    fine, coarse, bstd, zref = fine_coarse_case(W, H, 1.0, F, scenario)
    f = Sim(fine, mode="dense")
    f.set_run(1e9)
    c = Sim(coarse, mode="dense")
    for k in range(nsamp):
        f.simulate(60)
        restrict(f, c, F)
        c.compute_step()
        c.bc_apply()
        f.simulate(fine_steps)
        cl.add_sample(c, f, bstd, zref, F)
    f.close()
    c.close()
    return bstd, zref


def test_training_reduces_loss_and_beats_plain():
    from fip_amd.sim import Closure
    cl = Closure(64, 64, max_samples=16, seed=5)
    make_samples(cl, 12, scenario=1)
    make_samples(cl, 4, scenario=2)
    assert cl.num_samples() == 16
    p0 = cl.params
    plain, learned0 = cl.eval(12, 4)
    cl.params = np.zeros_like(p0)
    plain2, zero = cl.eval(12, 4)
    assert zero == plain2 == plain  # no correction: the plain step
    cl.params = p0
    losses = cl.train(150, batch=8, patch=32, lr=2e-3, first=0, n=12)
    assert np.isfinite(losses).all()
    assert losses[-20:].mean() < losses[:10].mean()
    plain3, learned = cl.eval(12, 4)
    print(f"held-out mse plain {plain3:.3e} learned {learned:.3e}", flush=True)
    assert learned < plain3
    cl.close()


def test_inference_hook_mass_and_fallback():
    from fip_amd.sim import Closure, Sim
    fine, coarse, bstd, zref = fine_coarse_case(256, 256, 1.0, 4, scenario=4, wet=0.5)
    cl = Closure(64, 64, 0, seed=9)
    rng = np.random.default_rng(2)
    cl.params = rand_params(rng, 0.3)
    plain = Sim(coarse, mode="dense")
    plain.set_run(1e9)
    plain.simulate(150)
    c = Sim(coarse, mode="dense")
    cl.attach(c, bstd, zref, sat=50.0, mass_limit=1e-4)
    c.set_run(1e9)
    c.simulate(150)
    st = cl.status(c)
    assert st["fallback"] == 0 and st["max_rel_mass_error"] < 1e-5, st
    assert not np.array_equal(c.field("qx"), plain.field("qx"))
    v0 = coarse.h.astype(np.float64)[1:-1, 1:-1].sum()
    v1 = c.field("h").astype(np.float64)[1:-1, 1:-1].sum()
    assert abs(v1 - v0) / v0 < 1e-4  # closed box: volume kept
    # saturation: every face keeps the plain step
    big = Closure(64, 64, 0, seed=9)
    big.params = rand_params(rng, 0.3) * 1000
    d = Sim(coarse, mode="dense")
    big.attach(d, bstd, zref, sat=1e-6)
    d.set_run(1e9)
    d.simulate(150)
    assert big.status(d)["saturated_faces"] > 0
    assert np.array_equal(d.field("qx").view(np.uint32), plain.field("qx").view(np.uint32))
    Closure.detach(c)
    for x in (plain, c, d):
        x.close()
    cl.close()
    big.close()


@pytest.mark.parametrize("act", ["gelu", "hsilu"])
def test_fused_step_matches_layerwise_forward(act):
    """One coarse step with the closure attached equals the plain step plus
    tanh(o) 0.1 |q| on interior faces, o from the layer-by-layer network on the plain
    step's state (tolerance: the fused kernel sums layer 0 in another order and applies
    the activation to fp32 pre-activations, which the layer-by-layer form rounds to bf16)."""
    from fip_amd.sim import Closure, Sim
    fine, coarse, bstd, zref = fine_coarse_case(384, 256, 1.0, 4, scenario=6, wet=0.5)
    coarse.h = coarse.h + np.float32(0.0)
    rng = np.random.default_rng(8)
    cl = Closure(96, 64, 0, seed=2, activation=act)
    cl.params = rand_params(rng, 0.3)
    a = Sim(coarse, mode="dense")
    a.set_run(1e9)
    a.simulate(30)
    b = Sim(coarse, mode="dense")
    b.set_run(1e9)
    b.simulate(29)
    cl.attach(b, bstd, zref, sat=50.0)
    # the attached step starts from the same state as a's last step
    b.simulate(1)
    ha, qxa, qya = a.field("h"), a.field("qx"), a.field("qy")
    feat = co.features(ha, qxa, qya, coarse.bed, bstd, coarse.kst, zref)[None]
    o = cl.forward(feat)[0]
    of = cl.fused_outputs(a, bstd, zref)
    err = np.abs(of - o)
    print("fused vs layerwise outputs: max", err.max(), "at", np.unravel_index(err.argmax(), err.shape),
          "scale", np.abs(o).max(), flush=True)
    assert err.max() <= 2e-2 * max(1.0, np.abs(o).max())
    H, W = ha.shape
    m = co.interior_face_mask(H, W)
    ok = np.abs(o) <= 50.0  # faces where the network saturates keep the plain step
    assert ok.any() and (~ok).any()
    exp_x = qxa + np.where((m[..., 0] == 1) & ok[..., 0], np.tanh(o[..., 0]) * 0.1 * np.abs(qxa), 0)
    exp_y = qya + np.where((m[..., 1] == 1) & ok[..., 1], np.tanh(o[..., 1]) * 0.1 * np.abs(qya), 0)
    got_x, got_y = b.field("qx"), b.field("qy")
    # faces whose output lies within the output tolerance of the saturation limit may fall
    # on either side of it in the two evaluations (they round differently); the others
    # must agree
    far = np.abs(np.abs(o) - 50.0) > 2e-2 * max(1.0, np.abs(o).max())
    cx = np.abs(exp_x - qxa).max()
    assert cx > 0 and far.mean() > 0.5
    assert np.abs(np.where(far[..., 0], got_x - exp_x, 0)).max() <= 0.05 * cx + 1e-7
    assert np.abs(np.where(far[..., 1], got_y - exp_y, 0)).max() <= 0.05 * np.abs(exp_y - qya).max() + 1e-7
    assert np.array_equal(b.field("h"), ha)
    for x in (a, b):
        x.close()
    cl.close()


@pytest.mark.parametrize("act", ["gelu", "hsilu"])
@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("shape", [(96, 64), (50, 37), (130, 75)])
def test_fused_kernel_forms_agree(shape, form, act):
    """The persistent-tile form (weights in LDS, prefetched tiles) and the row-march form
    (58-column strips, LDS row rings) give the tile form's outputs and corrections bit for
    bit, on a multiple and non-multiples of the 16 x 16 tile and of the 58-column strip."""
    from fip_amd.sim import Closure, Sim
    Wc, Hc = shape
    fine, coarse, bstd, zref = fine_coarse_case(Wc * 4, Hc * 4, 1.0, 4, scenario=7, wet=0.5)
    rng = np.random.default_rng(5)
    p = rand_params(rng, 0.3)
    cls, sims, outs = [], [], []
    for f in (0, form):
        cl = Closure(Wc, Hc, 0, seed=2, activation=act)
        cl.params = p
        cl.set_kernel(f)
        s = Sim(coarse, mode="dense")
        s.set_run(1e9)
        s.simulate(20)
        outs.append(cl.fused_outputs(s, bstd, zref))
        cl.attach(s, bstd, zref, sat=3.0)
        s.simulate(20)
        cls.append(cl)
        sims.append(s)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    for k in ("h", "qx", "qy"):
        assert np.array_equal(sims[0].field(k).view(np.uint32), sims[1].field(k).view(np.uint32)), k
    st0, st1 = cls[0].status(sims[0]), cls[1].status(sims[1])
    assert st0["saturated_faces"] == st1["saturated_faces"] > 0
    assert st0["fallback"] == st1["fallback"] == 0
    assert st1["max_rel_mass_error"] < 1e-5, st1
    for x in sims + cls:
        x.close()


def test_closure_ensemble_deterministic():
    """Two runs of a closure-corrected ensemble give the same bits: the network reads the
    uncorrected state of every tile's halo, and the correction is a separate launch. The
    mass check's figures are the same too (its sums are added in a fixed order), and the
    relative mass error stays at rounding level."""
    from fip_amd.sim import Closure, Ensemble
    rng = np.random.default_rng(12)
    p = rand_params(rng, 0.3)
    runs, status = [], []
    for _ in range(2):
        cases, stds, zrefs = [], [], []
        for sc in (21, 22, 23):
            fine, coarse, bstd, zref = fine_coarse_case(320, 256, 1.0, 4, scenario=sc, wet=0.5)
            cases.append(coarse)
            stds.append(bstd)
            zrefs.append(zref)
        cl = Closure(80, 64, 0, seed=2)
        cl.params = p
        e = Ensemble(cases, mode="dense", variant="fast", tile_rows=32)
        cl.attach_ensemble(e, np.stack(stds), np.array(zrefs, np.float32), sat=3.0)
        e.set_run([1e9] * 3)
        e.simulate(40)
        runs.append([e.field(k, f) for k in range(3) for f in ("h", "qx", "qy")])
        status.append([cl.status(ens=e, s=k) for k in range(3)])
        e.close()
        cl.close()
    for a, b in zip(*runs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert status[0] == status[1], status
    for st in status[0]:
        assert st["fallback"] == 0 and 0.0 <= st["max_rel_mass_error"] < 1e-6, st


def test_closure_ensemble_of_mixed_sizes_equals_single_domains():
    """An ensemble of coarse scenarios of different sizes with the closure attached (the
    held-out workload asks for coarse grids of 512^2 to 1024^2 in one batched ensemble):
    every scenario equals the same scenario run alone with the closure attached, bit for
    bit, after steps in which the correction acts."""
    from fip_amd.sim import Closure, Ensemble, Sim
    rng = np.random.default_rng(31)
    p = rand_params(rng, 0.3)
    sizes = [(320, 256), (256, 256), (448, 192)]
    cases, stds, zrefs = [], [], []
    for k, (W, H) in enumerate(sizes):
        _, coarse, bstd, zref = fine_coarse_case(W, H, 1.0, 4, scenario=40 + k, wet=0.5)
        cases.append(coarse)
        stds.append(bstd)
        zrefs.append(zref)
    cl = Closure(64, 64, 0, seed=2)
    cl.params = p
    e = Ensemble(cases, mode="dense", variant="fast", tile_rows=32)
    cl.attach_ensemble(e, np.concatenate([b.ravel() for b in stds]), np.array(zrefs, np.float32), sat=3.0)
    e.set_run([1e9] * len(cases))
    e.simulate(30)
    ens = [[e.field(k, f) for f in ("h", "qx", "qy")] for k in range(len(cases))]
    st = [cl.status(ens=e, s=k) for k in range(len(cases))]
    e.close()
    for k, case in enumerate(cases):
        s = Sim(case, mode="dense", variant="fast")
        cl.attach(s, stds[k], zrefs[k], sat=3.0)
        s.set_run(1e9)
        s.simulate(30)
        one = [s.field(f) for f in ("h", "qx", "qy")]
        one_st = cl.status(s)
        # the counts are the same; the mass sums add the same terms in another order (the
        # ensemble's slots are laid out for its largest scenario)
        assert (one_st["fallback"], one_st["saturated_faces"]) == (st[k]["fallback"], st[k]["saturated_faces"])
        assert abs(one_st["max_rel_mass_error"] - st[k]["max_rel_mass_error"]) < 1e-9
        Closure.detach(s)
        s.close()
        for a, b in zip(one, ens[k]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    cl.close()
    assert max(x["max_rel_mass_error"] for x in st) < 1e-6


def test_adam_step_matches_oracle():
    """The optimiser step of training (cl_adam) equals the oracle's binary32 Adam bit for
    bit over consecutive steps, moment estimates and bias corrections included, and again
    from fresh moments after new parameters are set."""
    from fip_amd.sim import Closure
    rng = np.random.default_rng(21)
    cl = Closure(16, 16, 0, seed=4)
    w = rand_params(rng, 0.3)
    cl.params = w
    for restart in range(2):  # setting the parameters restarts the optimiser
        if restart:
            w = rand_params(rng, 0.3)
            cl.params = w
        m = np.zeros_like(w)
        v = np.zeros_like(w)
        for t in range(1, 4):
            g = np.where(co.live_mask(), rng.normal(0, 1e-2, co.NPARAMS), 0).astype(np.float32)
            cl.adam_step(g, 1e-3)
            w, m, v = co.adam(w, m, v, g, t, 1e-3)
            got = cl.params
            bad = np.count_nonzero(got.view(np.uint32) != w.view(np.uint32))
            assert bad == 0, (restart, t, bad, np.abs(got - w).max())
    cl.close()


def test_mass_error_fallback_gives_the_plain_step():
    """A mass-balance limit below any error (-1: the relative error is never negative)
    switches the domain to the plain step at the first check (the second attached step:
    the first one records the volume); from there on the domain equals a run that took the
    same two corrected steps and then ran without the closure, bit for bit."""
    from fip_amd.sim import Closure, Sim
    fine, coarse, bstd, zref = fine_coarse_case(256, 256, 1.0, 4, scenario=4, wet=0.5)
    rng = np.random.default_rng(3)
    p = rand_params(rng, 0.3)
    runs = []
    for limit in (-1.0, 1e30):
        cl = Closure(64, 64, 0, seed=9)
        cl.params = p
        s = Sim(coarse, mode="dense")
        s.set_run(1e9)
        s.simulate(20)
        cl.attach(s, bstd, zref, sat=50.0, mass_limit=limit)
        s.simulate(1)
        assert cl.status(s)["fallback"] == 0
        s.simulate(1)
        st = cl.status(s)
        if limit > 0:
            assert st["fallback"] == 0
            Closure.detach(s)  # the reference: plain steps from the same corrected state
        else:
            assert st["fallback"] == 1, st
        s.simulate(60)
        if limit < 0:
            assert cl.status(s)["fallback"] == 1
        runs.append({(k, side): s.field(k, side) for k in ("h", "qx", "qy") for side in (0, 1)})
        s.close()
        cl.close()
    for key in runs[0]:
        assert np.array_equal(runs[0][key].view(np.uint32), runs[1][key].view(np.uint32)), key


def test_activation_is_checked_and_fixed_once_attached():
    """An unknown activation is refused, and neither the activation nor the kernel form
    can change while the closure is attached (a captured step loop keeps the kernels it
    was captured with, and the forms run the mass check at different points)."""
    from fip_amd.sim import Closure, Sim, lib
    with pytest.raises(ValueError):
        Closure(16, 16, 0, seed=1, activation="relu")
    fine, coarse, bstd, zref = fine_coarse_case(128, 128, 1.0, 4, scenario=2, wet=0.5)
    cl = Closure(32, 32, 0, seed=1, activation="hsilu")
    assert lib().fip_closure_set_activation(cl.c, 7) != 0
    s = Sim(coarse, mode="dense")
    s.set_run(1e9)
    cl.attach(s, bstd, zref)
    assert lib().fip_closure_set_activation(cl.c, 0) != 0
    assert lib().fip_closure_set_kernel(cl.c, 0) != 0
    assert cl.activation == "hsilu"
    s.simulate(3)
    Closure.detach(s)
    assert lib().fip_closure_set_activation(cl.c, 0) == 0 and cl.activation == "gelu"
    s.close()
    cl.close()
