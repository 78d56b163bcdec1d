"""Training data, training and evaluation of the learned subgrid closure, driven through
the C ABI (include/fip_closure.h); every step runs on the GPU.

Fine scenarios (dx = 1 m) follow the config-0 terrain recipe with 1 m-scale ridges: a
plane sloping 0.002, Gaussian hills with sigma ~ U(3, 20) cells, a 24-cell channel 3 m deep
along y = H/2 (aligned with the 4x blocks), `nridges` straight ridges 2 cells wide and
U(0.3, 1.5) m high, uniform kSt ~ U(20, 45); dry start; right open, top and bottom closed.
The inflow is a left type-q boundary over the channel rows plus one dry bank row: Q ramps
from 0 (60 s before t = 0) to Q0 ~ U(20, 60) m^3/s at t = 0 and stays there, with the
water level of the discharge distribution at the channel bed + 2 m. A discharge boundary
puts the same volume into the fine domain and its coarse twin (the bank row, dry on both
grids, is the `to` cell that src/boundary_condition.cpp:10 leaves out of the sum); with a
fixed-level (type z) inflow the two grids draw different volumes through the boundary,
which no interior flux correction can change, and the measured depth errors would be set
by the boundary rather than by the terrain. The coarse twin is the 4x restriction (bed =
block mean, the same discharge boundary on the coarse rows). Both run with a fixed dt
(0.05 s fine, 0.2 s coarse: one coarse step = four fine steps, the pairing the closure is
trained on).

Training pairs: every `sample_every` seconds of a fine run, its state is restricted onto
the coarse grid, the coarse solver takes one plain step, the fine solver four steps, and
(restricted input, plain coarse fluxes, restricted fine fluxes) is stored.
Evaluation: held-out scenarios run plain fine, plain coarse and learned coarse (closure
after every coarse step) to the same time; the coarse results are compared with the
restricted fine run (depth RMSE, max-depth error, wet-extent IoU).
"""
from __future__ import annotations

import math
import time

import numpy as np

from .workloads import BCSpec, Case, SplitMix64, add_hills, seed_for

F = 4
DT_FINE = 0.05


def _ridges(H: int, W: int, rng: SplitMix64, n: int) -> np.ndarray:
    """n straight ridges 2 cells wide, U(0.3, 1.5) m high, random position, direction and
    length (5 to 30 % of the grid); overlapping ridges keep the higher crest."""
    ridge = np.zeros((H, W), np.float64)
    for _ in range(n):
        x0, y0 = rng.uniform(0, W), rng.uniform(0, H)
        ang = rng.uniform(0, math.pi)
        length = rng.uniform(0.05, 0.3) * min(W, H)
        height = rng.uniform(0.3, 1.5)
        steps = int(length * 2) + 1
        t = np.linspace(-0.5, 0.5, steps)
        xs = (x0 + t * length * math.cos(ang)).astype(np.int64)
        ys = (y0 + t * length * math.sin(ang)).astype(np.int64)
        for dy in (0, 1):
            for dx in (0, 1):
                yy, xx = ys + dy, xs + dx
                ok = (yy >= 0) & (yy < H) & (xx >= 0) & (xx < W)
                ridge[yy[ok], xx[ok]] = np.maximum(ridge[yy[ok], xx[ok]], height)
    return ridge


def fine_case(scenario: int, W: int = 1024, H: int = 1024, nridges: int = 200, duration: float = 3600.0,
              seed_config: int = 3) -> Case:
    rng = SplitMix64(seed_for(seed_config, scenario))
    dx = 1.0
    x = np.arange(W, dtype=np.float64)
    bed = np.broadcast_to(0.002 * dx * (W - 1 - x), (H, W)).copy()
    nh = max(4, int(round(64 * W * H / (512.0 * 512.0))))
    add_hills(bed, rng, nh, (0.0, 5.0), (3.0, 20.0))
    c = (H // 2) // F * F
    bed[c - 12:c + 12, :] -= 3.0
    bed += _ridges(H, W, rng, nridges)
    kst = float(np.float32(rng.uniform(20.0, 45.0)))
    q0 = rng.uniform(20.0, 60.0)
    bed32 = bed.astype(np.float32)
    zc = float(np.float32(bed32[c - 12:c + 12, 0].min()) + np.float32(2.0))
    bcs = [BCSpec("q", "left", c - 12, c + 12, z_series=([0.0, duration], [zc, zc]),
                  q_series=([-60.0, 0.0, duration], [0.0, q0, q0])),
           BCSpec("open", "right"), BCSpec("close", "top"), BCSpec("close", "bottom")]
    return Case(name=f"fine_{scenario}", W=W, H=H, dx=dx, dt=DT_FINE, duration=duration, bed=bed32,
                kst=kst, variable_dt=False, bcs=bcs)


def coarse_of(fine: Case) -> tuple[Case, np.ndarray, float]:
    """The 4x coarse twin, the subgrid bed std and the feature reference level."""
    W, H = fine.W // F, fine.H // F
    b = fine.bed.astype(np.float64).reshape(H, F, W, F)
    cb = (b.sum(axis=(1, 3)) / (F * F)).astype(np.float32)
    bstd = np.sqrt(((b - b.mean(axis=(1, 3), keepdims=True)) ** 2).mean(axis=(1, 3))).astype(np.float32)
    bcs = [BCSpec(bc.type, bc.side, bc.frm // F if bc.frm is not None else None,
                  bc.to // F if bc.to is not None else None, z=bc.z, q=bc.q,
                  z_series=bc.z_series, q_series=bc.q_series) for bc in fine.bcs]
    coarse = Case(name=fine.name + "_c", W=W, H=H, dx=fine.dx * F, dt=fine.dt * F, duration=fine.duration,
                  bed=cb, kst=fine.kst, variable_dt=False, bcs=bcs)
    return coarse, bstd, float(cb.mean())


def block_mean(a: np.ndarray) -> np.ndarray:
    H, W = a.shape
    return (a.astype(np.float64).reshape(H // F, F, W // F, F).sum(axis=(1, 3)) / (F * F)).astype(np.float32)


def generate_samples(cl, scenarios, duration, sample_every=60.0, variant="fast", W=1024, log=print):
    """Adds (duration / sample_every) samples per scenario to the closure's store: the input
    of a sample is the restricted state of the fine run, the target the restricted fine
    discharges one coarse step (four fine steps) later."""
    from .sim import Sim, restrict
    per = int(round(sample_every / DT_FINE)) - F
    t0 = time.perf_counter()
    for k, s in enumerate(scenarios):
        if k and k % 32 == 0:
            log(f"samples: {cl.num_samples()} after {k} scenarios, {time.perf_counter() - t0:.1f} s")
        fine = fine_case(s, W=W, H=W, duration=duration)
        coarse, bstd, zref = coarse_of(fine)
        f = Sim(fine, mode="dense", variant=variant)
        f.set_run(1e12)
        c = Sim(coarse, mode="dense", variant=variant)
        for _ in range(int(duration // sample_every)):
            f.simulate(per)
            restrict(f, c, F)
            c.compute_step()
            c.bc_apply()
            f.simulate(F)
            cl.add_sample(c, f, bstd, zref, F)
        f.close()
        c.close()
    log(f"samples: {cl.num_samples()} in {time.perf_counter() - t0:.1f} s")


def evaluate(cl, scenarios, duration, variant="fast", W=1024, log=print, sat=2.65, nets=None):
    """Plain fine vs plain coarse vs learned coarse on held-out scenarios, as ensembles.
    W: the fine grid size, one for all scenarios or one per scenario (the ensembles then
    hold scenarios of different sizes); the ridge count scales with the area (200 on
    2048^2). nets: networks [(activation, params), ...] that cl runs in turn on the same
    scenarios (default: cl as it is); the first gives the learned_* figures, each other one
    a sub-object under its activation's name. Returns metrics and timings."""
    from . import metrics as co
    from .sim import Ensemble, tune_ensemble
    sizes = list(W) if isinstance(W, (list, tuple)) else [W] * len(scenarios)
    fines = [fine_case(s, W=w, H=w, duration=duration, nridges=max(1, round(200 * (w / 2048.0) ** 2)))
             for s, w in zip(scenarios, sizes)]
    pairs = [coarse_of(f) for f in fines]
    coarses = [p[0] for p in pairs]
    bstd = np.concatenate([p[1].ravel() for p in pairs])  # each scenario's in turn (sizes may differ)
    zref = np.array([p[2] for p in pairs], np.float32)
    out = {"scenarios": len(scenarios), "duration_s": duration,
           "fine": ", ".join(sorted({f"{f.W}x{f.H}" for f in fines})) + " dx=1",
           "coarse": ", ".join(sorted({f"{c.W}x{c.H}" for c in coarses})) + " dx=4"}

    def run(cases, closure, record=True):
        # each ensemble runs with the launch shape the ensemble autotuner picks for it
        best = tune_ensemble(cases, mode="dense", variant=variant, steps=10)
        e = Ensemble(cases, mode="dense", variant=variant, tile_rows=32)
        e.set_launch(best["waves_per_block"], 0)
        if record:
            out.setdefault("waves_per_block_fine_plain_learned", []).append(best["waves_per_block"])
        if closure is not None:
            closure.attach_ensemble(e, bstd, zref, sat=sat)
        e.set_run([duration] * len(cases))
        e.simulate(200)  # graphs built, warm
        t_warm = e.time(0)[0]  # every scenario of an ensemble here has the same fixed dt
        t0 = time.perf_counter()
        # in stretches of 30 simulated minutes, so that a long run reports its progress
        while e.simulate(int(round(1800.0 / cases[0].dt)))[0] > 0:
            log(f"{cases[0].W}x{cases[0].H}...: {e.time(0)[0]:.0f} s simulated, {time.perf_counter() - t0:.1f} s")
        wall = time.perf_counter() - t0
        # simulated hours per wall second over the timed steps (after the warm-up)
        return e, len(cases) * (duration - t_warm) / 3600.0 / wall, wall

    ef, rf, wf = run(fines, None)
    ep, rp, wp = run(coarses, None)
    n = len(scenarios)
    refs = [(block_mean(ef.field(k, "h")), block_mean(ef.field(k, "hmax"))) for k in range(n)]
    ef.close()

    def scores(e):
        rm, me, iou = [], [], []
        for k, (href, hmref) in enumerate(refs):
            h = e.field(k, "h")
            rm.append(co.depth_rmse(h, href))
            me.append(co.max_depth_error(e.field(k, "hmax"), hmref))
            iou.append(co.wet_iou(h, href))
        return {"depth_rmse_m": round(float(np.mean(rm)), 4), "max_depth_err_m": round(float(np.mean(me)), 4),
                "wet_iou": round(float(np.mean(iou)), 4)}

    for name, rate, wall in (("fine", rf, wf), ("plain_coarse", rp, wp)):
        out[name + "_sim_h_per_wall_s"] = round(rate, 3)
        out[name + "_wall_s"] = round(wall, 2)
    plain = scores(ep)
    ep.close()
    for i, (act, params) in enumerate(nets or [(cl.activation, None)]):
        cl.set_activation(act)
        if params is not None:
            cl.params = params
        el, rl, wl = run(coarses, cl, record=i == 0)
        learned = scores(el)
        fb, mass = 0, 0.0
        for k in range(n):
            st = cl.status(ens=el, s=k)
            fb += st["fallback"]
            mass = max(mass, st["max_rel_mass_error"])
        el.close()
        res = {"activation": act, "learned_coarse_sim_h_per_wall_s": round(rl, 3),
               "learned_coarse_wall_s": round(wl, 2), "speedup_learned_vs_fine": round(rl / rf, 2),
               **{"learned_" + k: v for k, v in learned.items()}, "fallback_scenarios": fb,
               "max_rel_mass_error": mass}
        if i == 0:
            out.update({"learned_activation": act,
                        "learned_coarse_sim_h_per_wall_s": res["learned_coarse_sim_h_per_wall_s"],
                        "learned_coarse_wall_s": res["learned_coarse_wall_s"],
                        "speedup_learned_vs_fine": res["speedup_learned_vs_fine"],
                        **{"plain_" + k: v for k, v in plain.items()},
                        **{k: v for k, v in res.items() if k.startswith("learned_") or k in
                           ("fallback_scenarios", "max_rel_mass_error")}})
        else:
            out[act] = {k: v for k, v in res.items() if k != "activation"}
    log(str(out))
    return out


def experiment(n_train=16, n_test=8, duration=1800.0, sample_every=60.0, steps=1500, batch=32, patch=128,
               lr=1e-3, W=1024, eval_W=2048, activations=("gelu",), eval_duration=None, log=print):
    """Generate samples (fine grids of W^2), train, evaluate on held-out scenarios with
    fine grids of eval_W^2, or eval_W[k]^2 for held-out scenario k (the network is
    convolutional, so it applies to any size). The
    held-out samples (four per held-out scenario, one a minute) come first in the store, the
    training samples (one every `sample_every` seconds of each training run) after them.
    The held-out runs last eval_duration seconds (default: duration). One network per
    activation, each trained from the same initial weights on the same
    samples; the first gives the top-level figures, each other one a sub-object under its
    name (training figures and the held-out evaluation)."""
    from .sim import Closure
    Wc = W // F
    nsamp = n_train * int(duration // sample_every) + n_test * 4
    cl = Closure(Wc, Wc, max_samples=nsamp, seed=11, activation=activations[0])
    p0 = cl.params
    generate_samples(cl, range(1000, 1000 + n_test), 240.0, 60.0, W=W, log=log)
    ntest = cl.num_samples()
    generate_samples(cl, range(n_train), duration, sample_every, W=W, log=log)
    runs, nets = [], []
    for act in activations:
        cl.set_activation(act)
        cl.params = p0
        plain0, learned0 = cl.eval(0, ntest)
        t0 = time.perf_counter()
        losses = cl.train(steps, batch=batch, patch=patch, lr=lr, first=ntest)
        twall = time.perf_counter() - t0
        plain1, learned1 = cl.eval(0, ntest)
        r = {"train_samples": cl.num_samples() - ntest, "train_steps": steps, "train_wall_s": round(twall, 2),
             "loss_first10": float(np.mean(losses[:10])), "loss_last50": float(np.mean(losses[-50:])),
             "heldout_flux_mse_plain": plain1, "heldout_flux_mse_learned": learned1,
             "heldout_flux_mse_untrained": learned0}
        log(f"{act}: {r}")
        runs.append(r)
        nets.append((act, cl.params))
    res = dict(runs[0])
    for act, r in zip(activations[1:], runs[1:]):
        res[act] = r
    res["eval"] = evaluate(cl, range(1000, 1000 + n_test), eval_duration or duration, W=eval_W, log=log,
                           nets=nets)
    cl.close()
    return res


def time_ensemble_step(steps=300, n=8, W=1024, warm=200, sat=1e30, seed=11, form=None, activation="gelu",
                       log=print):
    """Per-step wall time of the held-out coarse twins (n scenarios of (W/4)^2) as a plain
    ensemble and with a closure of seeded random weights attached (every face corrected:
    sat = 1e30). Returns (plain_us, learned_us)."""
    from .sim import Closure, Ensemble
    fines = [fine_case(1000 + s, W=W, H=W, duration=1800.0) for s in range(n)]
    pairs = [coarse_of(f) for f in fines]
    coarses = [p[0] for p in pairs]
    bstd = np.stack([p[1] for p in pairs])
    zref = np.array([p[2] for p in pairs], np.float32)
    cl = Closure(W // F, W // F, max_samples=0, seed=seed, activation=activation)
    if form is not None:
        cl.set_kernel(form)
    out = []
    for label, closure in (("plain", None), ("learned", cl)):
        e = Ensemble(coarses, mode="dense", variant="fast", tile_rows=32)
        if closure is not None:
            closure.attach_ensemble(e, bstd, zref, sat=sat)
        e.set_run([1e9] * n)
        e.simulate(warm)
        t0 = time.perf_counter()
        e.simulate(steps)
        e.time(0)
        us = (time.perf_counter() - t0) / steps * 1e6
        log(f"{label}: {us:.1f} us per step")
        out.append(us)
        e.close()
    cl.close()
    return tuple(out)


def time_training(steps=200, scenarios=2, W=1024, duration=600.0, batch=32, patch=128, lr=1e-3, log=print):
    """Wall time of `steps` training steps (batch x patch^2 coarse pixels each: forward,
    loss, backward, Adam) on samples of `scenarios` fine runs of W^2, after 10 warm-up
    steps. Returns a dict."""
    from .sim import Closure
    n = scenarios * int(duration // 60.0)
    cl = Closure(W // F, W // F, max_samples=n, seed=11)
    generate_samples(cl, range(scenarios), duration, 60.0, W=W, log=log)
    cl.train(10, batch=batch, patch=patch, lr=lr)
    t0 = time.perf_counter()
    losses = cl.train(steps, batch=batch, patch=patch, lr=lr, seed=9)
    wall = time.perf_counter() - t0
    cl.close()
    out = {"train_steps": steps, "batch": batch, "patch": patch, "samples": n, "wall_s": round(wall, 3),
           "ms_per_step": round(wall / steps * 1e3, 3), "pixels_per_step": batch * patch * patch,
           "loss_first": float(losses[0]), "loss_last": float(losses[-1])}
    log(str(out))
    return out


def drift(scenarios=range(1000, 1004), times=(300, 900, 1800, 3600, 7200, 10800), W=2048, log=print):
    """Plain coarse against the restricted plain fine run of the same held-out scenarios
    over time, without any closure: depth RMSE, wet-extent IoU and the wet fractions of
    both at each time. Returns a list of dicts."""
    from . import metrics as co
    from .sim import Sim
    out = []
    for s in scenarios:
        fine = fine_case(s, W=W, H=W, duration=float(times[-1]))
        coarse, _, _ = coarse_of(fine)
        f = Sim(fine, mode="dense", variant="fast")
        c = Sim(coarse, mode="dense", variant="fast")
        f.set_run(1e12)
        c.set_run(1e12)
        for T in times:
            f.simulate(int(round((T - f.time()[0]) / fine.dt)))
            c.simulate(int(round((T - c.time()[0]) / coarse.dt)))
            href = block_mean(f.field("h"))
            hc = c.field("h")
            rec = {"scenario": s, "t_s": T, "depth_rmse_m": round(co.depth_rmse(hc, href), 4),
                   "wet_iou": round(co.wet_iou(hc, href), 4),
                   "wet_fraction_fine": round(float((href > 0.01).mean()), 5),
                   "wet_fraction_coarse": round(float((hc > 0.01).mean()), 5)}
            log(str(rec))
            out.append(rec)
        f.close()
        c.close()
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time the coarse ensemble step with and without the closure, "
                                             "or (--train-steps) the training step")
    ap.add_argument("--train-steps", type=int, default=0, help="time this many training steps instead")
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--scenarios", type=int, default=8)
    ap.add_argument("--fine", type=int, default=1024, help="fine grid size (the coarse one is a quarter)")
    ap.add_argument("--warm", type=int, default=200)
    ap.add_argument("--form", type=int, default=None, help="fused kernel form (fip_closure_set_kernel)")
    ap.add_argument("--sat", type=float, default=1e30, help="saturation limit of the network outputs")
    ap.add_argument("--activation", default="gelu", choices=["gelu", "hsilu"], help="the network's activation")
    ap.add_argument("--drift", action="store_true",
                    help="plain coarse against the restricted fine run over 3 h instead (no closure)")
    a = ap.parse_args()
    if a.drift:
        drift()
    elif a.train_steps:
        time_training(a.train_steps)
    else:
        time_ensemble_step(a.steps, a.scenarios, a.fine, a.warm, sat=a.sat, form=a.form, activation=a.activation)
