"""Learned-closure workflow on one GPU: a plain fine run supplies training data and is the
judge; a coarse run with the trained closure is compared with it and with the plain coarse
run, and the closure is dropped (fallback to the plain coarse method) when its Nusselt
number misses the fine run's by more than a tolerance or when it does no better than the plain
coarse run on the half of the time unit that selects it.

All compute runs in libsdc_rbc.so (rbc_step, rbc_closure_*); this module only sequences it.
"""
import time

import numpy as np

from .initial import random_theta
from .rbc import Closure, Stepper


def _mk(cfg, precision):
    return Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                   precision=precision)


def _record(st, nsteps, every, advance, log=None, tag=""):
    """Advance nsteps in chunks of `every` with advance(k); returns (times of the samples,
    volume Nu, plate Nu (mean of both plates), the E_h spectra of the samples (rows, cut to a
    common length), number of spectra, wall s)."""
    t_nu, nu_v, nu_p = [], [], []
    spec, nspec = [], 0
    wall = 0.0
    done = 0
    while done < nsteps:
        k = min(every, nsteps - done)
        st.synchronize()
        t0 = time.perf_counter()
        advance(k)
        st.synchronize()
        wall += time.perf_counter() - t0
        done += k
        nu = st.nusselt()
        t_nu.append(done)
        nu_v.append(nu[2])
        nu_p.append(0.5 * (nu[0] + nu[1]))
        E, _ = st.energy_spectrum()
        spec.append(np.array(E, copy=True))
        nspec += 1
        if log and nspec % 10 == 0:
            log(tag, "step", done, "of", nsteps, "Nu", nu.tolist(), "wall", round(wall, 2))
    n = min(len(E) for E in spec)
    return np.array(t_nu), np.array(nu_v), np.array(nu_p), np.array([E[:n] for E in spec]), nspec, wall


def run(fine, coarse, *, T=1.0, spinup_steps=2000, samples=10, train_steps=2000, batch=65536, lr=1e-3,
        nz=0, seed=3, tol=0.02, every=200, fine_precision="f32", coarse_precision="f32", max_ratio=10.0,
        check_every=200, closure_every=1, variants=None, log=None):
    """The whole experiment.  fine / coarse: dicts with M, N, L, Lx, Ly, Ra, Pr, dt (same box,
    Ra, Pr).  closure_every: the network is evaluated on every closure_every-th coarse step and
    the steps between reuse its last correction (a list runs one corrected coarse run per value).
    variants: corrected runs as dicts {"every", "gain", "kfrac"} (gain times the correction on
    the modes with |k| <= kfrac k_max, none above), in place of closure_every.  A variant with
    "control": True is run and reported but never selected: with a tiny gain it is the plain
    coarse run under a small perturbation, whose errors show how much a single chaotic
    trajectory's time means vary.

    Errors are measured on both halves of the time unit: the first half (whose fine states also
    supply the training data) selects the reported corrected run -- the smallest of its Nu and
    spectrum errors relative to the plain coarse run's -- and the second half is reported, so
    that the choice does not see the reported figures.  Returns a dict of the measurements."""
    say = log or (lambda *a: None)
    out = {}
    F = _mk(fine, fine_precision)
    F.set_state(*random_theta(fine["M"], fine["N"], fine["L"], seed=seed, amp=1e-3))
    t0 = time.perf_counter()
    F.step_time(spinup_steps)
    F.synchronize()
    out["fine_spinup_s"] = time.perf_counter() - t0
    say("spin-up done", F.nusselt())
    Cc = _mk(coarse, coarse_precision)
    n_f = int(round(T / fine["dt"]))
    n_c = int(round(T / coarse["dt"]))
    cl = Closure(Cc, nz=nz, capacity_points=0)  # probe the point count
    npts = cl.npts
    cl.close()
    cl = Closure(Cc, nz=nz, capacity_points=samples * npts)
    cl.restrict_state(F)
    c0 = Cc.get_state()
    # fine run over T, sampling training data at evenly spaced times of its first half
    gap = max(1, (n_f // 2) // samples)
    t_sample = 0.0
    done_f = 0

    def fine_advance(k):
        # samples fall on multiples of `gap` inside the first half of the run
        nonlocal t_sample, done_f
        start = done_f
        F.step_time(k)
        done_f += k
        for s in range(start // gap + 1, done_f // gap + 1):
            if s <= samples and cl.count + npts <= samples * npts:
                F.synchronize()
                ts = time.perf_counter()
                cl.sample(F)
                t_sample += time.perf_counter() - ts
    tf, nuf, nupf, specf, nsf, wall_f = _record(F, n_f, every, fine_advance, say, "fine")
    wall_f -= t_sample
    out["fine"] = dict(steps=n_f, wall_s=wall_f, s_per_time_unit=wall_f / T, nu_series=nuf.tolist())
    say("fine run done", wall_f, "s; samples", cl.count // npts)
    # training on the GPU
    t0 = time.perf_counter()
    cl.fit_normalization()
    n_train = cl.count - npts if cl.count > npts else cl.count
    loss = cl.train(train_steps, batch=batch, lr=lr, seed=seed, first=0, n=n_train)
    train_s = time.perf_counter() - t0
    mse, zero = cl.evaluate(n_train, cl.count - n_train) if cl.count > n_train else (float("nan"), float("nan"))
    out["training"] = dict(points=int(cl.count), train_points=int(n_train), steps=train_steps, batch=batch,
                           wall_s=train_s, loss_first=float(loss[0]), loss_last=float(np.mean(loss[-20:])),
                           heldout_mse=mse, heldout_mse_no_closure=zero)
    say("training done", out["training"])

    def coarse_run(with_closure, k=1):
        Cc.set_state(*c0)
        Cc.reset_history()
        if with_closure:
            cl.set_interval(k)
            cl.enable(True, max_ratio=max_ratio)
            adv = lambda k: cl.step(k, check_every=check_every)  # noqa: E731
        else:
            cl.enable(False)
            adv = Cc.step_time
        # record at the same simulated times as the fine run
        r = _record(Cc, n_c, max(1, int(round(every * fine["dt"] / coarse["dt"]))), adv, say,
                    "coarse+closure" if with_closure else "coarse")
        status = cl.status()
        cl.enable(False)
        return r, status

    first = lambda a: a[:len(a) // 2]  # noqa: E731
    half = lambda a: a[len(a) // 2:]  # noqa: E731
    nu_f, nu_f1 = float(np.mean(half(nuf))), float(np.mean(first(nuf)))

    def spec_err(e, ref):
        nb = min(len(ref), len(e))
        kcut = max(2, nb - 1)
        return float(np.linalg.norm(e[1:kcut] - ref[1:kcut]) / np.linalg.norm(ref[1:kcut]))

    def errors(nu, spec):
        """(first-half, second-half) relative errors of the Nu mean and the mean spectrum."""
        nb = min(spec.shape[1], specf.shape[1])
        e1 = (abs(np.mean(first(nu)) - nu_f1) / nu_f1,
              spec_err(first(spec[:, :nb]).mean(0), first(specf[:, :nb]).mean(0)))
        e2 = (abs(np.mean(half(nu)) - nu_f) / nu_f, spec_err(half(spec[:, :nb]).mean(0), half(specf[:, :nb]).mean(0)))
        return e1, e2

    (tp, nup, nupp, specp, nsp, wall_p), _ = coarse_run(False)
    nu_p = float(np.mean(half(nup)))
    (p1_nu, p1_sp), (p2_nu, p2_sp) = errors(nup, specp)
    out["plain_coarse"] = dict(steps=n_c, wall_s=wall_p, nu=nu_p, nu_rel_err=p2_nu, spectrum_rel_err=p2_sp,
                               first_half=dict(nu_rel_err=p1_nu, spectrum_rel_err=p1_sp),
                               speedup_wall_per_time_unit=wall_f / wall_p if wall_p > 0 else None)
    if variants is None:
        variants = [dict(every=k, gain=1.0, kfrac=1.0)
                    for k in (closure_every if isinstance(closure_every, (list, tuple)) else [closure_every])]
    runs = []
    for v in variants:
        cl.set_damping(v.get("gain", 1.0), v.get("kfrac", 1.0))
        (tc, nuc, nucp, specc, nsc, wall_c), status = coarse_run(True, v.get("every", 1))
        nu_c = float(np.mean(half(nuc)))
        (c1_nu, c1_sp), (c2_nu, c2_sp) = errors(nuc, specc)
        # selection score on the first half: the larger of the two errors relative to plain's
        score = max(c1_nu / max(p1_nu, 1e-12), c1_sp / max(p1_sp, 1e-12)) if np.isfinite(nu_c) else np.inf
        runs.append(dict(network_every=v.get("every", 1), gain=v.get("gain", 1.0), kfrac=v.get("kfrac", 1.0),
                         control=bool(v.get("control", False)),
                         steps=n_c, wall_s=wall_c, nu=nu_c, nu_rel_err=c2_nu, spectrum_rel_err=c2_sp,
                         first_half=dict(nu_rel_err=c1_nu, spectrum_rel_err=c1_sp, score_vs_plain=score),
                         beats_plain=bool(c2_nu < p2_nu and c2_sp < p2_sp),
                         runtime_fallback=status[1], last_correction_ratio=status[2],
                         speedup_wall_per_time_unit=wall_f / wall_c if wall_c > 0 else None))
        say("closure run", runs[-1])
    cl.set_damping(1.0, 1.0)
    eligible = [i for i in range(len(runs)) if not runs[i]["control"]] or list(range(len(runs)))
    best = min(eligible, key=lambda i: runs[i]["first_half"]["score_vs_plain"])
    out["closure_coarse"] = runs[best]
    out["closure_variants"] = runs
    out["selected_variant"] = best
    out["nu_fine"] = nu_f
    nu_c, wall_c, fell_back = runs[best]["nu"], runs[best]["wall_s"], runs[best]["runtime_fallback"]
    out["speedup_wall_per_time_unit"] = wall_f / wall_c if wall_c > 0 else None
    # the closure is kept only when it is within the Nu tolerance of the fine run and better
    # than the plain coarse run on the selection half (score below 1)
    use_closure = ((not fell_back) and np.isfinite(nu_c) and abs(nu_c - nu_f) / nu_f <= tol
                   and runs[best]["first_half"]["score_vs_plain"] < 1.0)
    out["tolerance"] = tol
    out["fallback_to_plain"] = not use_closure
    out["result_nu"] = nu_c if use_closure else nu_p
    cl.close()
    Cc.close()
    F.close()
    return out
