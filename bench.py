"""Throughput of the Rayleigh-Benard IMEX step on one MI355X.

Workload (one "step" = one IMEX time step of the whole field): 512 x 512 horizontal grid,
32 vertical Galerkin modes per field, Ra = 1e6, Pr = 0.7, box 4 x 4 x 1, dt = 5e-5, from a
seeded theta ~ 1e-6 N(0,1) in the lowest 4 x 4 x 2 modes, the other fields zero (synthetic input,
resident in HBM before timing).  Prints one JSON line; its roofline object also carries the
rates of a plain 1 GiB fill and copy on the same GPU (--no-hbm-probe skips them).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f32|bf16x3|f64]
                  [--no-cpu] [--cpu-threads T]
  python bench.py --ensemble [--members 64] [--steps K] [--ens-k 10]   (batched workload)
  python bench.py --closure [--closure-T 1.0] [--precision f32|bf16x3]   (learned closure)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "timesteps/sec at 512²×32 (+HBM/MFMA roofline %); closure vs plain Nu err & wall time"
HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(M=512, N=512, L=32, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5)


def stage_bytes(st, cs, nprod):
    """Algorithmic HBM bytes per step of each stage (cs = bytes per complex value).
    P = kept horizontal modes, H = N * KX hybrid row entries, nq levels, L modes, nprod product
    fields between the row and forward y stages (h, d, uv, uw, vw, u.grad(theta))."""
    P = st.kx_kept * st.ky_kept
    H = WORKLOAD["N"] * st.kx_kept
    nq, L = st.nq, WORKLOAD["L"]
    return {
        "eval": (3 * L * P + 6 * nq * P) * cs,
        "fft_y_inv": (6 * nq * P + 6 * nq * H) * cs,
        "rows_products": (6 * nq * H + nprod * nq * H) * cs,
        # forward y-FFT reads the product rows and writes the 5 level combinations
        "fft_y_fwd": (nprod * nq * H + 5 * nq * P) * cs,
        "project": (5 * nq * P + 3 * L * P) * cs,
        # state, N and history in/out; factor records once per |k|^2 group
        "update": (9 * L * P + 6 * L * P) * cs + st.n_unique_k * (2 * L * L + L * L // 2) * (cs // 2),
        # make_real of the kx = 0 column and the F, G update
        "finish": 2 * 3 * st.ky_kept * L * cs + 8 * L * cs,
    }


FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (half of it without packed FP32)


def stage_flops(st):
    """Algorithmic FP32 operations per step of the x-row stage: per (level, y) row 6 inverse and
    6 forward real 512-point transforms (2.5 M log2 M each) and the products and combinations
    (uu, vv, ww, uv, uw, vw, u.grad(theta), h = uu + vv - 2 ww, d = uu - vv: 15 per point)."""
    M, N = WORKLOAD["M"], WORKLOAD["N"]
    lg = M.bit_length() - 1
    rows = st.nq * N
    return {"rows_products": rows * (12 * 2.5 * M * lg + 15 * M)}


def fft_product_stage(st, stages, cs):
    """The FFT + product stage (inverse y-FFT, x rows with the products, forward y-FFT with the
    level combinations) against HBM: algorithmic bytes (F_in + F_out) Nq K cs with K = (M/2+1) N
    (the fused formula), for this formulation (F_in = 6 level fields, F_out = 5 level
    combinations, Nq = the level count) and for the canonical one (F_in = 4, F_out = 9, Nq = 48)."""
    K = (WORKLOAD["M"] // 2 + 1) * WORKLOAD["N"]
    names = ("fft_y_inv", "rows_products", "fft_y_fwd")
    ms = sum(stages[k] for k in names)
    out = {"stages": list(names), "stage_ms": round(ms, 4), "K": K, "bytes_per_complex": cs}
    for tag, fi, fo, nq in (("this", 6, 5, st.nq), ("canonical", 4, 9, 48)):
        b = (fi + fo) * nq * K * cs
        gbs = b / (ms * 1e-3) / 1e9
        out[tag] = {"F_in": fi, "F_out": fo, "Nq": nq, "bytes": b, "achieved_GBs": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["frac"] = out["this"]["frac"]
    return out


# kernels of each stage on the 512 x 512 FP32 path (BF16X3 runs the <1> variants of the two
# matrix-core kernels, with the same memory traffic)
STAGE_KERNELS = {"eval": ["k_eval_mma<0, false>"], "fft_y_inv": ["k_fft_y_inv512"], "rows_products": ["k_rows512t"],
                 "fft_y_fwd": ["k_fft_y_fwd512c2<8>"],
                 "project": ["k_project_mma<0, 4, 1>"], "update": ["k_update32<float, 32>"],
                 "finish": ["k_finish<float>"]}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_512x512x32_f32.md")


def measured_traffic(stage):
    """HBM bytes per step of a stage from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (calibrated against a device copy); None when the stage's kernels are not in it."""
    try:
        txt = open(TRAFFIC_PROFILE).read()
        table = json.loads(txt.split("<!-- traffic-bytes ", 1)[1].split(" -->", 1)[0])
        return sum(table[k] for k in STAGE_KERNELS[stage])
    except (OSError, IndexError, KeyError, ValueError):
        return None


ISSUE_PROFILE = os.path.join(ROOT, "profiles", "sq_issue_512x512x32_f32.md")


def measured_valu_issue(stage):
    """VALU instructions issued per CU per clock (at most 1: packed FP32 is off) of a stage's
    kernels, from the committed rocprofv3 SQ-counter summary; None when they are not in it."""
    try:
        txt = open(ISSUE_PROFILE).read()
        table = json.loads(txt.split("<!-- valu-issue ", 1)[1].split(" -->", 1)[0])
        return {k: table[k] for k in STAGE_KERNELS[stage]}
    except (OSError, IndexError, KeyError, ValueError):
        return None


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB by the HIP runtime (its blit kernel reads and writes
    exactly that many bytes), used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), n) == 0 and hip.hipMalloc(ctypes.byref(dst), n) == 0
    assert hip.hipMemset(src, 1, n) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)


def hbm_probe(mib=1024, reps=20):
    """What this GPU's HBM does for plain streams: a fill (hipMemsetD32Async, writes only) and a
    device-to-device copy (reads and writes) of `mib` MiB each, timed with HIP events over `reps`
    calls; GB/s of bytes moved.  Context for the per-kernel rates beside the published 8 TB/s."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemsetD32Async.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
    hip.hipMemcpyDtoDAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    n = mib << 20
    src, dst, e0, e1 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), n) == 0 and hip.hipMalloc(ctypes.byref(dst), n) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0

    def timed(call):
        assert call() == 0 and hip.hipDeviceSynchronize() == 0
        assert hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            assert call() == 0
        assert hip.hipEventRecord(e1, None) == 0 and hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return ms.value / reps

    fill = timed(lambda: hip.hipMemsetD32Async(src, 0, n // 4, None))
    copy = timed(lambda: hip.hipMemcpyDtoDAsync(dst, src, n, None))
    for p in (src, dst):
        hip.hipFree(p)
    for e in (e0, e1):
        hip.hipEventDestroy(e)
    return {"MiB": mib, "fill_GBs": round(n / fill / 1e6, 1), "copy_GBs": round(2 * n / copy / 1e6, 1),
            "method": "hipMemsetD32Async and hipMemcpyDtoDAsync, HIP events, %d calls each" % reps}


def host_cpu_info():
    """nproc, the usable CPUs of this process, the CPU model and the socket count (from
    /proc/cpuinfo, as lscpu reports them)."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        model, sockets = None, set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model is None:
                model = line.split(":", 1)[1].strip()
            elif line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
        info["model"] = model
        info["sockets"] = len(sockets) or None
    except OSError:
        pass
    return info


def _c3_sampled(threads, budget_s):
    """One 512^2 x 32 oracle step with an evenly strided sample of the nonlinear term's (i, j)
    pair loop, extrapolated to all L^2 pairs (the full step takes minutes)."""
    from oracle import oracle as O
    from sdc_amd.initial import low_mode_theta
    W = WORKLOAD
    S = O.Stepper(W["M"], W["N"], W["L"], W["Lx"], W["Ly"], W["Ra"], W["Pr"], W["dt"], nthreads=threads)
    S.set_state(*low_mode_theta(W["M"], W["N"], W["L"], seed=3, amp=1e-6))
    # calibrate with a coarse sample, then pick the stride that fits the budget
    L2 = W["L"] ** 2
    probe = O.time_step_sampled(S, L2 // max(1, threads))
    per_pair = probe["pairs_sampled"] / max(1, threads)
    fixed = probe["phys"] + probe["rest"]
    n_pairs = max(threads, int((budget_s - fixed) / max(per_pair, 1e-9)))
    stride = max(1, L2 // n_pairs)
    r = O.time_step_sampled(S, stride)
    S.close()
    return r, stride


def _oracle_steps(cfg, state, steps, threads):
    """Wall seconds per step of the oracle on a configuration (setup excluded)."""
    from oracle import oracle as O
    S = O.Stepper(*cfg, nthreads=threads)
    S.set_state(*state)
    t0 = time.perf_counter()
    S.step(steps)
    dt = time.perf_counter() - t0
    S.close()
    return dt / steps


def cpu_baseline(threads, budget_s=15.0, c1_steps=100, c2_steps=2):
    """The oracle (FP64 C restatement of the reference algorithm, OpenMP) on this host, on all
    `threads` cores and on one thread: the headline workload as a sampled step, and the two
    smaller configurations (32^2 x 8 Ra 5000, 128^2 x 16 Ra 1e5) as whole steps."""
    from oracle import oracle as O
    from sdc_amd.initial import low_mode_theta, random_theta
    O.build()
    W = WORKLOAD
    r, stride = _c3_sampled(threads, budget_s)
    r1, stride1 = _c3_sampled(1, budget_s)
    c1 = (32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3)
    c2 = (128, 128, 16, 4.0, 4.0, 1e5, 0.7, 2e-4)
    s1 = low_mode_theta(32, 32, 8, seed=1, amp=1e-3)
    s2 = random_theta(128, 128, 16, seed=2, amp=1e-3)
    small = {}
    for name, cfg, st, n in (("32x32x8", c1, s1, c1_steps), ("128x128x16", c2, s2, c2_steps)):
        small[name] = {"steps_timed": n,
                       "timesteps_per_s": {f"{t}_threads": round(1.0 / _oracle_steps(cfg, st, n, t), 3)
                                           for t in sorted({threads, 1})}}
    return {
        "value": round(1.0 / r["s_per_step"], 6),
        "unit": "timesteps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one {W['M']}x{W['N']}x{W['L']} step of the FP64 oracle (OpenMP, {threads} threads) with every "
                   f"{stride}-th of the {W['L'] ** 2} (i,j) pairs of the nonlinear term, extrapolated to all pairs "
                   f"(sampled, not a whole step); phases s: physical fields {r['phys']:.2f}, sampled pairs "
                   f"{r['pairs_sampled']:.2f}, rest {r['rest']:.2f}"),
        "single_thread": {"value": round(1.0 / r1["s_per_step"], 6), "unit": "timesteps/s", "cores": 1,
                          "sample": f"as above on one thread, every {stride1}-th pair"},
        "whole_steps_other_configs": small,
        "host": host_cpu_info(),
        "threads_note": "threads = the GPU box's CPU share for one GPU (16), capped by the usable CPUs",
    }


def ensemble_configs(n=64, seed=4):
    """The batched workload: n runs with the grid drawn from {64^2, 128^2, 256^2}, L from
    {8, 16, 24}, Ra log-uniform in [1e4, 1e6], Pr = 0.7, box 4 x 4.  dt = 0.25 dx / U with the
    free-fall velocity U = (Ra Pr)^1/2 (in thermal-diffusion units) keeps the explicit AB2
    advection stable through the onset transient of every member."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cfgs = []
    for i in range(n):
        M = int(rng.choice([64, 128, 256]))
        L = int(rng.choice([8, 16, 24]))
        Ra = float(10 ** rng.uniform(4, 6))
        dt = 0.25 * (4.0 / M) / (Ra * 0.7) ** 0.5
        cfgs.append(dict(M=M, N=M, L=L, Lx=4.0, Ly=4.0, Ra=Ra, Pr=0.7, dt=dt, precision="f32"))
    return cfgs


def run_ensemble(a):
    """Member-steps per second of the ensemble workload: graph replay with the size classes
    batched (one launch per stage and class), graph replay with one branch per member, and
    eager launches."""
    import numpy as np
    from sdc_amd import Ensemble
    from sdc_amd.initial import random_theta
    cfgs = ensemble_configs(a.members)
    t0 = time.perf_counter()
    E = Ensemble(cfgs)
    t_create = time.perf_counter() - t0
    t0 = time.perf_counter()
    rep = E.tune()
    t_tune = time.perf_counter() - t0
    for i, (m, c) in enumerate(zip(E.members, cfgs)):
        m.set_state(*random_theta(c["M"], c["N"], c["L"], seed=1000 + i, amp=1e-3))
    E.step(1 + a.warmup, steps_per_launch=a.ens_k)  # Euler start, capture, warm-up
    batched = E.batched_members
    res = {}
    for mode, k in (("graph", a.ens_k), ("graph_unbatched", a.ens_k), ("eager", 0)):
        if mode == "graph_unbatched":
            E.set_batching(False)
            E.step(a.ens_k, steps_per_launch=a.ens_k)  # capture outside the timed region
        t0 = time.perf_counter()
        E.step(a.steps, steps_per_launch=k)
        res[mode] = time.perf_counter() - t0
    nu = np.array([m.nusselt()[2] for m in E.members])
    finite = int(np.isfinite(nu).sum())
    E.close()
    # the same draw with every grid at 64^2: members too small to fill the device, where the
    # batched classes pay off
    small = [dict(c, M=64, N=64, dt=0.25 * (4.0 / 64) / (c["Ra"] * 0.7) ** 0.5) for c in cfgs]
    small_rate = {}
    for mode, on in (("batched", True), ("unbatched", False)):
        S = Ensemble(small)
        S.set_batching(on)
        for i, (m, c) in enumerate(zip(S.members, small)):
            m.set_state(*random_theta(c["M"], c["N"], c["L"], seed=1000 + i, amp=1e-3))
        S.step(1 + 2 * a.ens_k, steps_per_launch=a.ens_k)
        t0 = time.perf_counter()
        S.step(a.steps, steps_per_launch=a.ens_k)
        small_rate[mode] = round(len(small) * a.steps / (time.perf_counter() - t0), 1)
        if on:
            small_rate["batched_members"] = S.batched_members
        S.close()
    n = len(cfgs)
    classes = sorted({(c["M"], c["L"]) for c in cfgs})
    out = {
        "metric": "ensemble member-steps/s (64 uneven RBC runs, size classes batched in one captured graph per launch)",
        "value": round(n * a.steps / res["graph"], 1),
        "unit": "member-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * res["graph"] / a.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded theta ~ 1e-3 N(0,1) per member, seed 1000 + member)",
        "config": {"workload": f"ensemble of {n} runs, grid in {{64^2,128^2,256^2}}, L in {{8,16,24}}, "
                               "Ra log-uniform [1e4,1e6], Pr=0.7, 4x4x1 box",
                   "steps_per_launch": a.ens_k, "size_classes": [list(c) for c in classes]},
        "batched_members": batched,
        "unbatched_graph_member_steps_per_s": round(n * a.steps / res["graph_unbatched"], 1),
        "eager_member_steps_per_s": round(n * a.steps / res["eager"], 1),
        "batched_speedup_vs_unbatched_graph": round(res["graph_unbatched"] / res["graph"], 3),
        "graph_speedup_vs_eager": round(res["eager"] / res["graph"], 3),
        "create_s": round(t_create, 2), "tune_s": round(t_tune, 2),
        "tuned_shapes": {f"{c['M']}^2xL{c['L']}": [int(r[0]), int(r[1]), round(float(r[2]), 4)]
                         for c, r in zip(cfgs, rep)},
        "small_grid_ensemble": {"workload": "the same 64 runs, every grid at 64^2 (L, Ra as drawn)",
                                "member_steps_per_s": small_rate},
        "members_finite": finite,
        "nu_vol_range": [round(float(np.nanmin(nu)), 6), round(float(np.nanmax(nu)), 6)],
    }
    print(json.dumps(out))


# The explicit AB2 advection limits the step: from a random theta at Ra = 1e6 the onset plumes
# reach |u| ~ 500, and u dt k_max must stay well below 1 (k_max = 2 pi 171 / 4 at 512^2, 43
# columns at 128^2).  Each grid runs at the largest step of a doubling search that stays stable
# over the whole time unit: 512^2 x 32 at 5e-6 and 128^2 x 16 at 2e-5 survive 0.2 time units
# from the developed state but blow up near t = 0.45 and t = 0.35; a factor 2 below both
# (the same CFL number on both grids) is used.
CLOSURE_FINE = dict(M=512, N=512, L=32, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=2.5e-6)
CLOSURE_COARSE = dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=1e-5)


def closure_variants(a):
    """Corrected runs from --closure-variants "every:gain:kfrac,..." (default: the network on every
    step and every 2nd step undamped, then damped and low-pass-filtered corrections), plus the
    controls of --closure-controls (same syntax; reported, never selected)."""
    out = []
    for spec, control in ((a.closure_variants, False), (a.closure_controls, True)):
        for item in filter(None, spec.split(",")):
            e, g, k = (item.split(":") + ["1", "1"])[:3]
            out.append(dict(every=int(e), gain=float(g), kfrac=float(k), control=control))
    return out


SMALL_FINE = dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=2.5e-5)
SMALL_COARSE = [dict(M=32, N=32, L=8, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=1e-4),
                dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=5e-5)]


def run_closure_small(a):
    """The closure experiment at a small size where the coarse grids are far from resolved:
    plain 128^2 x 16 at Ra = 1e5 (judge and data source) against 32^2 x 8 and 64^2 x 8 runs with
    and without the closure over one time unit; one JSON line per coarse grid."""
    from sdc_amd import closure
    for coarse in SMALL_COARSE:
        r = closure.run(SMALL_FINE, coarse, T=a.closure_T, spinup_steps=8000, samples=a.closure_samples,
                        train_steps=a.closure_train_steps, batch=65536, lr=1e-3, every=400,
                        fine_precision=a.precision, variants=closure_variants(a),
                        log=lambda *m: print("#", *m, file=sys.stderr, flush=True))
        print(json.dumps({"metric": "closure vs plain Nu err & spectrum err (small grids, Ra=1e5)",
                          "config": {"fine": SMALL_FINE, "coarse": coarse, "T": a.closure_T},
                          "nu_fine": r["nu_fine"], "plain_coarse": r["plain_coarse"],
                          "closure_coarse": r["closure_coarse"], "closure_variants": r["closure_variants"],
                          "selected_variant": r["selected_variant"], "fallback_to_plain": r["fallback_to_plain"],
                          "training": r["training"],
                          "fine": {k: v for k, v in r["fine"].items() if k != "nu_series"}}), flush=True)


def run_closure(a):
    """The learned-closure experiment: plain 512^2 x 32 run (judge and data source) vs a
    128^2 x 16 run with and without the trained closure over one time unit."""
    from sdc_amd import closure
    r = closure.run(CLOSURE_FINE, CLOSURE_COARSE, T=a.closure_T, spinup_steps=a.closure_spinup,
                    samples=a.closure_samples, train_steps=a.closure_train_steps, batch=65536, lr=1e-3, every=2000,
                    fine_precision=a.precision, variants=closure_variants(a),
                    log=lambda *m: print("#", *m, file=sys.stderr, flush=True))
    cc = r["closure_coarse"]
    out = {
        "metric": "closure vs plain Nu err & wall time (128^2x16 with closure vs plain 512^2x32, Ra=1e6)",
        "value": round(cc["nu_rel_err"], 5), "unit": "relative Nu error", "higher_is_better": False,
        "n_gpus": 1, "steps": r["fine"]["steps"], "warmup": a.closure_spinup,
        "ms_per_step": round(1e3 * cc["wall_s"] / cc["steps"], 4), "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 weights+f32", "data": "synthetic (fine run from seeded theta ~ 1e-3 N(0,1) in every kept mode)",
        "config": {"workload": f"RBC Ra=1e6 Pr=0.7 4x4x1, {a.closure_T} time unit(s) after "
                               f"{a.closure_spinup} fine spin-up steps",
                   "fine": CLOSURE_FINE, "coarse": CLOSURE_COARSE, "fine_precision": a.precision},
        "speedup_wall_per_time_unit": r["speedup_wall_per_time_unit"],
        "nu_fine": r["nu_fine"], "plain_coarse": r["plain_coarse"], "closure_coarse": cc,
        "closure_variants": r["closure_variants"],
        "training": r["training"], "fallback_to_plain": r["fallback_to_plain"], "tolerance": r["tolerance"],
        "fine": {k: v for k, v in r["fine"].items() if k != "nu_series"},
    }
    print(json.dumps(out))


def _max_rel(a, b, keep):
    """Largest relative difference over theta, f, g (kept modes) and F, G (each against its own
    maximum)."""
    errs = []
    for i in range(5):
        x, y = (a[i][:, keep], b[i][:, keep]) if i < 3 else (a[i], b[i])
        scale = np.abs(y).max()
        errs.append(float(np.abs(x - y).max() / scale) if scale > 0 else float(np.abs(x).max()))
    return max(errs)


def tracer_flops(P, L, npart):
    """Matrix-core FP32 operations of one spectral velocity evaluation (k_vel_points_mma) at npart
    particles over P kept modes: per (particle, mode) six length-L vertical sums (A = sum f C',
    Cc = sum f C, B = sum g S, real and imaginary parts), 12 L operations."""
    return npart * P * 12 * L


def run_tracers(a):
    """Tracer throughput at 128^2 x 16 (Ra 1e5, Pr 0.7, box 4 x 4) with 1e5 particles: the
    spectral midpoint RK2 (two exact velocity evaluations per step) on the FP32 state (the
    reported value) and on a BF16X3 copy of it, and the reference's RK4 with trilinear
    interpolation on the 128-level velocity grid."""
    import ctypes
    import sdc_amd
    from sdc_amd.initial import random_theta
    hip = ctypes.CDLL("libamdhip64.so")
    cfg = (128, 128, 16, 4.0, 4.0, 1e5, 0.7, 2e-4)
    npart, steps = 100_000, a.steps
    S = sdc_amd.Stepper(*cfg, precision="f32")
    S.set_state(*random_theta(128, 128, 16, seed=2, amp=1e-3))
    S.step_time(20)
    out = {"metric": "tracer steps/s (1e5 tracers, 128^2x16 RBC state)", "unit": "tracer-steps/s", "n_gpus": 1,
           "steps": steps, "warmup": a.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded theta ~ 1e-3 N(0,1), 20 steps)",
           "config": {"workload": "1e5 passive tracers in a 128x128 grid x 16 mode state, Ra=1e5, Pr=0.7, 4x4x1",
                      "particles": npart}}
    modes = {}
    # the BF16X3 state: the spectral evaluation's vertical sums as BF16 hi/lo split products
    Sb = sdc_amd.Stepper(*cfg, precision="bf16x3")
    Sb.set_state(*S.get_state())
    Tb = sdc_amd.Tracer(Sb, npart, buffer_size=steps + a.warmup + 2, mode="spectral_rk2", seed=2)
    Tb.init_positions("random")
    for _ in range(a.warmup):
        Tb.step_particles(1e-3)
    Sb.synchronize()
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Tb.step_particles(1e-3)
    Sb.synchronize()
    hip.hipDeviceSynchronize()
    dt_b = time.perf_counter() - t0
    Tb.close()
    Sb.close()
    modes["spectral_rk2_bf16x3"] = {"tracer_steps_per_s": round(steps / dt_b, 2), "ms_per_step": round(1e3 * dt_b / steps, 4),
                                    "particle_steps_per_s": round(npart * steps / dt_b, 1)}
    for mode in ("spectral_rk2", "rk4_trilinear"):
        T = sdc_amd.Tracer(S, npart, buffer_size=steps + a.warmup + 2, mode=mode, seed=2)
        T.init_positions("random")
        for _ in range(a.warmup):
            T.step_particles(1e-3)
        S.synchronize()
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            T.step_particles(1e-3)
        S.synchronize()
        hip.hipDeviceSynchronize()
        dt = time.perf_counter() - t0
        T.close()
        modes[mode] = {"tracer_steps_per_s": round(steps / dt, 2), "ms_per_step": round(1e3 * dt / steps, 4),
                       "particle_steps_per_s": round(npart * steps / dt, 1)}
    P = S.kx_kept * S.ky_kept
    fl = 2 * tracer_flops(P, 16, npart)
    tf = fl / (modes["spectral_rk2"]["ms_per_step"] * 1e-3) / 1e12
    out["value"] = modes["spectral_rk2"]["tracer_steps_per_s"]
    out["ms_per_step"] = modes["spectral_rk2"]["ms_per_step"]
    out["modes"] = modes
    out["roofline"] = {"bound": "mfma", "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(tf / FP32_PEAK_TFLOPS, 4),
                       "kernel": "k_vel_points_mma (two per spectral RK2 step)",
                       "algorithmic_flops_per_step": fl, "kept_modes": P,
                       "note": "FP32 matrix-core flops of the vertical sums over the whole tracer step time (both "
                               "evaluations, the slice sums and the position updates); the phases and the velocity "
                               "combination run on the VALU, which does not overlap the matrix cores here"}
    S.close()
    print(json.dumps(out))


def run_validate(a):
    """The accuracy forms, each against an FP64 solver on the same seeded input:
    C1 (32^2 x 8, Ra 5000) FP64 on the GPU against the FP64 CPU oracle, 1000 steps;
    C2 (128^2 x 16, Ra 1e5) FP32 against FP64 on the GPU (the FP64 path is pinned to the oracle),
       2000 steps, 1e5 spectral-RK2 tracers riding the FP32 run (at dt = 2.5e-5: the stated 2e-4 is
       beyond the explicit advection's limit, which the run records);
    C3 (512^2 x 32, Ra 1e6) BF16X3 and FP32 against FP64 on the GPU from a developed FP64 state."""
    import sdc_amd
    from oracle import oracle as O
    from sdc_amd.initial import kept_mask, low_mode_theta, random_theta
    say = lambda *m: print("#", *m, file=sys.stderr, flush=True)  # noqa: E731
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    O.build()
    out = {"metric": "validation: relative error of fields, Nu and spectra against FP64 runs"}

    # C1
    c1 = (32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3)
    st = low_mode_theta(32, 32, 8, seed=1, amp=1e-3)
    keep = kept_mask(32, 32)
    G = sdc_amd.Stepper(*c1, precision="f64")
    G.set_state(*st)
    o = O.Stepper(*c1, nthreads=threads)
    o.set_state(*st)
    nu_g, nu_o, err100 = [], [], None
    for k in range(10):
        G.step_time(100)
        o.step(100)
        if k == 0:
            err100 = _max_rel(G.get_state(), o.get_state(), keep)
        nu_g.append(G.nusselt().tolist())
        nu_o.append(o.nusselt().tolist())
    err1000 = _max_rel(G.get_state(), o.get_state(), keep)
    G.close()
    o.close()
    nu_g, nu_o = np.array(nu_g), np.array(nu_o)
    out["C1"] = {"config": "32x32x8 Ra=5000 Pr=1 box 4.032^2 dt=1e-3, low-mode theta seed 1, FP64 GPU vs FP64 oracle",
                 "field_rel_err_100_steps": err100, "field_rel_err_1000_steps": err1000,
                 "nu_rel_diff_max": float(np.abs(nu_g - nu_o).max() / np.abs(nu_o).max()),
                 "nu_volume_every_100_steps": nu_g[:, 2].tolist()}
    say("C1", out["C1"])

    # C2: first the stated dt = 2e-4 (FP64), recording when it stops being finite; then the
    # comparison at a stable step
    st = random_theta(128, 128, 16, seed=2, amp=1e-3)
    keep = kept_mask(128, 128)
    T = sdc_amd.Stepper(128, 128, 16, 4.0, 4.0, 1e5, 0.7, 2e-4, precision="f64")
    T.set_state(*st)
    blowup = None
    from sdc_amd import RbcNumericError
    for k in range(20):
        T.step_time(100)
        try:
            T.synchronize()
        except RbcNumericError as e:  # the end-of-step health word names the first non-finite step
            blowup = 100 * (k + 1)
            out["C2_stated_dt_health"] = str(e)
            break
    T.close()
    c2 = (128, 128, 16, 4.0, 4.0, 1e5, 0.7, 2.5e-5)
    A = sdc_amd.Stepper(*c2, precision="f32")
    B = sdc_amd.Stepper(*c2, precision="f64")
    A.set_state(*st)
    B.set_state(*st)
    tr = sdc_amd.Tracer(A, 100000, buffer_size=101, mode="spectral_rk2", seed=2)
    tr.init_positions("random")
    nu_a, nu_b = [], []
    t0 = time.perf_counter()
    for k in range(20):
        for _ in range(100):
            A.step_time(1)
            tr.step_particles(c2[7])
        tr.get_particle_positions(clear=True)  # the ring holds 100 steps (the reference's output cadence)
        B.step_time(100)
        nu_a.append(A.nusselt().tolist())
        nu_b.append(B.nusselt().tolist())
    wall = time.perf_counter() - t0
    Ea, _ = A.energy_spectrum()
    Eb, _ = B.energy_spectrum()
    n = min(len(Ea), len(Eb))
    nu_a, nu_b = np.array(nu_a), np.array(nu_b)
    mom = tr.moments()
    out["C2"] = {"config": "128x128x16 Ra=1e5 Pr=0.7 box 4^2, random theta 1e-3 seed 2, 2000 steps at dt=2.5e-5, FP32 "
                           "vs FP64 on the GPU, 1e5 spectral-RK2 tracers",
                 "stated_dt_2e-4_fp64_nonfinite_by_step": blowup,
                 "field_rel_err_2000_steps": _max_rel(A.get_state(), B.get_state(), keep),
                 "nu_rel_diff_max": float(np.abs(nu_a - nu_b).max() / np.abs(nu_b).max()),
                 "spectrum_rel_err": float(np.linalg.norm(Ea[:n] - Eb[:n]) / np.linalg.norm(Eb[:n])),
                 "nu_volume_every_100_steps_f32": nu_a[:, 2].tolist(),
                 "tracer_moments": [float(x) for x in mom], "wall_s_f32_with_tracers_plus_f64": round(wall, 2)}
    tr.close()
    A.close()
    B.close()
    say("C2", {k: v for k, v in out["C2"].items() if "every" not in k})

    # C3
    c3 = (512, 512, 32, 4.0, 4.0, 1e6, 0.7, 2.5e-6)
    keep = kept_mask(512, 512)
    F = sdc_amd.Stepper(*c3, precision="f64")
    F.set_state(*random_theta(512, 512, 32, seed=3, amp=1e-3))
    t0 = time.perf_counter()
    F.step_time(2000)
    F.synchronize()
    dev = F.get_state()
    say("C3 developed state", round(time.perf_counter() - t0, 1), "s")
    res = {}
    ref_nu, ref_E = None, None
    F.reset_history()
    F.step_time(200)
    ref = F.get_state()
    ref_nu = F.nusselt()
    ref_E, _ = F.energy_spectrum()
    F.close()
    for prec in ("bf16x3", "f32"):
        S = sdc_amd.Stepper(*c3, precision=prec)
        S.set_state(*dev)
        S.step_time(200)
        E, _ = S.energy_spectrum()
        n = min(len(E), len(ref_E))
        res[prec] = {"field_rel_err": _max_rel(S.get_state(), ref, keep),
                     "nu_rel_err": float(np.abs(S.nusselt() - ref_nu).max() / np.abs(ref_nu).max()),
                     "spectrum_rel_err": float(np.linalg.norm(E[:n] - ref_E[:n]) / np.linalg.norm(ref_E[:n]))}
        S.close()
    out["C3"] = {"config": "512x512x32 Ra=1e6 Pr=0.7 box 4^2; FP64 2000 steps from random theta 1e-3 seed 3, then 200 "
                           "steps in each form; dt = 2.5e-6 (5e-5 is unstable through the onset transient)",
                 "nu_fp64": ref_nu.tolist(), **res}
    say("C3", out["C3"])
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200; 500 with --ensemble)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="f32", choices=["f32", "bf16x3", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grid", type=int, nargs=3, metavar=("M", "N", "L"), default=None,
                    help="another grid for profiling (the reported workload is 512 512 32)")
    ap.add_argument("--no-variants", action="store_true", help="skip timing the other precision form")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-torch", action="store_true", help="bracket the timed region with the library's "
                    "stream synchronisation only, no device-wide wait (for profiler runs)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MB",
                    help="also run one device-to-device copy of MB mebibytes (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-hbm-probe", action="store_true",
                    help="skip timing a 1 GiB fill and copy (the measured HBM rates beside the published peak)")
    ap.add_argument("--ensemble", action="store_true", help="run the batched-ensemble workload instead")
    ap.add_argument("--closure", action="store_true", help="run the learned-closure experiment instead")
    ap.add_argument("--closure-small", action="store_true",
                    help="run the closure experiment on small grids (128^2 x 16 fine, 32^2 x 8 and 64^2 x 8 coarse)")
    ap.add_argument("--validate", action="store_true", help="run the accuracy forms (C1, C2, C3) instead")
    ap.add_argument("--tracers", action="store_true", help="run the tracer workload (1e5 tracers at 128^2 x 16)")
    ap.add_argument("--closure-T", type=float, default=1.0)
    ap.add_argument("--closure-variants", default="1:1:1,2:1:1,1:0.5:1,1:0.5:0.5,1:0.25:0.5,2:0.5:0.33",
                    help="corrected runs as every:gain:kfrac (network interval, gain of the correction, "
                         "fraction of k_max below which it is added); the first half of the time unit "
                         "selects the reported one")
    ap.add_argument("--closure-controls", default="1:0.001:1",
                    help="control runs (every:gain:kfrac): a tiny gain makes the plain coarse run under a small "
                         "perturbation, which shows the trajectory-to-trajectory spread of the error measures")
    ap.add_argument("--closure-spinup", type=int, default=8000)
    ap.add_argument("--closure-samples", type=int, default=10)
    ap.add_argument("--closure-train-steps", type=int, default=2000)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--ens-k", type=int, default=10, help="steps captured per graph launch")
    a = ap.parse_args()
    if a.steps is None:
        a.steps = 500 if a.ensemble else (50 if a.tracers else 200)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU layout of this "
              "path is described in DESIGN.md and not measured here.", file=sys.stderr)
        sys.exit(2)

    if a.ensemble:
        run_ensemble(a)
        return
    if a.closure_small:
        run_closure_small(a)
        return
    if a.closure:
        run_closure(a)
        return
    if a.validate:
        run_validate(a)
        return
    if a.tracers:
        run_tracers(a)
        return

    import sdc_amd
    from sdc_amd.initial import low_mode_theta
    if a.grid:
        WORKLOAD.update(M=a.grid[0], N=a.grid[1], L=a.grid[2])
    W = WORKLOAD
    st = sdc_amd.Stepper(W["M"], W["N"], W["L"], W["Lx"], W["Ly"], W["Ra"], W["Pr"], W["dt"],
                         precision=a.precision, device=0)
    st.set_state(*low_mode_theta(W["M"], W["N"], W["L"], seed=3, amp=1e-6))
    # device-wide wait around the timed region: hipDeviceSynchronize of the HIP runtime the
    # library loaded (torch is not imported: its bundled HIP runtime does not see the GPU of the
    # test machines, and a second runtime in the process breaks hipMalloc)
    tsync = lambda: None  # noqa: E731
    if not a.no_torch:
        try:
            import ctypes
            _hip = ctypes.CDLL("libamdhip64.so")
            tsync = lambda: _hip.hipDeviceSynchronize()  # noqa: E731
        except OSError:
            pass
    seed_state = low_mode_theta(W["M"], W["N"], W["L"], seed=3, amp=1e-6)

    def reseed(s):
        """Fresh seeded state and an Euler start for each timed phase.  The seed is theta at
        1e-6 (the amplitude of the reference's default random start) in the lowest 4 x 4 x 2
        modes.  At Ra = 1e6 the convective instability grows about e-fold per 31 steps of
        dt = 5e-5, and the saturated flow is beyond the explicit advection's stable range at
        512^2 for this dt; from 1e-6 the timed steps stay in the linear regime and finite.  The
        cost of a step does not depend on the data."""
        s.set_state(*seed_state)
        s.reset_history()

    def finite_after(s):
        from sdc_amd import RbcNumericError
        try:
            s.synchronize()
            return True
        except RbcNumericError:
            return False

    # timed region: plain rbc_step, no events inside, bracketed by stream + device waits
    st.step_time(a.warmup)
    finite_after(st)
    reseed(st)
    tsync()
    t0 = time.perf_counter()
    st.step_time(a.steps)
    finite = finite_after(st)
    tsync()
    t1 = time.perf_counter()
    sec = t1 - t0
    value = a.steps / sec
    # a separate pass with HIP events around every stage on the library's stream
    reseed(st)
    nprof = min(a.steps, 100)
    stages = st.profile_steps(nprof)
    finite_prof = finite_after(st)

    cs = 16 if a.precision == "f64" else 8
    nprod = 6
    nbytes = stage_bytes(st, cs, nprod)
    dom = max(stages, key=lambda k: stages[k])
    achieved = nbytes[dom] / (stages[dom] * 1e-3) / 1e9
    full = a.precision != "f64" and not a.grid
    traffic = measured_traffic(dom) if full else None
    issue = measured_valu_issue(dom) if full else None
    hbm = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
           "algorithmic_bytes": nbytes[dom]}
    # the dominant kernel's bound from its measured VALU issue rate (rocprofv3 SQ counters): at
    # >= 0.8 wave64 VALU instructions per CU-clock it is the vector ALU that limits it
    valu_bound = issue is not None and max(issue.values()) >= 0.8
    if valu_bound:
        fl = stage_flops(st)[dom]
        tf = fl / (stages[dom] * 1e-3) / 1e12
        roofline = {"bound": "valu", "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(tf / FP32_PEAK_TFLOPS, 4), "algorithmic_flops": fl, "hbm": hbm}
    else:
        roofline = dict(hbm, bound="hbm")
    roofline.update({"traffic": traffic, "kernel": dom, "valu_issue_per_cu_cycle": issue,
                     "stage_ms": {k: round(v, 4) for k, v in stages.items()},
                     "stage_GBs": {k: round(nbytes[k] / (stages[k] * 1e-3) / 1e9, 1)
                                   for k in stages if stages[k] > 0},
                     "fft_product_stage": fft_product_stage(st, stages, cs),
                     "stage_events_steps": nprof})
    nq, kx, ky, ng = st.nq, st.kx_kept, st.ky_kept, st.n_unique_k
    # the boundary's host buffers: one state in (reference layout, FP64 host arrays), one step, one
    # state out, including PCIe and the host-side packing; reported beside `value`, which excludes it
    reseed(st)
    state = st.get_state()
    st.synchronize()
    tb = time.perf_counter()
    st.set_state(*state)
    st.step_time(1)
    state = st.get_state()
    boundary_ms = 1e3 * (time.perf_counter() - tb)
    # the same exchange on device buffers in the reference layout (rbc_set_state_device /
    # rbc_get_state_device): state in and out without a step, mean of 20.  Device memory comes
    # from hipMalloc of the library's HIP runtime.
    dev_boundary_us = None
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        ct = np.complex128 if a.precision == "f64" else np.complex64
        ptrs = []
        for x in state:
            h = np.ascontiguousarray(x.astype(ct))
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), h.nbytes) == 0
            assert hip.hipMemcpy(p, h.ctypes.data_as(ctypes.c_void_p), h.nbytes, 1) == 0
            ptrs.append(p.value)
        st.set_state_device(*ptrs)
        st.get_state_device(*ptrs)
        st.synchronize()
        tb = time.perf_counter()
        for _ in range(20):
            st.set_state_device(*ptrs)
            st.get_state_device(*ptrs)
        st.synchronize()
        dev_boundary_us = 1e6 * (time.perf_counter() - tb) / 20
        for p in ptrs:
            hip.hipFree(ctypes.c_void_p(p))
    except OSError:
        pass
    st.close()
    # the other single-GPU precision forms on the same workload (plain step_time, no stage events)
    variants = {}
    if not a.no_variants:
        for prec in ("f32", "bf16x3"):
            if prec == a.precision:
                continue
            sv = sdc_amd.Stepper(W["M"], W["N"], W["L"], W["Lx"], W["Ly"], W["Ra"], W["Pr"], W["dt"],
                                 precision=prec, device=0)
            reseed(sv)
            sv.step_time(a.warmup)
            finite_after(sv)
            reseed(sv)
            tsync()
            tv = time.perf_counter()
            sv.step_time(a.steps)
            fin = finite_after(sv)
            tsync()
            tv = time.perf_counter() - tv
            variants[prec] = {"timesteps_per_s": round(a.steps / tv, 2), "ms_per_step": round(1e3 * tv / a.steps, 4),
                              "state_finite": fin}
            sv.close()
    if a.copy_calibrate > 0:
        copy_probe(a.copy_calibrate)
    if not a.no_hbm_probe:
        roofline["hbm_measured"] = hbm_probe()

    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        cpu = cpu_baseline(threads)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * sec / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "bf16x3": "bf16x3+f32", "f64": "f64"}[a.precision],
        "data": "synthetic (seeded theta ~ 1e-6 N(0,1) in the lowest 4x4x2 modes, other fields zero)",
        "config": {"workload": f"RBC IMEX step, {W['M']}x{W['N']} grid x {W['L']} vertical modes, Ra=1e6, Pr=0.7, "
                               "4x4x1 box, dt=5e-5", "M": W["M"], "N": W["N"], "L": W["L"], "nq_levels": nq,
                   "kept_modes": [kx, ky], "unique_k2_groups": ng, "precision": a.precision,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "precision_variants": variants,
        "boundary_roundtrip_ms": round(boundary_ms, 2),
        "device_boundary_roundtrip_us": round(dev_boundary_us, 1) if dev_boundary_us is not None else None,
        "state_finite_after_timed_steps": finite,
        "state_finite_after_profiled_steps": finite_prof,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
