"""Benchmark of the lattice hot path (one leapfrog step = kick + drift + wrap).

Workload (the line's `value`): the precision form's lattice, 1024^3 FP64
(dx = 0.5, dt = 0.05, V = phi^2/2 - 1.2 phi^4/4 + 0.25 phi^6/6 in the
reference's units, phi0 = 1.2 plus seeded vacuum fluctuations), K timed steps
with the fields resident in HBM, timed with HIP events on the stream the step
kernels run on.  Before the timed steps the library's timed choice of the
multi-step kernel (two or three leapfrog steps per pass over HBM, the same
bits as single steps) runs to completion on untimed steps, whatever --warmup
is; the line reports how many steps it took.  `roofline` reports the
dominant kernel against its own algorithmic traffic (Phi and Pi read and
written once per launch) and, beside it, the site-update rate against the
one-step bound (32 B per site-update).

The same FP64 lattice, an FP32 lattice from the same initial condition (FP32
fields, FP64 compensated sums) and an FP64 lattice started one ulp away run
on to --precision-steps steps: the FP32 form's rate over the same K steps,
its ratio to FP64, and per output its energy difference to the FP64 run with
the reference's staggered energy and with the time-synchronised one, field
and spectrum errors, beside the one-ulp control.

Extras: 256^3 and 512^3 FP64 and FP32 rates, the one-step kernel, the
reference's loop at 256^3 with its outputs (averages, spectrum, island
tracking every 200 steps), the batch of 64 realizations, the adjoint, and the
CPU baseline (the reference's own evolve() on the host cores).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu] [--no-extra]
                [--kernel S:Z  (fix S steps per pass with z block Z instead of the timed choice)]
                [--analysis N  (time the analysis calls at N^3 only)]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import ctypes as C

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak
METRIC = "site-updates/s & % HBM roofline, FP32-comp vs FP64 at dE/E<=1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def model(lam=1.0, g=0.2, phi0=1.0):
    """The potential V = phi^2/2 - lam phi^4/4 + g phi^6/6 in the reference's
    units (Lambda = 6 lam, Kappa^2 = Lambda^2 / (120 g), phi_ref = sqrt(Lambda) phi)."""
    L = 6.0 * lam
    return dict(a=0.5, mass_mode=1, m2=1.0, lambda_mode=-1, kappa=float(np.sqrt(L * L / (120.0 * g))),
                phi0=float(np.sqrt(L)) * phi0, Lambda=L, lam=lam, g=g)


def lattice(pn, N, m, precision):
    return pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"], precision=precision)


def launch_count(chunks, per_pass):
    """Launches for steps run as evolve calls of the given sizes with a kernel
    advancing `per_pass` steps per launch: (multi-step launches, steps left to
    shorter launches).  pn_evolve runs the rest of a call by pairs, then singly."""
    return sum(n // per_pass for n in chunks), sum(n % per_pass for n in chunks)


def kernel_used(pn, L):
    if not L.two_step():
        return {"steps_per_pass": 1, "zblock": None}
    if L.three_step():
        zb = C.c_int()
        L._chk(L.L.pn_get_option(L.h, pn.PN_OPT_STEP3_ZBLOCK, C.byref(zb)))
        return {"steps_per_pass": 3, "zblock": zb.value}
    return {"steps_per_pass": 2, "zblock": L.step2_zblock()}


def fix_kernel(L, kernel):
    if kernel[0] == 3:
        L.three_step(True, zblock=kernel[1])
    else:
        L.three_step(False)
        L.step2_zblock(kernel[1])


def tune(lats, dtk, dtd, s, warmup, stop=None, chunk=64):
    """Advances the lattices together, untimed, until each has made its timed
    choice of the multi-step kernel (PN_OPT_KERNEL_TIMED) and at least
    `warmup` steps have run; evolve calls end at multiples of `stop` (the
    output grid) when given.  Returns the step reached."""
    s0 = s
    while any(L.two_step() and not L.kernel_timed() for L in lats) or s - s0 < warmup:
        n = chunk
        if stop:
            n = min(n, stop - s % stop)
        for L in lats:
            L.evolve(dtk[s:s + n], dtd[s:s + n])
        s += n
        if s - s0 > 4000:
            raise RuntimeError("the kernel choice did not finish in 4000 steps")
    return s


def run_leg(pn, N, precision, steps, warmup, seed, m, two_step=True, kernel=None, chunk=None):
    """Seeded initialize==3 IC on the device, the timed kernel choice on
    untimed steps, then `steps` timed steps (in evolve calls of `chunk` steps,
    or one call).  Returns (event ms over the timed steps, the evolve call
    sizes, kernel used, untimed steps before the timing)."""
    tune_chunk = 64
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 4000 + warmup + steps)
    with lattice(pn, N, m, precision) as L:
        if not two_step:
            L.two_step(False)
        elif kernel:
            fix_kernel(L, kernel)
        L.initial_conditions(pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"]))
        s = tune([L], dtk, dtd, 0, warmup, chunk=tune_chunk)
        chunks, total = [], 0.0
        left = steps
        while left:
            n = min(left, chunk or left)
            total += L.evolve_timed(dtk[s:s + n], dtd[s:s + n])
            chunks.append(n)
            s += n
            left -= n
        used = kernel_used(pn, L)
    return total, chunks, used, s - steps


def profiled_value(name, key):
    """A per-launch figure measured with rocprofv3 --pmc and written into
    profiles/<name>.md as `<key> = <number>`: `traffic_bytes_per_launch`
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy), `fp64_flop_per_launch`
    (64 x (2 FMA + ADD + MUL + TRANS) FP64 wave instructions) or
    `valu_busy_percent`.  None when the summary or the line is absent."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    if not os.path.exists(path):
        return None
    m = re.search(r"%s = ([0-9.]+)" % key, open(path).read())
    return float(m.group(1)) if m else None


def kernel_names(t, per_pass, N):
    """(kernel name for the JSON line, profile summary name) of the kernel that
    advances `per_pass` steps per launch for element type t."""
    p = "fp64" if t == "double" else "fp32"
    c = "CoefF64" if t == "double" else "CoefF32"
    if per_pass == 3:
        return ("k_step3<%s, %s, POW2, W, RT, FAST> (three leapfrog steps per pass)" % (t, c),
                "k_step3_%s_%dcube" % (p, N))
    if per_pass == 2:
        return ("k_step2<%s, %s, POW2, W, RT, LP, FAST> (two leapfrog steps per pass)" % (t, c),
                "k_step2_%s_%dcube" % (p, N))
    return "k_step_vec<%s, %s, POW2, W, 4, ZB, 0> (one fused step)" % (t, c), "k_step_vec_%s_%dcube" % (p, N)


def roofline(kernel, N, ms_total, chunks, per_pass, esize, profile=None):
    """Roofline of the dominant kernel: algorithmic bytes per launch (Phi and
    Pi read once and written once: 4 * esize bytes per site) over the average
    launch duration, plus the site-update rate against the one-step bound."""
    nmain, nrest = launch_count(chunks, per_pass)
    steps = sum(chunks)
    sites = float(N) ** 3
    algo = 4.0 * esize * sites
    # the few steps left to shorter launches count pro rata
    per_launch_ms = ms_total / (nmain + nrest / float(per_pass))
    achieved = algo / (per_launch_ms * 1e-3) / 1e9
    upd = sites * steps / (ms_total * 1e-3)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
           "steps_per_launch": per_pass, "algorithmic_bytes_per_launch": int(algo),
           "launches": nmain, "avg_launch_ms": per_launch_ms,
           "one_step_bound": {"bytes_per_site_update": 4 * esize,
                              "site_updates_per_s_at_peak": HBM_PEAK_GBS * 1e9 / (4 * esize),
                              "frac": upd * 4 * esize / 1e9 / HBM_PEAK_GBS}}
    if profile:
        t = profiled_value(profile, "traffic_bytes_per_launch")
        if t:
            out["traffic"] = int(t)
            out["traffic_over_algorithmic"] = t / algo
            out["traffic_source"] = "profiles/%s.md" % profile
        flop = profiled_value(profile, "fp64_flop_per_launch")
        if flop and esize == 8:
            # FP64 FLOP counted by rocprofv3 over the live launch time
            tf = flop / (per_launch_ms * 1e-3) / 1e12
            out["fp64_valu"] = {"flop_per_launch": int(flop), "achieved": tf,
                                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": tf / FP64_PEAK_TFLOPS,
                                "valu_busy_percent": profiled_value(profile, "valu_busy_percent"),
                                "source": "profiles/%s.md" % profile}
    return out


def precision_leg(pn, N, K, warmup, total_steps, out_every, seed=1234, control=True, kernel=None):
    """The headline and the precision form.  N^3, lambda = 1.2, g = 0.25,
    phi0 = 1.2 plus seeded vacuum fluctuations.  An FP64 lattice, an FP32
    lattice from the same initial condition and (control) an FP64 lattice
    with every phi one ulp higher step together: untimed until both timed
    kernel choices are made, then K steps timed on the FP64 and the FP32
    lattice (evolve calls ending on the output grid), then on to total_steps.
    Every out_every steps: energies (the reference's, with pi half a step
    behind phi, and the time-synchronised one, pi kicked by dtau/2), the FP32
    field's max and L2 errors against FP64, the spectrum's relative error
    below QS, and the same for the control."""
    m = model(lam=1.2, g=0.25, phi0=1.2)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], total_steps + 4000 + K + warmup)
    half = 0.5 * 0.1 * m["a"]
    ic = pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"])
    rows = []
    ms = {"fp64": 0.0, "fp32": 0.0}
    chunks = []
    with lattice(pn, N, m, pn.PN_FP64) as L64, lattice(pn, N, m, pn.PN_FP32) as L32:
        Lc = lattice(pn, N, m, pn.PN_FP64) if control else None
        try:
            if kernel:
                fix_kernel(L64, kernel)
            t0 = time.time()
            L64.initial_conditions(ic)
            L32.initial_conditions(ic)
            log("precision %d^3: initial conditions in %.1fs" % (N, time.time() - t0))
            if Lc is not None:
                phi, pi = L64.get_fields()
                Lc.set_fields(np.nextafter(phi, np.inf), pi)
                del phi, pi
            lats = [L for L in (L64, L32, Lc) if L is not None]

            def record(step):
                r = dict(step=step)
                e = {k: L.volume_averages().energy for k, L in (("64", L64), ("32", L32))}
                # at step 0 pi and phi share the initial time: no kick
                h = half if step > 0 else 0.0
                es = {k: L.volume_averages_synced(h).energy for k, L in (("64", L64), ("32", L32))}
                d = L32.field_diff(L64)
                s64, s32 = L64.spectrum(), L32.spectrum()
                p = np.arange(s64["nbins"]) * s64["binsize"]
                sel = (s64["counts"] > 0) & (p < ic.QS)
                r.update(e64=e["64"], e32=e["32"], dE=abs(e["32"] - e["64"]) / abs(e["64"]),
                         e64_sync=es["64"], dE_sync=abs(es["32"] - es["64"]) / abs(es["64"]),
                         max_dphi=d["max_dphi"], max_dphi_rel=d["max_dphi"] / d["max_phi"],
                         l2_dphi=d["l2_dphi"], l2_dpi=d["l2_dpi"],
                         spec=float(np.max(np.abs(s32["F"][sel] - s64["F"][sel]) / s64["F"][sel])))
                if Lc is not None:
                    dc = Lc.field_diff(L64)
                    r.update(ulp_dE=abs(Lc.volume_averages().energy - e["64"]) / abs(e["64"]),
                             ulp_dE_sync=abs(Lc.volume_averages_synced(h).energy - es["64"]) / abs(es["64"]),
                             ulp_l2_dphi=dc["l2_dphi"], ulp_max_dphi=dc["max_dphi"])
                rows.append(r)
                log("precision %d^3 step %d: fp32 dE/E %.3g (synced %.3g), max|dphi| %.3g, L2 %.3g, "
                    "spectrum %.3g" % (N, step, r["dE"], r["dE_sync"], d["max_dphi"], d["l2_dphi"], r["spec"]))

            record(0)
            s = tune(lats, dtk, dtd, 0, warmup, stop=out_every)
            if s % out_every == 0:
                record(s)
            tuned_at = s
            log("precision %d^3: kernel choice done after %d untimed steps" % (N, s))
            timed_end = s + K
            while s < max(total_steps, timed_end):
                n = min(out_every - s % out_every, max(total_steps, timed_end) - s)
                if s < timed_end:
                    n = min(n, timed_end - s)
                    ms["fp64"] += L64.evolve_timed(dtk[s:s + n], dtd[s:s + n])
                    ms["fp32"] += L32.evolve_timed(dtk[s:s + n], dtd[s:s + n])
                    chunks.append(n)
                    if Lc is not None:
                        Lc.evolve(dtk[s:s + n], dtd[s:s + n])
                else:
                    for L in lats:
                        L.evolve(dtk[s:s + n], dtd[s:s + n])
                s += n
                if s % out_every == 0:
                    record(s)
            k64, k32 = kernel_used(pn, L64), kernel_used(pn, L32)
        finally:
            if Lc is not None:
                Lc.close()
    sites = float(N) ** 3
    v64 = sites * K / (ms["fp64"] * 1e-3)
    v32 = sites * K / (ms["fp32"] * 1e-3)

    def within(key):
        bad = [r["step"] for r in rows if r[key] > 1e-6]
        return int(s if not bad else max([r["step"] for r in rows if r["step"] < bad[0]], default=0))

    keys = ["step", "dE", "dE_sync", "max_dphi", "max_dphi_rel", "l2_dphi", "l2_dpi", "spec"]
    if control:
        keys += ["ulp_dE", "ulp_dE_sync", "ulp_l2_dphi", "ulp_max_dphi"]
    e64 = np.array([r["e64"] for r in rows])
    return {
        "fp64": {"ms": ms["fp64"], "chunks": chunks, "kernel": k64, "value": v64},
        "fp32": {"ms": ms["fp32"], "chunks": chunks, "kernel": k32, "value": v32},
        "untimed_steps_before_timing": tuned_at,
        "table": {"N": N, "steps": s, "out_every": out_every, "lambda": 1.2, "g": 0.25, "phi0": 1.2,
                  "seed": seed, "fp64_kernel": k64, "fp32_kernel": k32,
                  "fp64_ms_per_step": ms["fp64"] / K, "fp32_ms_per_step": ms["fp32"] / K,
                  "fp64_site_updates_per_s": v64, "fp32_site_updates_per_s": v32,
                  "ratio_fp32_vs_fp64": v32 / v64,
                  "steps_with_dE_over_E_le_1e-6": within("dE"),
                  "steps_with_synced_dE_over_E_le_1e-6": within("dE_sync"),
                  "energy_drift_fp64": float(np.max(np.abs(e64 - e64[0])) / abs(e64[0])),
                  "per_output": {k: [float("%.3g" % r[k]) for r in rows] for k in keys},
                  "per_output_keys": "dE = |E32 - E64| / E64 with the reference's energy (pi half a step "
                                     "behind phi); dE_sync = the same with pi kicked by dtau/2 to phi's "
                                     "time; max_dphi = max |phi32 - phi64|; max_dphi_rel = max_dphi / "
                                     "max |phi64|; l2_dphi, l2_dpi = relative L2 field errors; spec = max "
                                     "over shells below QS of |F32 - F64| / F64; ulp_* = the same for an "
                                     "FP64 run started with every phi one ulp higher, against the FP64 run"}}


def cfg1_leg(pn, N=256, steps=10000, out_every=200, seed=1234):
    """The reference's time loop at N^3 FP64 (main.cpp:52-96 with the config's
    cadence): every out_every steps the averages and extrema
    (Print_AveragesExtrema), the shell spectrum (Print_fp) and an island
    analysis with tracking (Full_Analysis at 2.57583 sqrt(<phi^2>),
    main.cpp:70), no files.  Steps timed with events, each output on the
    host clock (the calls return their results)."""
    m = model()
    dtk, dtd, taus = pn.time_factors(0.0, 0.1, m["a"], 4000 + steps)
    t_out = {"volume": 0.0, "spectrum": 0.0, "islands": 0.0}
    with lattice(pn, N, m, pn.PN_FP64) as L, pn.Tracker() as T:
        L.initial_conditions(pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"]))
        s0 = tune([L], dtk, dtd, 0, 0)
        s = s0
        ms_steps = 0.0
        n_out = 0
        t_wall = time.perf_counter()
        while s < s0 + steps:
            n = min(out_every, s0 + steps - s)
            ms_steps += L.evolve_timed(dtk[s:s + n], dtd[s:s + n])
            s += n
            t0 = time.perf_counter()
            avg, ext = L.volume_analysis()
            t1 = time.perf_counter()
            L.spectrum()
            t2 = time.perf_counter()
            T.update(L, 0, 2.57583 * np.sqrt(avg.phi2), taus[s - 1])
            t3 = time.perf_counter()
            t_out["volume"] += t1 - t0
            t_out["spectrum"] += t2 - t1
            t_out["islands"] += t3 - t2
            n_out += 1
        wall = time.perf_counter() - t_wall
        tracked = len(T.islands())
        used = kernel_used(pn, L)
    return {"N": N, "steps": steps, "out_every": out_every, "outputs": n_out,
            "ms_per_step_steps_only": ms_steps / steps,
            "ms_per_step_with_outputs": wall * 1e3 / steps,
            "ms_per_output": {k: v * 1e3 / max(n_out, 1) for k, v in t_out.items()},
            "site_updates_per_s_with_outputs": float(N) ** 3 * steps / wall,
            "tracked_islands_at_end": tracked, "kernel": used}


def analysis_only(pn, N, reps=10):
    """Times each analysis call at N^3 FP64 (cfg1 model, 200 steps after a
    seeded initial condition): for profiling the analysis kernels."""
    m = model()
    dtk, dtd, taus = pn.time_factors(0.0, 0.1, m["a"], 200)
    out = {"N": N}
    with lattice(pn, N, m, pn.PN_FP64) as L:
        L.initial_conditions(pn.ic_params(1234, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"]))
        L.evolve(dtk, dtd)
        thr = 2.57583 * np.sqrt(L.volume_averages().phi2)
        calls = {"volume_analysis": lambda: L.volume_analysis(),
                 "volume_averages_synced": lambda: L.volume_averages_synced(0.025),
                 "spectrum": lambda: L.spectrum()}
        if N <= 512:  # catalogs of millions of islands above that
            calls["islands"] = lambda: L.islands(0, thr, as_array=True, view=True)
            calls["islands_with_labels"] = lambda: L.islands(0, thr, labels=True)
        for name, f in calls.items():
            f()
            t0 = time.perf_counter()
            for _ in range(reps):
                f()
            out[name + "_ms"] = (time.perf_counter() - t0) * 1e3 / reps
            log("analysis %d^3 %s: %.3f ms" % (N, name, out[name + "_ms"]))
        if N > 512:
            return out
        # Full_Analysis repeated on the same field: every island is tracked
        # from the second call on (re-grown from its peak, none retired)
        with pn.Tracker() as T:
            T.update(L, 0, thr, 0.0)
            t0 = time.perf_counter()
            for k in range(reps):
                T.update(L, 0, thr, 1.0 + k)
            out["tracker_update_ms"] = (time.perf_counter() - t0) * 1e3 / reps
            out["tracked_islands"] = len(T.islands())
        log("analysis %d^3 tracker update: %.3f ms (%d islands)" % (N, out["tracker_update_ms"],
                                                                  out["tracked_islands"]))
    return out


def batch_leg(pn, steps, out_every=500):
    """64 realizations: 32 x 64^3, 24 x 128^3, 8 x 192^3, seeds 0..63, lambda = 1,
    g drawn uniformly from [0.1, 0.4] with numpy's default_rng(0); FP64, one
    launch per two or three steps for all of them (multi-step tiles; the
    autotuner keeps the faster mode and the planes per tile per size class),
    HIP graphs of 100 or 96 steps.  Every `out_every` steps (outside
    the timing) each realization gets its energy, the false-vacuum check, its
    shell spectrum and its island catalog (energy density above 4x its mean,
    islands.cpp:783-786, 6-connected).  Returns a dict for the JSON line."""
    sizes = [64] * 32 + [128] * 24 + [192] * 8
    g = np.random.default_rng(0).uniform(0.1, 0.4, len(sizes))
    reals = [dict(N=N, a=0.5, mass_mode=1, m2=1.0, lambda_mode=-1,
                  kappa=float(np.sqrt(36.0 / (120.0 * gg)))) for N, gg in zip(sizes, g)]
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, steps)
    with pn.Batch(reals) as B:
        for r in range(len(sizes)):
            B.view(r).initial_conditions(pn.ic_params(r, a=0.5, Lambda=6.0, phi0=float(np.sqrt(6.0))))
        tiles = B.autotune(trial_steps=5)  # planes per tile, and two or three steps per launch
        per_launch = 3 if B.three_step() else (2 if B.two_step() else 1)
        e0 = np.array([B.view(r).volume_averages().energy for r in range(len(sizes))])
        decayed = np.zeros(len(sizes), dtype=np.int32)
        total = 0.0
        t_analysis = 0.0
        n_out = 0
        s = 0
        while s < steps:
            n = min(out_every, steps - s)
            total += B.evolve_timed(dtk[s:s + n], dtd[s:s + n], use_graph=True)
            s += n
            t0 = time.time()
            av = [B.view(r).volume_averages() for r in range(len(sizes))]
            # false_vacuum_decayed (fieldanalysis.cpp:887-893) per realization
            newly = np.array([(a.interaction < 0 or a.interaction != a.interaction) for a in av])
            if np.any(newly & (decayed == 0)):
                decayed |= newly.astype(np.int32)
                B.set_decayed(decayed)
            spectra = [B.view(r).spectrum() for r in range(len(sizes))]
            cats = [B.view(r).islands(1, 4.0 * av[r].energy, as_array=True)[0] for r in range(len(sizes))]
            t_analysis += time.time() - t0
            n_out += 1
        e1 = np.array([a.energy for a in av])
        isl = [len(c) for c in cats]
        vmax = [int(c["nn"].max()) if len(c) else 0 for c in cats]
        pars = [float(sp["F"].sum()) for sp in spectra]
    sites = float(sum(N ** 3 for N in sizes))
    return {"value": sites * steps / (total * 1e-3), "unit": "site-updates/s",
            "realizations": len(sizes), "sites": int(sites), "steps": steps,
            "ms_per_step": total / steps, "steps_per_launch": per_launch, "tiles_rows_zblock": tiles,
            "one_step_bound_frac": 32.0 * sites / (total / steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "decayed": int(decayed.sum()),
            "energy_drift_max": float(np.max(np.abs(e1 - e0) / np.abs(e0))),
            "outputs": n_out, "analysis_s_per_output": t_analysis / max(n_out, 1),
            "final_islands_per_realization": isl, "final_largest_island_sites": vmax,
            "final_spectrum_total_power": [float("%.6g" % v) for v in pars]}


def adjoint_leg(pn, N=128, steps=4000, K=100, lam=1.0, g=0.2, check_steps=1000, eps=1e-6):
    """Reverse-mode gradient of the smooth oscillon fraction in physical units
    (V = phi^2/2 - lam phi^4/4 + g phi^6/6): seeded vacuum IC divided by
    sqrt(Lambda), checkpoints every K steps.  Timed over `steps` steps.  The
    gradient is checked against central differences in lambda, g and one
    random phi0 direction over `check_steps` steps, where the dynamics is not
    yet chaotic enough to swamp a difference quotient; at `steps` steps the
    lambda derivative is compared with differences at eps = 1e-6 ... 1e-8,
    which converge to it as eps shrinks."""
    m = model(lam, g)
    with lattice(pn, N, m, pn.PN_FP64) as L:
        L.initial_conditions(pn.ic_params(7, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"]))
        phi, pi = L.get_fields()
    s = float(np.sqrt(m["Lambda"]))
    phi, pi = phi / s, pi / s
    out = {"N": N, "steps": steps, "checkpoint_every": K}

    def fwd(l_, g_, ph, dtk, dtd):
        return pn.forward_observable(pn.adj_params(N, c2=1.0, c4=-l_, c6=g_), ph, pi, dtk, dtd)

    # timing run over the full horizon
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    p = pn.adj_params(N, c2=1.0, c4=-lam, c6=g, checkpoint_every=K)
    F, gphi, _, gc, res = pn.adjoint(p, phi, pi, dtk, dtd)
    out.update({"F": F, "dF_dlambda": -gc[1], "dF_dg": gc[2], "forward_ms": res.forward_ms,
                "backward_ms": res.backward_ms, "backward_over_forward": res.backward_ms / res.forward_ms,
                "peak_device_bytes": res.peak_bytes})
    # every phi_n resident in HBM (checkpoint_every >= steps): no recompute;
    # the same kernels in the same order, so the same bits
    pr = pn.adj_params(N, c2=1.0, c4=-lam, c6=g, checkpoint_every=steps)
    Fr, gphir, _, gcr, resr = pn.adjoint(pr, phi, pi, dtk, dtd)
    out["resident"] = {"forward_ms": resr.forward_ms, "backward_ms": resr.backward_ms,
                       "backward_over_forward": resr.backward_ms / resr.forward_ms,
                       "peak_device_bytes": resr.peak_bytes,
                       "same_bits_as_checkpointed": bool(Fr == F and np.array_equal(gphir, gphi)
                                                         and np.array_equal(gcr, gc))}
    del gphir
    conv = {}
    for e in (1e-6, 1e-7, 1e-8):
        conv["%g" % e] = (fwd(lam + e, g, phi, dtk, dtd) - fwd(lam - e, g, phi, dtk, dtd)) / (2 * e)
    out["fd_dF_dlambda_vs_eps"] = conv
    out["rel_err_fd_lambda_eps1e-8"] = abs(conv["1e-08"] - (-gc[1])) / abs(gc[1])
    # gradient check over the shorter horizon
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], check_steps)
    F, gphi, _, gc, _ = pn.adjoint(pn.adj_params(N, c2=1.0, c4=-lam, c6=g, checkpoint_every=K),
                                   phi, pi, dtk, dtd)
    dF_dlam, dF_dg = -gc[1], gc[2]
    chk = {"steps": check_steps, "F": F, "dF_dlambda": dF_dlam, "dF_dg": dF_dg}
    # central differences carry an eps^2 truncation term: shown at two eps
    for e in (eps, eps / 10):
        fd_lam = (fwd(lam + e, g, phi, dtk, dtd) - fwd(lam - e, g, phi, dtk, dtd)) / (2 * e)
        fd_g = (fwd(lam, g + e, phi, dtk, dtd) - fwd(lam, g - e, phi, dtk, dtd)) / (2 * e)
        chk["rel_err_fd_lambda_eps%g" % e] = abs(fd_lam - dF_dlam) / abs(dF_dlam)
        chk["rel_err_fd_g_eps%g" % e] = abs(fd_g - dF_dg) / abs(dF_dg)
    v = np.random.default_rng(3).standard_normal(N ** 3)
    fd_v = (fwd(lam, g, phi + eps * v, dtk, dtd) - fwd(lam, g, phi - eps * v, dtk, dtd)) / (2 * eps)
    ad_v = float(np.dot(gphi, v))
    chk["rel_err_fd_phi0_direction_eps%g" % eps] = abs(fd_v - ad_v) / abs(ad_v)
    out["fd_check"] = chk
    return out


def _host_info(threads):
    model_name = "unknown CPU"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model_name = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu": model_name, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def _host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n))


def cpu_baseline_reference(N, m, target_s=12.0):
    """The reference's own evolve() (oracle/_ref/ref_timing, built from the
    reference sources with its flags) over MPI z-slabs on the host's cores,
    one rank per core: a probe of a few steps sizes a run of about target_s.
    None when the build or mpirun is missing or the run fails."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_timing")
    mpirun = "/opt/conda/bin/mpirun"
    if not (os.path.exists(exe) and os.path.exists(mpirun)):
        return None
    ranks = _host_threads()
    while ranks > 1 and N % ranks:
        ranks -= 1
    args = ["-Nt", str(N), "-at", repr(m["a"]), "-M", repr(m["m2"]), "-Mmode", str(m["mass_mode"]),
            "-L", repr(m["Lambda"]), "-Lmode", str(m["lambda_mode"]), "-Kappa", repr(m["kappa"]),
            "-T0", "0", "-dt", "0.1", "-phi0_0", repr(m["phi0"])]

    def run(steps):
        r = subprocess.run([mpirun, "-np", str(ranks), exe, str(steps)] + args, capture_output=True,
                           text=True, timeout=300, cwd="/tmp")
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-400:])
        return float(r.stdout.strip().splitlines()[-1])

    try:
        probe = run(2)
        steps = max(2, int(target_s / max(probe / 2.0, 1e-6)))
        sec = run(steps)
    except (RuntimeError, ValueError, OSError, subprocess.TimeoutExpired) as e:
        log("reference CPU baseline unavailable: %s" % e)
        return None
    return {"value": float(N) ** 3 * steps / sec, "unit": "site-updates/s", "cores": ranks,
            "kind": "reference",
            "sample": "%d^3 FP64 leapfrog of the workload's model, %d timed steps (%.1f s) of the "
                      "reference's own evolve() (dynamics.cpp:196-237, halo exchange comm.cpp:63-78) "
                      "built from its sources with its flags -O3 -msse4.2 -fprefetch-loop-arrays "
                      "-funroll-all-loops -ffast-math, mpirun -np %d z-slabs" % (N, steps, sec, ranks),
            "host": _host_info(ranks)}


def cpu_baseline(N, m, target_s=12.0):
    ref = cpu_baseline_reference(N, m, target_s)
    if ref is not None:
        return ref
    path = os.path.join(ROOT, "oracle", "libcpu_baseline.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.cb_run.restype = C.c_double
    lib.cb_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                           C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    threads = _host_threads()
    ck = C.c_double()
    probe = lib.cb_run(N, threads, 1, 2, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"],
                       m["kappa"], 0.1, m["phi0"], C.byref(ck))
    steps = max(2, int(target_s / max(probe / 2.0, 1e-6)))
    sec = lib.cb_run(N, threads, 2, steps, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"],
                     m["kappa"], 0.1, m["phi0"], C.byref(ck))
    return {"value": float(N) ** 3 * steps / sec, "unit": "site-updates/s", "cores": threads,
            "kind": "port",
            "sample": "%d^3 FP64 leapfrog, %d timed steps (%.1f s) of the reference loop nest "
                      "(AoS + z halo, pos()/mod() per site, kick then drift, z slabs over %d threads), "
                      "reference flags -O3 -ffast-math" % (N, steps, sec, threads),
            "host": _host_info(threads)}


def extra_rate(pn, N, precision, steps, warmup, m, chunk=None, two_step=True):
    ms, chunks, used, untimed = run_leg(pn, N, precision, steps, warmup, 1234, m, two_step=two_step,
                                        chunk=chunk)
    t = "double" if precision == pn.PN_FP64 else "float"
    kname, prof = kernel_names(t, used["steps_per_pass"], N)
    return {"site_updates_per_s": float(N) ** 3 * steps / (ms * 1e-3), "ms_per_step": ms / steps,
            "steps": steps, "kernel": used, "untimed_steps_before_timing": untimed,
            "roofline": roofline(kname, N, ms, chunks, used["steps_per_pass"], 8 if t == "double" else 4,
                                 prof)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed steps of the 1024^3 workload")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps at least (the kernel choice always completes first)")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--precision-steps", type=int, default=10000)
    ap.add_argument("--precision-every", type=int, default=1000)
    ap.add_argument("--no-control", action="store_true", help="no one-ulp FP64 control run")
    ap.add_argument("--cfg1-steps", type=int, default=10000)
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--kernel", default=None,
                    help="S:Z fixes the FP64 multi-step kernel (S = 2 or 3 steps per pass, z block Z) "
                         "instead of the library's timed choice (profiling runs)")
    ap.add_argument("--analysis", type=int, default=0, help="only time the analysis calls at N^3")
    ap.add_argument("--adjoint-only", action="store_true", help="only the 128^3 adjoint leg")
    ap.add_argument("--cfg1-only", action="store_true",
                    help="only the reference's loop with outputs at 256^3 (--cfg1-steps steps)")
    ap.add_argument("--fp32-only", action="store_true",
                    help="only the FP32 form at --size for --steps steps (for profiling)")
    ap.add_argument("--single-step", action="store_true", help="time the one-step FP64 kernel only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="the workload and its precision form only")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB (to scale rocprof byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": a.numel() * 8,
                          "copy_calibration_bytes_written": b.numel() * 8, "dispatches": 10}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the lattice fits one MI355X; "
              "multi-GPU z-slab sharding is a design, not measured here)", file=sys.stderr)
        sys.exit(2)

    import phinum_amd as pn

    N, K, W = args.size, args.steps, args.warmup
    kern = tuple(int(v) for v in args.kernel.split(":")) if args.kernel else None
    if args.analysis:
        print(json.dumps(analysis_only(pn, args.analysis)), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(pn)), flush=True)
        return
    if args.cfg1_only:
        print(json.dumps(cfg1_leg(pn, 256, args.cfg1_steps)), flush=True)
        return
    m = model(lam=1.2, g=0.25, phi0=1.2)
    if args.fp32_only or args.single_step:
        prec = pn.PN_FP32 if args.fp32_only else pn.PN_FP64
        print(json.dumps(extra_rate(pn, N, prec, K, W, m, two_step=not args.single_step)), flush=True)
        return
    log("fp64/fp32 %d^3: kernel choice, %d timed steps, on to %d steps" % (N, K, args.precision_steps))
    t0 = time.time()
    P = precision_leg(pn, N, K, W, args.precision_steps, args.precision_every,
                      control=not args.no_control, kernel=kern)
    log("precision leg done in %.1fs wall" % (time.time() - t0))
    f64, f32 = P["fp64"], P["fp32"]
    value = f64["value"]
    kname, prof = kernel_names("double", f64["kernel"]["steps_per_pass"], N)
    kname32, prof32 = kernel_names("float", f32["kernel"]["steps_per_pass"], N)
    tab = P["table"]
    line = {
        "metric": METRIC, "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": f64["ms"] / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded initialize==3 vacuum field (seed 1234, phi0=1.2 in physical units, "
                "modes below k_max/2), generated on the device",
        "config": {"workload": "fp64 leapfrog %d^3 (the precision form's lattice), dx=0.5, dt=0.05, "
                               "lambda=1.2, g=0.25, %d timed steps" % (N, K),
                   "N": N, "kernel": f64["kernel"],
                   "untimed_steps_before_timing": P["untimed_steps_before_timing"],
                   "kernel_choice_complete_before_timing": True},
        "roofline": roofline(kname, N, f64["ms"], f64["chunks"], f64["kernel"]["steps_per_pass"], 8, prof),
        "fp32_comp": {
            "value": f32["value"], "unit": "site-updates/s", "ms_per_step": f32["ms"] / K,
            "ratio_vs_fp64": f32["value"] / value, "kernel": f32["kernel"],
            "steps_with_dE_over_E_le_1e-6": tab["steps_with_dE_over_E_le_1e-6"],
            "steps_with_synced_dE_over_E_le_1e-6": tab["steps_with_synced_dE_over_E_le_1e-6"],
            "roofline": roofline(kname32, N, f32["ms"], f32["chunks"], f32["kernel"]["steps_per_pass"], 4,
                                 prof32)},
        "precision_%d" % N: tab,
    }
    if not args.no_extra:
        mc = model()
        for n5 in (256, 512):
            log("fp64 and fp32 %d^3" % n5)
            line["fp64_%d" % n5] = extra_rate(pn, n5, pn.PN_FP64, 2000, 10, mc)
            line["fp32_%d" % n5] = extra_rate(pn, n5, pn.PN_FP32, 2000, 10, mc)
            line["fp32_%d" % n5]["ratio_vs_fp64"] = (line["fp32_%d" % n5]["site_updates_per_s"] /
                                                    line["fp64_%d" % n5]["site_updates_per_s"])
        log("fp64 256^3 one-step kernel")
        line["single_step_kernel_256"] = extra_rate(pn, 256, pn.PN_FP64, 1000, 10, mc, two_step=False)
        log("the reference's loop with outputs, 256^3")
        line["cfg1_loop_256"] = cfg1_leg(pn, 256, args.cfg1_steps)
        log("batch of 64 realizations")
        line["batch_64"] = batch_leg(pn, args.batch_steps)
        log("adjoint 128^3")
        line["adjoint_128"] = adjoint_leg(pn)
    if not args.no_cpu:
        log("cpu baseline")
        cb = cpu_baseline(N, m)
        if cb:
            line["cpu_baseline"] = cb
            line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
