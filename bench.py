"""Benchmark of the lattice hot path (one leapfrog step = kick + drift + wrap).

Workload: the 256^3 FP64 drop-in configuration (dx = 0.5, dt = 0.05,
V = phi^2/2 - phi^4/4 + 0.2 phi^6/6 in the reference's units, 10^4 steps with
energy outputs every 200 steps).  `value` is FP64 lattice site-updates per
second with the fields resident in HBM, timed with HIP events on the stream
the step kernels run on.  The steps run as the multi-step (temporally
blocked) kernel the library picks by timing during the warmup (two or three
steps per pass over HBM, bit-identical to single steps); `roofline` reports
that kernel against its own algorithmic traffic (Phi and Pi read and written
once per launch) and, beside it, the site-update rate against the one-step
bound (32 B per site-update).  The line also carries the single-step kernel, the
FP32 form (FP32 fields, FP64 compensated energy sums) with its energy
difference to the FP64 run, the 512^3 case, the batched ensemble, the
adjoint, and the CPU baseline.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu] [--no-extra]
                [--kernel S:Z  (fix S steps per pass with z block Z instead of the timed choice)]
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
                phi0=float(np.sqrt(L)) * phi0, Lambda=L)


def launches(steps, chunk, per_pass):
    """Launch counts for `steps` run in chunks of `chunk` with the kernel
    advancing `per_pass` (1, 2 or 3) steps per launch: (main launches, steps
    left to shorter launches).  pn_evolve runs the rest of a chunk by pairs,
    then singly."""
    full, rest = divmod(steps, chunk)
    sizes = [chunk] * full + ([rest] if rest else [])
    return sum(n // per_pass for n in sizes), sum(n % per_pass for n in sizes)


def run_leg(pn, N, precision, steps, warmup, out_every, seed, m, two_step=True, kernel=None):
    """Seeded initialize==3 IC on the device, then `steps` timed steps in chunks
    of `out_every` with an FP64 energy between chunks (outside the timing).
    The multi-step kernel is the library's own choice (timed during a warmup of
    at least 118 steps) unless `kernel` = (steps per pass, z block) fixes it.
    Returns (total event ms over `steps`, energies at each output, the kernel
    used as {"steps_per_pass", "zblock"})."""
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], warmup + steps)
    energies = []
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"],
                    precision=precision) as L:
        if not two_step:
            L.two_step(False)
        elif kernel:
            if kernel[0] == 3:
                L.three_step(True, zblock=kernel[1])
            else:
                L.three_step(False)
                L.step2_zblock(kernel[1])
        L.initial_conditions(pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"]))
        if warmup:
            L.evolve(dtk[:warmup], dtd[:warmup])
        energies.append(L.volume_averages().energy)
        total = 0.0
        s = warmup
        while s < warmup + steps:
            n = min(out_every, warmup + steps - s)
            total += L.evolve_timed(dtk[s:s + n], dtd[s:s + n])
            s += n
            energies.append(L.volume_averages().energy)
        if not L.two_step():
            used = {"steps_per_pass": 1, "zblock": None}
        elif L.three_step():
            zb = C.c_int()
            L._chk(L.L.pn_get_option(L.h, pn.PN_OPT_STEP3_ZBLOCK, C.byref(zb)))
            used = {"steps_per_pass": 3, "zblock": zb.value}
        else:
            used = {"steps_per_pass": 2, "zblock": L.step2_zblock()}
    return total, np.array(energies), used


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


def profiled_traffic(name):
    t = profiled_value(name, "traffic_bytes_per_launch")
    return int(t) if t else None


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


def roofline(kernel, N, ms_total, steps, chunk, per_pass, esize, profile=None):
    """Roofline of the dominant kernel: algorithmic bytes per launch (Phi and
    Pi read once and written once: 4 * esize bytes per site) over the average
    launch duration, plus the site-update rate against the one-step bound."""
    nmain, nrest = launches(steps, chunk, per_pass)
    sites = float(N) ** 3
    algo = 4.0 * esize * sites
    # the few steps of a chunk left to shorter launches count pro rata
    per_launch_ms = ms_total / (nmain + nrest / float(per_pass))
    steps_per_launch = per_pass
    achieved = algo / (per_launch_ms * 1e-3) / 1e9
    upd = sites * steps / (ms_total * 1e-3)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
           "steps_per_launch": steps_per_launch, "algorithmic_bytes_per_launch": int(algo),
           "avg_launch_ms": per_launch_ms,
           "one_step_bound": {"bytes_per_site_update": 4 * esize,
                              "site_updates_per_s_at_peak": HBM_PEAK_GBS * 1e9 / (4 * esize),
                              "frac": upd * 4 * esize / 1e9 / HBM_PEAK_GBS}}
    if profile:
        t = profiled_traffic(profile)
        if t:
            out["traffic"] = t
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


def precision_leg(pn, N, steps, out_every, seed=1234):
    """Precision form: N^3, lambda = 1.2, g = 0.25, phi0 = 1.2 plus seeded
    vacuum fluctuations, `steps` steps.  Phi and Pi stored and stepped in FP32
    (energy and spectrum sums in FP64, compensated) against the same run in
    FP64 from the same initial condition.  Per output: relative energy
    difference, max |dphi| (absolute and over max |phi|), relative L2 field
    errors, and the spectrum's relative error (max over the shells below the
    initial cutoff QS, where the initial condition put power, of
    |F32 - F64| / F64).  A third run, FP64 from the FP64 initial condition with
    every phi moved up by one ulp, measures how fast this chaotic system
    separates trajectories that start 1e-16 apart."""
    m = model(lam=1.2, g=0.25, phi0=1.2)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    ic = pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=m["phi0"])
    rows = []
    ms = {"fp64": 0.0, "fp32": 0.0}

    def lattice(prec):
        return pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"], precision=prec)

    with lattice(pn.PN_FP64) as L64, lattice(pn.PN_FP32) as L32, lattice(pn.PN_FP64) as Lc:
        t0 = time.time()
        L64.initial_conditions(ic)
        L32.initial_conditions(ic)
        log("precision %d^3: initial conditions in %.1fs" % (N, time.time() - t0))
        phi, pi = L64.get_fields()
        Lc.set_fields(np.nextafter(phi, np.inf), pi)
        del phi, pi

        def record(step):
            e64, e32, ec = (L.volume_averages().energy for L in (L64, L32, Lc))
            d = L32.field_diff(L64)
            dc = Lc.field_diff(L64)
            s64, s32 = L64.spectrum(), L32.spectrum()
            p = np.arange(s64["nbins"]) * s64["binsize"]
            sel = (s64["counts"] > 0) & (p < ic.QS)
            spec = float(np.max(np.abs(s32["F"][sel] - s64["F"][sel]) / s64["F"][sel]))
            rows.append(dict(step=step, e64=e64, e32=e32, dE=abs(e32 - e64) / abs(e64),
                             max_dphi=d["max_dphi"], max_dphi_rel=d["max_dphi"] / d["max_phi"],
                             l2_dphi=d["l2_dphi"], l2_dpi=d["l2_dpi"], spec=spec,
                             ulp_dE=abs(ec - e64) / abs(e64), ulp_l2_dphi=dc["l2_dphi"],
                             ulp_max_dphi=dc["max_dphi"]))
            log("precision %d^3 step %d: fp32 dE/E %.3g, max|dphi| %.3g, L2 %.3g, spectrum %.3g; "
                "fp64+1ulp dE/E %.3g, L2 %.3g" % (N, step, rows[-1]["dE"], d["max_dphi"], d["l2_dphi"], spec,
                                                   rows[-1]["ulp_dE"], dc["l2_dphi"]))

        record(0)
        s = 0
        while s < steps:
            n = min(out_every, steps - s)
            ms["fp64"] += L64.evolve_timed(dtk[s:s + n], dtd[s:s + n])
            ms["fp32"] += L32.evolve_timed(dtk[s:s + n], dtd[s:s + n])
            Lc.evolve(dtk[s:s + n], dtd[s:s + n])
            s += n
            record(s)
        on64, on32 = ("three-step" if L.three_step() else "two-step" if L.two_step() else "one-step"
                      for L in (L64, L32))
    dE = np.array([r["dE"] for r in rows])
    over = np.nonzero(dE > 1e-6)[0]
    e64 = np.array([r["e64"] for r in rows])
    e32 = np.array([r["e32"] for r in rows])
    v64 = float(N) ** 3 * steps / (ms["fp64"] * 1e-3)
    v32 = float(N) ** 3 * steps / (ms["fp32"] * 1e-3)
    return {"N": N, "steps": steps, "out_every": out_every, "lambda": 1.2, "g": 0.25, "phi0": 1.2,
            "seed": seed, "fp64_kernel": on64, "fp32_kernel": on32,
            "fp64_ms_per_step": ms["fp64"] / steps, "fp32_ms_per_step": ms["fp32"] / steps,
            "fp64_site_updates_per_s": v64, "fp32_site_updates_per_s": v32, "ratio_fp32_vs_fp64": v32 / v64,
            "steps_with_dE_over_E_le_1e-6": int(steps if over.size == 0 else rows[over[0] - 1]["step"]),
            "energy_drift_fp64": float(np.max(np.abs(e64 - e64[0])) / abs(e64[0])),
            "energy_drift_fp32": float(np.max(np.abs(e32 - e32[0])) / abs(e32[0])),
            "per_output": {k: [float("%.3g" % r[k]) for r in rows]
                           for k in ("step", "dE", "max_dphi", "max_dphi_rel", "l2_dphi", "l2_dpi", "spec",
                                     "ulp_dE", "ulp_l2_dphi", "ulp_max_dphi")},
            "per_output_keys": "dE = |E32 - E64| / E64; max_dphi = max |phi32 - phi64|; max_dphi_rel = "
                               "max_dphi / max |phi64|; l2_dphi, l2_dpi = relative L2 field errors; spec = "
                               "max over shells below QS of |F32 - F64| / F64; ulp_* = the same for an FP64 "
                               "run started with every phi one ulp higher, against the FP64 run"}


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
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"]) as L:
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
        probe = run(4)
        steps = max(4, int(target_s / max(probe / 4.0, 1e-6)))
        sec = run(steps)
    except (RuntimeError, ValueError, OSError, subprocess.TimeoutExpired) as e:
        log("reference CPU baseline unavailable: %s" % e)
        return None
    return {"value": float(N) ** 3 * steps / sec, "unit": "site-updates/s", "cores": ranks,
            "kind": "reference",
            "sample": "%d^3 FP64 leapfrog, %d timed steps (%.1f s) of the reference's own evolve() "
                      "(dynamics.cpp:196-237, halo exchange comm.cpp:63-78) built from its sources with "
                      "its flags -O3 -msse4.2 -fprefetch-loop-arrays -funroll-all-loops -ffast-math, "
                      "mpirun -np %d z-slabs" % (N, steps, sec, ranks),
            "host": _host_info(ranks)}


def cpu_baseline(N, m, target_s=12.0):
    ref = cpu_baseline_reference(N, m, target_s)
    if ref is not None:
        return ref
    import ctypes as C
    path = os.path.join(ROOT, "oracle", "libcpu_baseline.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.cb_run.restype = C.c_double
    lib.cb_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                           C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    threads = _host_threads()
    ck = C.c_double()
    probe = lib.cb_run(N, threads, 1, 3, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"],
                       m["kappa"], 0.1, m["phi0"], C.byref(ck))
    steps = max(3, int(target_s / max(probe / 3.0, 1e-6)))
    sec = lib.cb_run(N, threads, 2, steps, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"],
                     m["kappa"], 0.1, m["phi0"], C.byref(ck))
    return {"value": float(N) ** 3 * steps / sec, "unit": "site-updates/s", "cores": threads,
            "kind": "port",
            "sample": "%d^3 FP64 leapfrog, %d timed steps (%.1f s) of the reference loop nest "
                      "(AoS + z halo, pos()/mod() per site, kick then drift, z slabs over %d threads), "
                      "reference flags -O3 -ffast-math" % (N, steps, sec, threads),
            "host": _host_info(threads)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=150)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--out-every", type=int, default=200)
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--precision-size", type=int, default=1024)
    ap.add_argument("--precision-steps", type=int, default=10000)
    ap.add_argument("--precision-every", type=int, default=1000)
    ap.add_argument("--fp32-only", action="store_true",
                    help="only the FP32 form at --size for --steps steps (for profiling)")
    ap.add_argument("--precision-only", action="store_true",
                    help="only the precision form (FP32 against FP64 at --precision-size)")
    ap.add_argument("--single-step", action="store_true", help="time the one-step kernel only")
    ap.add_argument("--kernel", default=None,
                    help="S:Z fixes the multi-step kernel (S = 2 or 3 steps per pass, z block Z) "
                         "instead of the library's timed choice (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="FP64 workload only")
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

    if args.precision_only:
        print(json.dumps(precision_leg(pn, args.precision_size, args.precision_steps,
                                       args.precision_every)), flush=True)
        return
    N, K, W = args.size, args.steps, args.warmup
    two = not args.single_step
    kern = tuple(int(v) for v in args.kernel.split(":")) if args.kernel else None
    m = model()
    if args.fp32_only:
        ms32, _, k32 = run_leg(pn, N, pn.PN_FP32, K, W, args.out_every, 1234, m, kernel=kern)
        print(json.dumps({"fp32_ms_per_step": ms32 / K, "kernel": k32}), flush=True)
        return
    log("fp64 %d^3: %d warmup + %d timed steps" % (N, W, K))
    t0 = time.time()
    ms64, e64, k64 = run_leg(pn, N, pn.PN_FP64, K, W, args.out_every, 1234, m, two_step=two, kernel=kern)
    log("fp64 done in %.1fs wall, %.4f ms/step, kernel %s" % (time.time() - t0, ms64 / K, k64))
    sites = float(N) ** 3
    value = sites * K / (ms64 * 1e-3)
    kname, prof = kernel_names("double", k64["steps_per_pass"], N)
    line = {
        "metric": METRIC, "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": ms64 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded initialize==3 vacuum field (seed 1234, phi0=1 in physical units, "
                "modes below k_max/2), generated on the device",
        "config": {"workload": "fp64 leapfrog %d^3, dx=0.5, dt=0.05, lambda=1, g=0.2, %d steps, "
                               "energy every %d steps" % (N, K, args.out_every),
                   "N": N, "out_every": args.out_every, "kernel": k64},
        "roofline": roofline(kname, N, ms64, K, args.out_every, k64["steps_per_pass"], 8, prof),
        "energy_drift_fp64": float(np.max(np.abs(e64 - e64[0])) / abs(e64[0])),
    }
    if not args.no_extra:
        log("fp64 %d^3 single-step kernel" % N)
        K1 = min(K, 2000)
        ms1, _, _ = run_leg(pn, N, pn.PN_FP64, K1, 10, args.out_every, 1234, m, two_step=False)
        line["single_step_kernel"] = {
            "value": sites * K1 / (ms1 * 1e-3), "unit": "site-updates/s", "steps": K1,
            "ms_per_step": ms1 / K1,
            "roofline": roofline("k_step_vec<double, CoefF64, true, 2, 4, 32, 0>", N, ms1, K1,
                                 args.out_every, 1, 8, "k_step_vec_fp64_%dcube" % N)}
        log("fp32 %d^3" % N)
        ms32, e32, k32 = run_leg(pn, N, pn.PN_FP32, K, W, args.out_every, 1234, m)
        kname32, prof32 = kernel_names("float", k32["steps_per_pass"], N)
        v32 = sites * K / (ms32 * 1e-3)
        rel = np.abs(e32 - e64) / np.abs(e64)
        over = np.nonzero(rel > 1e-6)[0]
        line["fp32_comp"] = {
            "value": v32, "unit": "site-updates/s", "ms_per_step": ms32 / K,
            "ratio_vs_fp64": v32 / value,
            "dE_over_E_vs_fp64_max": float(np.max(rel)),
            # outputs are every out_every steps after the warmup; the two
            # chaotic trajectories separate roughly exponentially
            "steps_with_dE_over_E_le_1e-6": int(K if over.size == 0 else (over[0] - 1) * args.out_every),
            "dE_over_E_vs_fp64_by_output": [float("%.3g" % r) for r in rel[::max(1, len(rel) // 10)]],
            "energy_drift_fp32": float(np.max(np.abs(e32 - e32[0])) / abs(e32[0])),
            "kernel": k32,
            "roofline": roofline(kname32, N, ms32, K, args.out_every, k32["steps_per_pass"], 4, prof32),
        }
        N5 = 512
        log("fp64 %d^3" % N5)
        ms5, _, k5 = run_leg(pn, N5, pn.PN_FP64, 200, 150, 200, 5, m)
        kname5, prof5 = kernel_names("double", k5["steps_per_pass"], N5)
        line["fp64_512"] = {"ms_per_step": ms5 / 200, "kernel": k5,
                            "site_updates_per_s": float(N5) ** 3 * 200 / (ms5 * 1e-3),
                            "roofline": roofline(kname5, N5, ms5, 200, 200, k5["steps_per_pass"], 8, prof5)}
        log("precision form %d^3" % args.precision_size)
        line["precision_%d" % args.precision_size] = precision_leg(
            pn, args.precision_size, args.precision_steps, args.precision_every)
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
