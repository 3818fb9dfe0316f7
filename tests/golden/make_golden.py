"""Writes the golden fixtures in this directory from the reference's own
hot-path sources, compiled unmodified with strict IEEE flags
(oracle/_ref/ref_harness, built by oracle/Makefile where the reference tree
exists).  The fixtures are data only: seeded inputs and the reference's
outputs for them.

    python tests/golden/make_golden.py          (everything)
    python tests/golden/make_golden.py info     (the parameter summary only)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402


def _fields(N, seed, phi0, amp):
    rng = np.random.default_rng(seed)
    return phi0 + amp * rng.standard_normal(N ** 3), 0.2 * rng.standard_normal(N ** 3)


CASES = [
    # name, model kwargs, N, steps, seed, phi0, amp, islands, thr, T0, dtbase
    ("cfg0_model_n8_s40", dict(lam=1.0, g=0.2), 8, 40, 1, 2.449489742783178, 0.4, True, -1.0, 0.0, 0.1),
    ("cfg2_model_n10_s25", dict(lam=1.2, g=0.25), 10, 25, 2, 3.219937887599697, 0.5, True, -1.0, 0.0, 0.1),
    ("islands_n12_s0", dict(lam=1.0, g=0.2), 12, 0, 3, 0.0, 1.0, True, 2.2, 0.0, 0.1),
    ("t0_offset_n6_s60", dict(lam=1.0, g=0.3), 6, 60, 4, 1.5, 0.3, False, -1.0, 3.7, 0.07),
]


# island tracking: Full_Analysis at the start and after each block of steps
TRACK_CASES = [
    # name, g, N, seed, amp, ball radius, analyses, steps between, thr (<0: 2.57583 sqrt(<phi^2>))
    ("track_n16_noise", 0.2, 16, 5, 0.9, 0, 6, 5, -1.0),
    ("track_n20_noise", 0.2, 20, 7, 1.0, 0, 10, 5, -1.0),
    ("track_n24_bubble", 0.1, 24, 8, 0.5, 6, 10, 100, -1.0),
    ("track_n16_oscillon", 0.2, 16, 11, 0.9, 0, 50, 20, 0.35 * 6.0 ** 0.5),
]


def track_inputs(N, seed, amp, ball):
    """Noise around phi = 0 (reference units), optionally a ball at the
    true-vacuum value 2.979 sqrt(6) of the g = 0.1 potential."""
    rng = np.random.default_rng(seed)
    phi = amp * rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    if ball:
        z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        r2 = ((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2).ravel()
        phi = np.where(r2 < ball ** 2, 2.979 * np.sqrt(6.0), phi)
    return phi, pi


def pack_track(track):
    """tracked list -> fixed arrays (hist padded with NaN)."""
    n = len(track)
    hist = np.full((n, 5, 8), np.nan)
    for i, (_, _, _, h) in enumerate(track):
        for j, e in enumerate(h):
            hist[i, j] = e
    return dict(t_index=np.array([t[0] for t in track], dtype=np.int64),
                t_status=np.array([t[1] for t in track], dtype=np.int32),
                t_disp=np.array([t[2] for t in track], dtype=np.int32),
                t_nhist=np.array([len(t[3]) for t in track], dtype=np.int32), t_hist=hist)


def main_track():
    d = os.path.join(HERE, "track")
    os.makedirs(d, exist_ok=True)
    for name, g, N, seed, amp, ball, R, nst, thr in TRACK_CASES:
        m = O.model(N, g=g)
        phi, pi = track_inputs(N, seed, amp, ball)
        r = O.run_ref(m, phi, pi, nst, islands=R, thr=thr)
        files = {k: np.array(v) for k, v in r["files"].items()}
        np.savez_compressed(
            os.path.join(d, name + ".npz"),
            N=N, a=m["a"], mass_mode=m["mass_mode"], m2=m["m2"], lambda_mode=m["lambda_mode"],
            kappa=m["kappa"], Lambda=m["Lambda"], analyses=R, steps=nst, thr=r["thr"],
            phi_in=phi, pi_in=pi, file_names=np.array(sorted(files)),
            **{"file_" + k.replace(".", "_"): v for k, v in files.items()}, **pack_track(r["track"]))
        print("wrote", name, "tracked:", len(r["track"]), "files:", sorted(files))


def main_config():
    """SaveConfig (fieldanalysis.cpp:252-282) of the cfg0 fixture's final state."""
    d = os.path.join(HERE, "config")
    os.makedirs(d, exist_ok=True)
    N = 8
    m = O.model(N)
    phi, pi = _fields(N, 1, 2.449489742783178, 0.4)
    r = O.run_ref(m, phi, pi, 40, save_config=7)
    np.savez_compressed(os.path.join(d, "saveconfig_n8_s40.npz"), N=N, steps=40, name=7,
                        phi_out=r["phi"], pi_out=r["pi"],
                        field_text=np.array(r["files"]["field-7-0.out"]),
                        moment_text=np.array(r["files"]["moment-7.out"]))
    print("wrote saveconfig_n8_s40")


# the drop-in driver's parameters in tests/test_gpu_dropin.py
def driver_info_args():
    N = 16
    m = O.model(N)
    return ["-runs", "1", "-Nt", str(N), "-at", "0.5", "-M", "1", "-Mmode", "1", "-L", "6", "-Lmode", "-1",
            "-Kappa", "%.17g" % m["kappa"], "-T0", "0", "-TMax", "2.0", "-dt", "0.1", "-TSO", "0.5", "-TMO", "1.0",
            "-Noise", "0.5", "-Ampl", "0", "-QS", "%.17g" % (3.0 ** 0.5 / 0.5), "-Q2", "%.17g" % (3.0 ** 0.5 / 0.5),
            "-phi0_0", "%.17g" % 6.0 ** 0.5, "-seed", "1234", "-islands", "1"]


def main_info():
    """Print_Info_Console (stdout) and Print_Info_File (info.txt) of the
    reference for the drop-in driver test's parameters (fieldanalysis.cpp:81-247)."""
    import json
    import subprocess
    import tempfile
    d = os.path.join(HERE, "info")
    os.makedirs(d, exist_ok=True)
    args = driver_info_args()
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([O.REF_HARNESS, "--info", "-"] + args, cwd=td, capture_output=True, text=True,
                           timeout=60)
        if r.returncode != 0:
            raise SystemExit(r.stderr)
        info = open(os.path.join(td, "output", "0", "info.txt")).read()
    with open(os.path.join(d, "driver_info.json"), "w") as f:
        json.dump({"args": args, "console": r.stdout, "info_txt": info}, f, indent=1)
    print("wrote info/driver_info.json")
    # command-line edge cases of the -key value parser (cfile.c:55-73,252-278):
    # a later duplicate, an unknown key, a key without a value at the end,
    # exponent notation, a negative value first, an over-long key, a
    # fractional value for an integer key, no arguments at all
    cases = {"duplicate": args + ["-Nt", "24"], "unknown_key": args + ["-foo", "3"],
             "key_without_value": args + ["-TSO"], "exponent": args + ["-TMax", "1e3"],
             "negative_first": ["-Lmode", "-1"] + args, "long_key": args + ["-" + "x" * 100, "5"],
             "fraction_for_int": args + ["-Nt", "24.7"], "no_arguments": []}
    out = {}
    for name, a in cases.items():
        with tempfile.TemporaryDirectory() as td:
            r = subprocess.run([O.REF_HARNESS, "--info", "-"] + a, cwd=td, capture_output=True, text=True,
                               timeout=60)
            if r.returncode != 0:
                raise SystemExit(r.stderr)
            out[name] = {"args": a, "console": r.stdout,
                         "info_txt": open(os.path.join(td, "output", "0", "info.txt")).read()}
    with open(os.path.join(d, "parse_cases.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote info/parse_cases.json")


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/ref_harness is not built (needs the reference tree)")
    for name, mk, N, steps, seed, phi0, amp, isl, thr, T0, dtb in CASES:
        m = O.model(N, **mk)
        phi, pi = _fields(N, seed, phi0, amp)
        r = O.run_ref(m, phi, pi, steps, islands=isl, thr=thr, T0=T0, dtbase=dtb)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            N=N, steps=steps, T0=T0, dtbase=dtb, a=m["a"], mass_mode=m["mass_mode"], m2=m["m2"],
            lambda_mode=m["lambda_mode"], kappa=m["kappa"], Lambda=m["Lambda"],
            phi_in=phi, pi_in=pi, phi_out=r["phi"], pi_out=r["pi"], avg=r["avg"], ext=r["ext"],
            extpos=r["extpos"], islands=np.int32(isl), thr=r["thr"], catalog=r["cat"],
            labels=r["labels"])
        print("wrote", name, "islands:", r["cat"].shape[0])


if __name__ == "__main__":
    if sys.argv[1:] == ["info"]:
        main_info()
    else:
        main()
        main_track()
        main_config()
        main_info()
