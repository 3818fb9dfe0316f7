"""Writes the golden fixtures in this directory from the reference's own
hot-path sources, compiled unmodified with strict IEEE flags
(oracle/_ref/ref_harness, built by oracle/Makefile where the reference tree
exists).  The fixtures are data only: seeded inputs and the reference's
outputs for them.

    python tests/golden/make_golden.py
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
    main()
