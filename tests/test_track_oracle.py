"""The island-tracking checker (oracle/oracle_track.cpp) against the
reference's own Full_Analysis run repeatedly on evolving fields: the
committed fixtures (tests/golden/track/, from oracle/_ref/ref_harness in
tracking mode) and, where the reference tree is built, fresh harness runs.
The tracked list (numbers, statuses, five-entry histories) and the three text
files must be identical."""
import glob
import os

import numpy as np
import pytest

import _oracle as O

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "track", "*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_tracking_oracle_reproduces_reference_fixture(path):
    m, z, track, files = O.load_track_fixture(path)
    got, gfiles = O.track_replay(m, z["phi_in"], z["pi_in"], int(z["analyses"]), int(z["steps"]),
                                 float(z["thr"]))
    assert got == track
    assert gfiles == files


def test_fixtures_cover_statuses():
    # between them the fixtures hold plain islands, an expanding bubble, an
    # oscillon, merges/retirements and more than five analyses per island
    st, names, longest = set(), set(), 0
    for p in FIX:
        _, z, track, files = O.load_track_fixture(p)
        st |= {t[1] for t in track}
        names |= set(files)
        longest = max([longest] + [len(t[3]) for t in track])
        assert files["non_critical_islands.txt"]  # islands were retired
    assert st == {0, 1, 2}
    assert longest == 5
    assert "expanding_bubbles.txt" in names
    assert len(FIX) == 4


@pytest.mark.skipif(not O.have_ref(), reason="reference harness not built (needs the reference tree)")
@pytest.mark.parametrize("N,seed,amp,analyses,steps", [(12, 21, 1.1, 8, 5), (18, 22, 0.8, 6, 10)])
def test_tracking_oracle_matches_reference_harness(N, seed, amp, analyses, steps):
    m = O.model(N)
    rng = np.random.default_rng(seed)
    phi = amp * rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    r = O.run_ref(m, phi, pi, steps, islands=analyses)
    got, files = O.track_replay(m, phi, pi, analyses, steps, r["thr"])
    assert got == r["track"]
    assert files == r["files"]
