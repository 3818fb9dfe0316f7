"""The checkpoint text of SaveConfig (fieldanalysis.cpp:252-282) as restated
for the checks of the drop-in driver (tests/_oracle.py config_text,
moment_text), against the reference's own SaveConfig output for the cfg0
fixture state (tests/golden/config/) and, where the reference tree is built,
a fresh harness run."""
import os

import numpy as np
import pytest

import _oracle as O

FIX = os.path.join(os.path.dirname(__file__), "golden", "config", "saveconfig_n8_s40.npz")


def test_config_text_matches_reference_fixture():
    z = np.load(FIX)
    assert O.config_text(z["phi_out"], z["pi_out"]) == str(z["field_text"])
    _, _, tau = O.time_factors(int(z["steps"]))
    assert O.moment_text(tau[-1]) == str(z["moment_text"])
    # a wrong digit count or a transposed pair would not match
    assert O.config_text(z["pi_out"], z["phi_out"]) != str(z["field_text"])
    assert O.config_text(z["phi_out"], z["pi_out"], fp32=True) != str(z["field_text"])


@pytest.mark.skipif(not O.have_ref(), reason="reference harness not built (needs the reference tree)")
def test_config_text_matches_reference_harness():
    N = 6
    m = O.model(N)
    rng = np.random.default_rng(9)
    phi = 1.3 * rng.standard_normal(N ** 3)
    pi = 0.2 * rng.standard_normal(N ** 3)
    r = O.run_ref(m, phi, pi, 17, save_config=3, T0=1.5)
    assert O.config_text(r["phi"], r["pi"]) == r["files"]["field-3-0.out"]
    _, _, tau = O.time_factors(17, T0=1.5)
    assert O.moment_text(tau[-1]) == r["files"]["moment-3.out"]
