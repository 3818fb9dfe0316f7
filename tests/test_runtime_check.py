"""The library refuses to run on a HIP runtime or rocFFT other than the ones
it was built against.  PyTorch's ROCm wheels bundle their own HIP runtime and
rocFFT under the same sonames; a process that imports torch before loading
the library binds the library's calls to those copies.  Checked on the CPU in
a child process (the check comes before any device call)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import torch  # loads torch's bundled libamdhip64 and librocfft first
sys.path.insert(0, %r)
import phinum_amd as pn
maps = [l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l]
try:
    pn.Lattice(8)
    out = {"created": True}
except pn.PhiNumError as e:
    out = {"created": False, "code": e.code, "msg": str(e)}
out["hip"] = sorted(set(maps))
print(json.dumps(out))
"""


def test_library_refuses_another_hip_runtime():
    try:
        import torch  # noqa: F401
    except ImportError:
        pytest.skip("torch is not importable")
    r = subprocess.run([sys.executable, "-c", CHILD % ROOT], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    if not any("/torch/" in p for p in out["hip"]):
        pytest.skip("torch does not bundle its own HIP runtime here")
    assert not out["created"]
    assert out["code"] == -4  # PN_ERR_STATE
    assert "HIP runtime" in out["msg"] or "rocFFT" in out["msg"]
