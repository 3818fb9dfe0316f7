"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import phinum_amd as pn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(pn_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    names = _declared()
    assert len(names) >= 15
    L = pn.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version_and_pure_helpers():
    L = pn.lib()
    assert L.pn_abi_version() == 1
    # nTBins of momentumspace.cpp:40-47 for the 64^3 ... 1024^3 lattices at a = 0.5
    assert [L.pn_spectrum_nbins(n, 0.5) for n in (64, 128, 192, 256, 1024)] == [37, 73, 108, 143, 567]


def test_create_without_gpu_fails_cleanly():
    # argument errors are reported before any device work
    p = pn.PnParams(2, 0.5, 1, 1.0, -1, 2.0, 0, 0)
    h = ctypes.c_void_p()
    rc = pn.lib().pn_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == pn.PN_ERR_ARG
    assert b"N=2" in pn.lib().pn_last_error(None)
