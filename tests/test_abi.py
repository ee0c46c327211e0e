"""The C ABI library loads and exports every function include/jgpm.h declares."""
import ctypes
import os
import re

import jupiter_amd as J

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "jgpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(jgpm_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(J.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_strerror_and_no_device_status():
    L = J.lib()
    for s in range(0, -12, -1):
        assert L.jgpm_strerror(s)
    h = ctypes.c_void_p()
    st = L.jgpm_init(0, ctypes.byref(h))
    if st == 0:  # a GPU is present
        L.jgpm_shutdown(h)
    else:
        assert st == -11
    assert L.jgpm_init(0, None) == -1


def test_pattern_names():
    assert J.PATTERNS["P13"] == 13 and J.PATTERNS["TC"] == 0
    assert J.pattern_id("P7") == 7
