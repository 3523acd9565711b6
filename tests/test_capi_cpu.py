"""The C-ABI library loads on a machine without a GPU and exports every entry
point declared in include/*.h; host-only paths (Boys table fill, argument and
layout validation that returns before any launch) behave as documented."""

import ctypes as C
import os
import re

import numpy as np

import oracle
from vlxmi import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(vlxmi_\w+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    L = runtime.lib()
    names = _declared()
    assert len(names) >= 36 + 54 + 6 + 6
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(runtime.exported_symbols()) <= names


def test_header_is_current():
    from vlxmi.gen import build
    with open(build.HEADER) as f:
        assert f.read() == build.header_text()


def test_boys_table_fill_matches_oracle():
    """Fine rows: F_0..F_14 and exp(-T) at T = 0.02 k, k = 0..2000 (T in [0, 40]);
    then coarse rows: F_0..F_15 at T = 0.4 k, k = 0..100."""
    table, ft = runtime.boys_tables_host()
    t = table.reshape(-1, 16)
    nfine, ncoarse = 2001, 101
    assert t.shape[0] == nfine + ncoarse
    dt = 40.0 / (nfine - 1)
    for k in (0, 1, 37, nfine // 2, nfine - 2, nfine - 1):
        F = oracle.boys(14, dt * k)
        assert np.max(np.abs(t[k, :15] - F) / F) < 2e-15
        assert abs(t[k, 15] - np.exp(-dt * k)) <= 1e-16 * max(1.0, np.exp(-dt * k))
    for k in (0, 1, 2, 50, 99, 100):
        F = oracle.boys(15, 0.4 * k)
        assert np.max(np.abs(t[nfine + k] - F) / F) < 2e-15, k
    assert np.allclose(ft, 1.0 / (2 * np.arange(32) + 1))


def test_launchers_validate_layout_before_launch():
    L = runtime.lib()
    fn = L.vlxmi_launch_computeCoulombFockSPPP
    # bra count 4 is not a whole number of sp blocks (3 entries each) -> layout error, no launch
    args = [None, None, 1, None, 1, None, None, None, None, None, None, 4, None, None, None, 9, None, None, 1e-12]
    assert fn(*args, 1, None) == -2
    # empty lists: nothing to do
    args[11] = 0
    assert fn(*args, 1, None) == 0
    # non-empty lists with a null output -> argument error, no launch
    args[11] = 3
    assert fn(*args, 1, None) == -1
    kf = L.vlxmi_launch_computeExchangeFockSSSS
    kargs = [None, None, None, 0, None, None, 1, 1.0, None, 1, None, None, None, None, None, None, None, 0.0, 1e-12]
    assert kf(*kargs, None) == 0
    assert "block order" in L.vlxmi_status_string(-2).decode()
