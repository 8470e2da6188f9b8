"""CPU checks of the boundary: the engine library loads and exports every
entry point declared in include/*.h, and the Python binding declares the same
set.  No compute call is made (no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(sl_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"sl_semi_local_comb", "sl_batch_create", "sl_batch_run", "sl_batch_run_per_pair"} <= names


def test_library_exports_every_declared_symbol():
    from semilocal_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsemilocal_hip.so not built (run python -m semilocal_amd.build)")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from semilocal_amd import _lib

    assert declared_symbols() == set(_lib.SIGNATURES)


def test_product_path_has_no_oracle_import():
    # the package must never reach into the checker
    for path in glob.glob(os.path.join(ROOT, "semilocal_amd", "**", "*.py"), recursive=True):
        src = open(path).read()
        assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, flags=re.M), path


def test_symbols_are_checked_not_cast():
    # distinct symbols must never collapse to one by a silent narrowing cast
    import numpy as np

    from semilocal_amd._lib import _perm, as_symbols

    assert as_symbols(np.array([1, 2], np.int64))[1] == 4
    assert as_symbols(np.array([True, False]))[1] == 1
    with pytest.raises(ValueError):
        as_symbols(np.array([1, 2**32 + 1], np.int64))
    with pytest.raises(ValueError):
        as_symbols(np.array([1.2, 1.7]))
    with pytest.raises(ValueError):
        as_symbols(np.array([2**31], np.uint32))
    with pytest.raises(ValueError):
        _perm(np.array([0.0, 1.0]))
    with pytest.raises(ValueError):
        _perm(np.array([0, 2**40], np.int64))
