"""CPU checks of the drop-in boundary: libflare_mi.so loads, exports every
entry point include/flare_mi.h declares, links nothing from the oracle, and
the Python mirror fails loudly without the library (no CPU fallback).
No compute call is made here (no GPU in this suite)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flare_mi.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fm_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "fm_predict" in syms and "fm_structure_create" in syms and "fm_last_error" in syms
    from flare_amd import _lib
    assert sorted(_lib.EXPORTS) == syms


def test_library_exports_every_symbol():
    from flare_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libflare_mi.so not built")
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_symbols():
        assert name in exported
    # The product never carries the oracle.
    assert not any(s.startswith("ref_") for s in exported)


def test_library_is_built_for_gfx950():
    from flare_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libflare_mi.so not built")
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_missing_library_raises(monkeypatch):
    from flare_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libflare_mi.so")
    with pytest.raises(_lib.FlareMIError):
        _lib.load()


def test_b2_settings_validation():
    from flare_amd._C_flare import B2
    b = B2("chebyshev", "quadratic", [0, 5.0], [], [2, 4, 3])
    assert b.n_descriptors == 8 * 9 // 2 * 4
    assert b.cutoffs.shape == (2, 2)
    with pytest.raises(ValueError):
        B2("bessel", "quadratic", [0, 5.0], [], [1, 4, 3])._settings()
    with pytest.raises(ValueError):
        B2("chebyshev", "polynomial", [0, 5.0], [], [1, 4, 3])._settings()
