"""CPU-side checks of the C ABI library: it loads, exports every entry point
declared in include/sassena_amd.h, and its host-only planning answers
(no GPU call is made here)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sassena_amd.h")).read()
    return sorted(set(re.findall(r"\b(sx_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from sassena_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    from sassena_amd import _lib
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name


@pytest.mark.parametrize("NF,L", [(1, 512), (200, 512), (256, 512), (257, 1024), (1000, 2048), (1500, 3072),
                                  (2000, 4096), (2500, 5120), (3000, 6144), (4096, 8192), (5000, 10240),
                                  (6000, 12288), (8192, 16384), (8193, 20480), (10240, 20480), (10241, 24576),
                                  (13000, 32768), (20000, 40960), (31000, 65536), (50000, 102400),
                                  (100000, 204800), (163840, 327680), (163841, 0)])
def test_pad_length_plan(NF, L):
    """The correlation length is >= 2NF (linear correlation for tau < NF):
    one-block FFT halves up to 10,240 points, then N1 rows of a one-block
    length (2 N1 Lh), up to 16 rows of 10,240."""
    import sassena_amd
    assert sassena_amd.pad_length(NF) == L
    if L:
        assert L >= 2 * NF


def test_version_string():
    import sassena_amd
    assert "gfx950" in sassena_amd.version()
