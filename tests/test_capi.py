"""The C-ABI library loads and exports every entry point include/kapdqmc.h declares.
No compute call is made (no GPU here); only argument validation that returns before any
HIP call is exercised."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kapdqmc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(kdq_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    from kapdqmc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from kapdqmc_amd.build import build
        build()
    return _lib.load()


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "kdq_create" in syms and "kdq_phy_update" in syms and len(syms) >= 15


def test_every_declared_symbol_is_exported(lib):
    from kapdqmc_amd import _lib
    for name in declared_symbols():
        assert hasattr(lib, name), f"{name} missing from {_lib.LIB_PATH}"
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_argument_errors_without_device(lib):
    from kapdqmc_amd import _lib
    h = C.c_void_p()
    assert lib.kdq_create(None, None, 0, C.byref(h)) == _lib.KDQ_EARG
    assert b"nchains" in lib.kdq_last_error(None)
    assert lib.kdq_set_fields(None, 0, None) == _lib.KDQ_EARG
    assert lib.kdq_phy_update(None, 1, 0, None) == _lib.KDQ_EARG


def test_model_validation_before_any_device_call(lib):
    from kapdqmc_amd import _lib
    md = (_lib.ModelDesc * 1)()
    rd = (_lib.RunDesc * 1)()
    md[0].kind = 5
    h = C.c_void_p()
    assert lib.kdq_create(md, rd, 1, C.byref(h)) == _lib.KDQ_EARG
    assert b"unknown model kind" in lib.kdq_last_error(None)


def test_no_cpu_fallback_when_library_missing(tmp_path):
    from kapdqmc_amd import _lib
    with pytest.raises(ImportError):
        _lib._lib_saved = _lib._lib
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "missing.so"))
        finally:
            _lib._lib = _lib._lib_saved


def test_complex_model_matrices_validated_before_any_device_call(lib):
    """Imaginary parts come all six or none, t-U only, FP64 only (checked before HIP)."""
    import numpy as np
    from kapdqmc_amd import _lib
    from kapdqmc_amd.dqmc import _descriptors
    from kapdqmc_amd.models import tU_Hubbard_Para
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=0.5, Θquench=0.0, Lattice="SQUARE90",
                        site=[3, 2], BatchSize=5, flux=np.pi)
    h = C.c_void_p()
    md, rd, keep = _descriptors([m], [1], [0], 8, 0, precision=_lib.PREC_MIXED)
    assert bool(md[0].K_im) and bool(md[0].Pt_im)          # the imaginary parts are passed, not dropped
    assert lib.kdq_create(md, rd, 1, C.byref(h)) == _lib.KDQ_EARG
    assert b"FP64" in lib.kdq_last_error(None)
    md, rd, keep = _descriptors([m], [1], [0], 8, 0)
    md[0].Pt_im = None
    assert lib.kdq_create(md, rd, 1, C.byref(h)) == _lib.KDQ_EARG
    assert b"all six" in lib.kdq_last_error(None)
