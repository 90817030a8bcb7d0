"""C-ABI surface: the library loads without a GPU, exports every function include/sdg.h declares,
and reports argument errors with status codes (no compute without a GPU)."""
import ctypes
import os
import re

import pytest

from subrosadg_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "sdg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdg_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = header_functions()
    assert len(names) >= 15
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert sorted(L.SIGNATURES) == names


def test_abi_version():
    assert L.load().sdg_abi_version() == L.ABI_VERSION


def test_invalid_problem_is_rejected_before_device_work():
    m = L.Mesh(L.ELEM_QUADRANGLE, 2, (4, 4), (0, 0), (1, 1))
    lib = L.load()
    h = ctypes.c_void_p()
    bad = L.problem_desc()
    bad.abi_version = 99
    assert lib.sdg_create(ctypes.byref(bad), m.handle, ctypes.byref(h)) == L.E_INVALID
    assert b"abi_version" in lib.sdg_last_error(None)
    bad = L.problem_desc(equation=L.EQ_NS, viscous_flux=L.VISC_NONE, transport=L.TRANSPORT_CONSTANT)
    assert lib.sdg_create(ctypes.byref(bad), m.handle, ctypes.byref(h)) == L.E_INVALID
    bad = L.problem_desc(precision=L.PREC_MIXED)
    assert lib.sdg_create(ctypes.byref(bad), m.handle, ctypes.byref(h)) == L.E_UNSUPPORTED
    assert lib.sdg_create(None, m.handle, ctypes.byref(h)) == L.E_INVALID


def test_export_rejects_small_buffer_and_unknown_field():
    m = L.Mesh(L.ELEM_QUADRANGLE, 1, (2, 2), (0, 0), (1, 1))
    lib = L.load()
    buf = (ctypes.c_double * 2)()
    assert lib.sdg_mesh_export(m.handle, 0, buf, 2) == L.E_INVALID
    assert lib.sdg_mesh_field_size(m.handle, 999) == 0
    assert lib.sdg_mesh_export(m.handle, 999, buf, 2) == L.E_INVALID


def test_create_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    m = L.Mesh(L.ELEM_QUADRANGLE, 2, (4, 4), (0, 0), (1, 1))
    with pytest.raises(L.SDGError, match="no HIP device"):
        L.Solver(L.problem_desc(), m)
