"""CPU-only checks of the C ABI boundary: libmela.so loads, exports every
function include/mela.h declares, and the host-side validation (which runs
before any device work) returns the reference's messages in the reference's
order.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mela.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(mela_[a-z0-9_]+)\s*\(", src))
    names.discard("mela_alloc_logical_fn")
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    import mex_linear_algebra_amd as m
    return m.load()


def test_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string(lib):
    assert b"gfx950" in lib.mela_version()


def _err(lib):
    return lib.mela_last_error().decode(), lib.mela_last_error_id().decode()


def test_matmul_validation_order(lib):
    # matmul_gf2.c:78-88: sparse first, then inner dimensions, then classes
    A = np.zeros(6, np.uint8)
    st = lib.mela_matmul_gf2(None, A.ctypes.data, 0x100, 2, 3, A.ctypes.data, 2, 4, 2, None)
    assert st == 1 and _err(lib)[0] == "Sparse matrices not supported. Convert to full using full()."
    st = lib.mela_matmul_gf2(None, A.ctypes.data, 2, 2, 3, A.ctypes.data, 2, 4, 2, None)
    assert st == 2 and _err(lib)[0] == "Inner dimensions must agree."
    st = lib.mela_matmul_m4ri(None, A.ctypes.data, 0, 2, 3, A.ctypes.data, 2, 3, 2, None)
    assert st == 3 and _err(lib)[0] == "Inputs must be logical or double matrices."
    st = lib.mela_matmul_m4ri(None, A.ctypes.data, 0, 2, 3, A.ctypes.data, 1, 3, 2, None)
    assert st == 5  # valid arguments, null context


def test_rank_and_null_messages(lib):
    r = C.c_double()
    st = lib.mela_rank_gf2(None, None, 2, 3, 3, C.byref(r))
    assert st == 3 and _err(lib) == ("Input must be a logical or double matrix.",
                                     "bitrank_mex:inputNotLogical")
    st = lib.mela_rank_m4ri(None, None, 2, 3, 3, C.byref(r))
    assert st == 3 and _err(lib)[0] == "Input must be double or logical."
    st = lib.mela_rank_m4ri(None, None, 0x101, 3, 3, C.byref(r))
    assert st == 1
    import mex_linear_algebra_amd as m
    nul = m.ALLOC_FN(lambda u, a, b: None)
    f = C.c_size_t()
    st = lib.mela_null_gf2(None, None, 0x100, 3, 3, nul, None, C.byref(f))
    assert st == 1 and _err(lib)[0].startswith("Sparse matrices not supported")
    st = lib.mela_null_m4ri(None, None, 2, 3, 3, nul, None, C.byref(f))
    assert st == 3 and _err(lib)[0] == "H must be double or logical."


def test_python_mirror_classifies_inputs():
    import mex_linear_algebra_amd as m
    _, code, r, c = m._classify(np.zeros((3, 4), bool))
    assert (code, r, c) == (m.MELA_LOGICAL, 3, 4)
    _, code, *_ = m._classify(np.zeros((3, 4)))
    assert code == m.MELA_DOUBLE
    _, code, *_ = m._classify(np.zeros((3, 4), np.int32))
    assert code == m.MELA_OTHER
    buf, *_ = m._classify(np.arange(12.0).reshape(3, 4))
    assert buf.flags.f_contiguous and buf[1, 2] == 6.0
