"""The six MEX gateways (mex/mela_gateway.c) driven through the mex.h shim,
as MATLAB would call them: arity and validation errors with the reference's
messages (CPU), and results against the oracle (GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MEXDIR = os.path.join(ROOT, "mex_linear_algebra_amd", "mex")
FNS = ["mela_matmul_gf2", "mela_matmul_m4ri", "mela_rank_gf2", "mela_rank_m4ri",
       "mela_null_gf2", "mela_null_m4ri"]
MEXFN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p))


@pytest.fixture(scope="module")
def shim():
    path = os.path.join(MEXDIR, "libmexshim.so")
    if not os.path.exists(path):
        pytest.skip("gateways not built")
    import mex_linear_algebra_amd as m
    m.load()  # libmela.so first, so the gateways resolve against it
    return _shim_types(C.CDLL(path, mode=C.RTLD_GLOBAL))


def _shim_types(S):
    S.shim_create.restype = C.c_void_p
    S.shim_create.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_int, C.c_void_p]
    S.shim_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int,
                            C.POINTER(C.c_void_p)]
    S.shim_last_msg.restype = C.c_char_p
    S.shim_last_id.restype = C.c_char_p
    S.mxGetM.argtypes = [C.c_void_p]
    S.mxGetM.restype = C.c_size_t
    S.mxGetN.argtypes = [C.c_void_p]
    S.mxGetN.restype = C.c_size_t
    S.mxGetData.argtypes = [C.c_void_p]
    S.mxGetData.restype = C.c_void_p
    S.mxGetPr.argtypes = [C.c_void_p]
    S.mxGetPr.restype = C.c_void_p
    S.mxIsDouble.argtypes = [C.c_void_p]
    S.mxIsDouble.restype = C.c_bool
    S.mxIsComplex.argtypes = [C.c_void_p]
    S.mxIsComplex.restype = C.c_bool
    return S


def _il_shim():
    """The stand-in built for MATLAB's interleaved complex API (mex -R2018a),
    loaded privately so that it and its gateways do not bind to the default
    build's symbols."""
    path = os.path.join(MEXDIR, "il", "libmexshim.so")
    if not os.path.exists(path):
        pytest.skip("interleaved-complex gateways not built")
    return _shim_types(C.CDLL(path, mode=C.RTLD_LOCAL))


def gateway(name, il=False):
    path = os.path.join(MEXDIR, "il" if il else "", name + ".mexa64")
    L = C.CDLL(path, mode=(os.RTLD_DEEPBIND | C.RTLD_LOCAL) if il else C.RTLD_LOCAL)
    return C.cast(L.mexFunction, C.c_void_p).value, L


def mx(shim, A, sparse=False):
    A = np.asarray(A)
    if A.dtype == np.complex128:
        buf = np.asfortranarray(A)  # (re, im) pairs in column-major order
        return shim.shim_create(3, A.shape[0], A.shape[1], int(sparse), buf.ctypes.data if buf.size else None)
    if A.dtype == np.bool_:
        cls, buf = 0, np.asfortranarray(A).view(np.uint8)
    elif A.dtype == np.float64:
        cls, buf = 1, np.asfortranarray(A)
    else:
        cls, buf = 2, np.asfortranarray(A.astype(np.int32))
    return shim.shim_create(cls, A.shape[0], A.shape[1], int(sparse), buf.ctypes.data if buf.size else None)


def call(shim, name, nlhs, *args, il=False):
    fn, keep = gateway(name, il)
    prhs = (C.c_void_p * max(len(args), 1))(*args)
    plhs = (C.c_void_p * max(nlhs, 1))()
    raised = shim.shim_call(fn, nlhs, plhs, len(args), prhs)
    if raised:
        return None, shim.shim_last_msg().decode(), shim.shim_last_id().decode()
    return plhs, None, None


def test_arity_messages(shim):
    a = mx(shim, np.zeros((2, 2), bool))
    assert call(shim, "mela_matmul_gf2", 1, a)[1] == "Need two input matrices A,B."
    assert call(shim, "mela_matmul_m4ri", 2, a, a)[1] == "Too many outputs."
    assert call(shim, "mela_rank_gf2", 1)[1:] == ("One input required.", "bitrank_mex:nargin")
    assert call(shim, "mela_rank_m4ri", 1, a, a)[1] == "One input required."
    assert call(shim, "mela_null_gf2", 1)[1] == "One input required: H."
    assert call(shim, "mela_null_m4ri", 2, a)[1] == "One output required."


def test_validation_messages(shim):
    a = mx(shim, np.zeros((2, 3), bool))
    b = mx(shim, np.zeros((4, 2), bool))
    i32 = mx(shim, np.zeros((3, 3), np.int32))
    sp = mx(shim, np.zeros((3, 3), bool), sparse=True)
    assert call(shim, "mela_matmul_gf2", 1, a, b)[1] == "Inner dimensions must agree."
    assert call(shim, "mela_matmul_m4ri", 1, sp, a)[1].startswith("Sparse matrices not supported")
    assert call(shim, "mela_matmul_gf2", 1, a, i32)[1] == "Inputs must be logical or double matrices."
    assert call(shim, "mela_rank_gf2", 1, i32)[1:] == ("Input must be a logical or double matrix.",
                                                       "bitrank_mex:inputNotLogical")
    assert call(shim, "mela_rank_m4ri", 1, i32)[1] == "Input must be double or logical."
    assert call(shim, "mela_null_gf2", 1, sp)[1].startswith("Sparse matrices not supported")
    assert call(shim, "mela_null_m4ri", 1, i32)[1] == "H must be double or logical."


def _logical(shim, ptr):
    m, n = shim.mxGetM(ptr), shim.mxGetN(ptr)
    buf = (C.c_uint8 * max(m * n, 1)).from_address(shim.mxGetData(ptr))
    return np.frombuffer(buf, dtype=np.uint8, count=m * n).reshape((m, n), order="F").astype(bool)


@pytest.mark.gpu
def test_gateways_on_device(shim, orc):
    rng = np.random.default_rng(17)
    A = rng.integers(0, 2, (70, 90)).astype(bool)
    B = rng.integers(-3, 4, (90, 40)).astype(np.float64)
    want = orc.matmul(A, B, "gf2")
    for fn in ("mela_matmul_gf2", "mela_matmul_m4ri"):
        plhs, err, _ = call(shim, fn, 1, mx(shim, A), mx(shim, B))
        assert err is None, err
        assert np.array_equal(_logical(shim, plhs[0]), want)
    for fn in ("mela_rank_gf2", "mela_rank_m4ri"):
        plhs, err, _ = call(shim, fn, 1, mx(shim, A))
        assert err is None
        r = C.cast(shim.mxGetData(plhs[0]), C.POINTER(C.c_double))[0]
        assert r == orc.rank(A, "gf2")
    for fn in ("mela_null_gf2", "mela_null_m4ri"):
        plhs, err, _ = call(shim, fn, 1, mx(shim, A))
        assert err is None
        assert np.array_equal(_logical(shim, plhs[0]), orc.null(A, "gf2"))


def test_complex_storage_in_both_apis(shim):
    """The stand-in's complex doubles: mxGetData / mxGetPr give the real parts
    under the separate API and (re, im) pairs under the interleaved one; both
    count as double and complex."""
    Z = np.array([[1.5 + 2j, -3 + 0.5j], [4 - 1j, 7 + 7j]])
    for S, il in ((shim, False), (_il_shim(), True)):
        a = mx(S, Z)
        assert S.mxIsDouble(a) and S.mxIsComplex(a)
        n = 8 if il else 4
        got = np.ctypeslib.as_array(C.cast(S.mxGetData(a), C.POINTER(C.c_double)), shape=(n,))
        want = np.asfortranarray(Z).ravel(order="F")
        want = np.column_stack([want.real, want.imag]).ravel() if il else want.real
        assert np.array_equal(got, want)


@pytest.mark.gpu
def test_gateways_complex_double_use_real_part(shim, orc):
    """Complex double inputs count by their real parts (the reference reads
    them through mxGetData / mxGetPr, null_gf2.c:170, rank_m4ri.c:309), under
    both MATLAB array APIs: each gateway's result equals the oracle's for the
    real part alone."""
    rng = np.random.default_rng(29)
    Re = rng.integers(-3, 4, (60, 75)).astype(np.float64)
    Im = rng.integers(-3, 4, (60, 75)).astype(np.float64)
    Z = Re + 1j * Im
    B = rng.integers(0, 2, (75, 30)).astype(bool)
    R = (np.abs(Re.astype(np.int64)) & 1).astype(bool)
    for S, il in ((shim, False), (_il_shim(), True)):
        plhs, err, _ = call(S, "mela_matmul_gf2", 1, mx(S, Z), mx(S, B), il=il)
        assert err is None, err
        assert np.array_equal(_logical(S, plhs[0]), orc.matmul(R, B, "gf2"))
        plhs, err, _ = call(S, "mela_rank_m4ri", 1, mx(S, Z), il=il)
        assert err is None, err
        assert C.cast(S.mxGetData(plhs[0]), C.POINTER(C.c_double))[0] == orc.rank(R, "gf2")
        plhs, err, _ = call(S, "mela_null_gf2", 1, mx(S, Z), il=il)
        assert err is None, err
        assert np.array_equal(_logical(S, plhs[0]), orc.null(R, "gf2"))
