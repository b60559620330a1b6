"""The oracle's SPARSKIT and band-solver restatement against the reference's
own routines, compiled from the reference tree into oracle/_ref
(oracle/ref.mk).  Skipped where that library is not built (it is built only
where the reference tree is present).

Same inputs, same arithmetic order: the factors, triangular solves, band
solutions and BiCGStab iterates agree bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle

from test_oracle_cpu import _fem_matrix

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libspkref.so")
ptr = pyoracle.ptr

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the reference tree)")


@pytest.fixture(scope="module")
def ref():
    return C.CDLL(REF)


@pytest.fixture(scope="module")
def lib64():
    return pyoracle.load("f64")


def _i(v):
    return C.byref(C.c_int(v))


def _d(v):
    return C.byref(C.c_double(v))


def _csr1(A):
    """1-based CSR arrays as the Fortran routines take them."""
    return (A.data.copy(), (A.indices + 1).astype(np.int32), (A.indptr + 1).astype(np.int32))


@pytest.mark.parametrize("n_target,seed", [(300, 1), (1200, 4)])
def test_ilut_bitwise(ref, lib64, n_target, seed):
    """ilut(lfil=5, droptol=1e-2), spk_ilut.f90:78, as spk_solve_system calls it."""
    A = _fem_matrix(n_target=n_target, seed=seed)
    n = A.shape[0]
    iwk = 2 * A.nnz
    a, ja, ia = _csr1(A)
    alu_r = np.zeros(iwk + 1)
    jlu_r = np.zeros(iwk + 1, np.int32)
    ju_r = np.zeros(n + 1, np.int32)
    w = np.zeros(n + 1)
    jw = np.zeros(2 * n, np.int32)
    ierr = C.c_int(0)
    ref.ilut_(_i(n), ptr(a), ptr(ja), ptr(ia), _i(5), _d(0.01), ptr(alu_r), ptr(jlu_r), ptr(ju_r), _i(iwk),
              ptr(w), ptr(jw), C.byref(ierr))
    assert ierr.value == 0
    alu_o = np.zeros(iwk + 1)
    jlu_o = np.zeros(iwk + 1, np.int32)
    ju_o = np.zeros(n + 2, np.int32)
    ao, jao, iao = A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32)
    assert lib64.oc_ilut(n, ptr(ao), ptr(jao), ptr(iao), 5, C.c_double(0.01), ptr(alu_o), ptr(jlu_o),
                         ptr(ju_o), iwk) == 0
    m = jlu_r[n] - 1                              # entries used: jlu(n+1) - 1
    # Fortran alu(1..m) <-> oracle alu[1..m]
    assert np.array_equal(jlu_r[:m], jlu_o[1:m + 1])
    assert np.array_equal(ju_r[:n], ju_o[1:n + 1])
    assert np.array_equal(alu_r[:m], alu_o[1:m + 1])


def test_ilu0_and_lusol_bitwise(ref, lib64):
    """ilu0 (spk_ilut.f90:1832) and lusol (:2309)."""
    A = _fem_matrix(n_target=800, seed=2)
    n = A.shape[0]
    a, ja, ia = _csr1(A)
    size = A.nnz + n + 2
    alu_r = np.zeros(size)
    jlu_r = np.zeros(size, np.int32)
    ju_r = np.zeros(n + 1, np.int32)
    iw = np.zeros(n, np.int32)
    ierr = C.c_int(0)
    ref.ilu0_(_i(n), ptr(a), ptr(ja), ptr(ia), ptr(alu_r), ptr(jlu_r), ptr(ju_r), ptr(iw), C.byref(ierr))
    assert ierr.value == 0
    alu_o = np.zeros(size + 1)
    jlu_o = np.zeros(size + 1, np.int32)
    ju_o = np.zeros(n + 2, np.int32)
    ao, jao, iao = A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32)
    assert lib64.oc_ilu0(n, ptr(ao), ptr(jao), ptr(iao), ptr(alu_o), ptr(jlu_o), ptr(ju_o)) == 0
    m = jlu_r[n] - 1
    assert np.array_equal(alu_r[:m], alu_o[1:m + 1]) and np.array_equal(jlu_r[:m], jlu_o[1:m + 1])
    y = np.random.default_rng(5).uniform(-1, 1, n)
    xr = np.zeros(n)
    xo = np.zeros(n)
    ref.lusol_(_i(n), ptr(y), ptr(xr), ptr(alu_r), ptr(jlu_r), ptr(ju_r))
    lib64.oc_lusol(n, ptr(y), ptr(xo), ptr(alu_o), ptr(jlu_o), ptr(ju_o))
    assert np.array_equal(xr, xo)


def test_bcgstab_bitwise(ref, lib64):
    """runrc + bcgstab (spk_itaux.f90, spk_iters.f90:1061) with the
    parameters of spk_solve_system (sparskit_admin.f90:266-304): ipar(2) = 1,
    ipar(3) = 1, work 8n, itermax 1000, fpar(1) = fpar(2) = 1e-8 (a default
    real literal), the ILUT preconditioner, a non-zero initial guess."""
    A = _fem_matrix(n_target=1500, seed=6)
    n = A.shape[0]
    iwk = 2 * A.nnz
    a, ja, ia = _csr1(A)
    alu = np.zeros(iwk + 1)
    jlu = np.zeros(iwk + 1, np.int32)
    ju = np.zeros(n + 1, np.int32)
    ierr = C.c_int(0)
    ref.ilut_(_i(n), ptr(a), ptr(ja), ptr(ia), _i(5), _d(0.01), ptr(alu), ptr(jlu), ptr(ju), _i(iwk),
              ptr(np.zeros(n + 1)), ptr(np.zeros(2 * n, np.int32)), C.byref(ierr))
    rng = np.random.default_rng(7)
    xt = rng.uniform(-1, 1, n)
    rhs = A @ xt
    guess = xt + 1e-3 * rng.uniform(-1, 1, n)
    ipar = np.zeros(16, np.int32)
    fpar = np.zeros(16)
    ipar[:6] = [0, 1, 1, 8 * n, 1, 1000]
    tol = float(np.float32(1e-8))
    fpar[0] = fpar[1] = tol
    sol = np.zeros(n)
    wk = np.zeros(8 * n)
    bcg = C.cast(ref.bcgstab_, C.c_void_p)
    ref.runrc_(_i(n), ptr(rhs), ptr(sol), ptr(ipar), ptr(fpar), ptr(wk), ptr(guess), ptr(a), ptr(ja), ptr(ia),
               ptr(alu), ptr(jlu), ptr(ju), bcg)
    assert ipar[0] == 0, f"reference bcgstab stopped with ipar(1) = {ipar[0]}"
    # the oracle on the same matrix, preconditioner and guess (0-based CSR)
    alu_o = np.zeros(iwk + 2)
    jlu_o = np.zeros(iwk + 2, np.int32)
    ju_o = np.zeros(n + 2, np.int32)
    alu_o[1:] = alu
    jlu_o[1:] = jlu
    ju_o[1:] = ju
    ao, jao, iao = A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32)
    xo = guess.copy()
    its = C.c_int(0)
    rc = lib64.oc_bcgstab(n, ptr(ao), ptr(jao), ptr(iao), ptr(alu_o), ptr(jlu_o), ptr(ju_o), ptr(rhs), ptr(xo),
                          C.c_double(tol), C.c_double(tol), 1000, C.byref(its))
    assert rc == 0
    assert 1 + 2 * its.value == ipar[6]                 # ipar(7): matvecs
    assert np.array_equal(xo, sol)


@pytest.mark.parametrize("seed", range(4))
def test_penta_bitwise(ref, lib64, seed):
    """penta_fact / penta_solve (penta.f90:104-321) on random diagonally
    dominant bands."""
    rng = np.random.default_rng(seed)
    for _ in range(10):
        n = int(rng.integers(4, 41))
        val = rng.uniform(-1, 1, (n, 5))
        val[:, 2] = np.abs(val).sum(1) * rng.choice([-1.0, 1.0], n)
        r = rng.uniform(-1, 1, n)
        vr = val.copy()
        vo = val.copy()
        sr = np.zeros(n)
        so = np.zeros(n)
        ref.penta_fact_(_i(n), ptr(vr))
        ref.penta_solve_(_i(n), ptr(vr), ptr(r), ptr(sr))
        assert lib64.oc_penta_fact(n, ptr(vo)) == 0
        lib64.oc_penta_solve(n, ptr(vo), ptr(r), ptr(so))
        assert np.array_equal(vr, vo) and np.array_equal(sr, so)
