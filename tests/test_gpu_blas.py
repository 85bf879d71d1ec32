"""The reference's BLAS-like entry points (org/qmckl_blas.org): qmckl_dgemm, qmckl_dgemv,
qmckl_dger, qmckl_adjugate and the size-safe proxies.  Column-major operands, checked against
numpy in FP64 and, for the adjugate, against the determinants the reference's own test asserts
(org/qmckl_blas.org:2300-2420: 47.622, 397.2748, 4096.8996 for its N = 3, 4, 5 matrices) and
the identity A adj(A) = det(A) I.  The argument checks run on the CPU (they return before the
GPU is touched)."""
import ctypes

import numpy as np
import pytest

import qmckl_amd as q

def ref_matrix(n):
    """The reference test's matrix (its Python generator, org/qmckl_blas.org:2306-2314)."""
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(i):
            A[i, j] = 0.1 * (i + 1) - (j + 1)
        A[i, i] = 1.0 + (i + 1)
        for j in range(i + 1, n):
            A[i, j] = 0.1 * (i + 1) + (j + 1)
    return A


def fortran(M, ld):
    """M stored column-major with leading dimension ld (rows past M's filled with a marker)."""
    r, c = M.shape
    buf = np.full((c, ld), 7.25)
    buf[:, :r] = M.T
    return buf.reshape(-1)


def unfortran(buf, r, c, ld):
    return buf.reshape(c, ld)[:, :r].T


@pytest.fixture
def ctx():
    c = q.context_create()
    yield c
    q.context_destroy(c)


def test_argument_checks(ctx):
    L = q.lib()
    a = np.zeros(16)
    p = q.ptr(a)
    assert L.qmckl_dgemm(ctx, b"N", b"N", 0, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2) == 4
    assert L.qmckl_dgemm(ctx, b"N", b"N", 2, 0, 2, 1.0, p, 2, p, 2, 0.0, p, 2) == 5
    assert L.qmckl_dgemm(ctx, b"N", b"N", 2, 2, 0, 1.0, p, 2, p, 2, 0.0, p, 2) == 6
    assert L.qmckl_dgemm(ctx, b"N", b"N", 2, 2, 2, 1.0, p, 0, p, 2, 0.0, p, 2) == 9
    assert L.qmckl_dgemm(ctx, b"N", b"N", 2, 2, 2, 1.0, p, 2, p, 0, 0.0, p, 2) == 11
    assert L.qmckl_dgemm(ctx, b"N", b"N", 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 0) == 14
    assert L.qmckl_dgemm(ctx, b"X", b"N", 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2) == 2
    assert L.qmckl_dgemm_safe(ctx, b"N", b"N", 2, 2, 2, 1.0, p, 0, 2, p, 4, 2, 0.0, p, 4, 2) == 9
    assert L.qmckl_dgemm_safe(ctx, b"N", b"N", 2, 2, 2, 1.0, p, 4, 2, p, 3, 2, 0.0, p, 4, 2) == 12
    assert L.qmckl_dgemm_safe(ctx, b"N", b"N", 2, 2, 2, 1.0, p, 4, 2, p, 4, 2, 0.0, p, 4, 0) == 17
    assert L.qmckl_dgemv(ctx, b"N", 0, 2, 1.0, p, 2, p, 1, 0.0, p, 1) == 3
    assert L.qmckl_dgemv(ctx, b"N", 2, 2, 1.0, p, 0, p, 1, 0.0, p, 1) == 7
    assert L.qmckl_dgemv(ctx, b"N", 2, 2, 1.0, p, 2, p, 0, 0.0, p, 1) == 9
    assert L.qmckl_dgemv(ctx, b"N", 2, 2, 1.0, p, 2, p, 1, 0.0, p, 0) == 12
    assert L.qmckl_dger(ctx, 2, 0, 1.0, p, 1, p, 1, p, 2) == 3
    assert L.qmckl_dger(ctx, 2, 2, 1.0, p, 0, p, 1, p, 2) == 6
    assert L.qmckl_dger(ctx, 2, 2, 1.0, p, 1, p, 1, p, 1) == 10
    d = ctypes.c_double(0.0)
    assert L.qmckl_adjugate(ctx, 0, p, 2, p, 2, ctypes.byref(d)) == 2
    assert L.qmckl_adjugate(ctx, 3, p, 2, p, 3, ctypes.byref(d)) == 4
    assert L.qmckl_adjugate_safe(ctx, 3, p, 8, 3, p, 9, 3, ctypes.byref(d)) == 4
    assert L.qmckl_adjugate_safe(ctx, 3, p, 9, 2, p, 9, 3, ctypes.byref(d)) == 5
    assert L.qmckl_dgemm(0, b"N", b"N", 2, 2, 2, 1.0, p, 2, p, 2, 0.0, p, 2) != 0  # QMCKL_NULL_CONTEXT


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k,beta", [(1, 1, 1, 0.0), (7, 5, 3, 0.5), (130, 67, 45, -1.25), (64, 128, 300, 0.0)])
def test_dgemm_vs_numpy(ctx, ta, tb, m, n, k, beta):
    L = q.lib()
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    opA, opB, C0 = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
    A = opA.T if ta == "T" else opA
    B = opB.T if tb == "T" else opB
    lda, ldb, ldc = A.shape[0] + 3, B.shape[0] + 1, m + 2
    a, b, c = fortran(A, lda), fortran(B, ldb), fortran(C0, ldc)
    if beta == 0.0:
        c.reshape(n, ldc)[:, :m] = np.nan  # beta = 0 does not read C
    alpha = 0.75
    q.check(ctx, L.qmckl_dgemm(ctx, ta.encode(), tb.encode(), m, n, k, alpha, q.ptr(a), lda, q.ptr(b), ldb, beta,
                               q.ptr(c), ldc))
    ref = alpha * (opA @ opB) + (beta * C0 if beta != 0.0 else 0.0)
    got = unfortran(c, m, n, ldc)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * k
    assert np.all(c.reshape(n, ldc)[:, m:] == 7.25)  # the rows past m are left as they were
    c2 = fortran(C0, ldc)
    q.check(ctx, L.qmckl_dgemm_safe(ctx, ta.encode(), tb.encode(), m, n, k, alpha, q.ptr(a), a.size, lda, q.ptr(b),
                                    b.size, ldb, beta, q.ptr(c2), c2.size, ldc))
    if beta != 0.0:
        assert np.array_equal(unfortran(c2, m, n, ldc), got)


@pytest.mark.gpu
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,n,incx,incy", [(5, 3, 1, 1), (300, 77, 2, 3)])
def test_dgemv_and_dger_vs_numpy(ctx, trans, m, n, incx, incy):
    L = q.lib()
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    lx, ly = (m, n) if trans == "T" else (n, m)
    x, y0 = rng.standard_normal(lx), rng.standard_normal(ly)
    xs = np.full(incx * (lx - 1) + 1, 3.5)
    xs[::incx] = x
    ys = np.full(incy * (ly - 1) + 1, 3.5)
    ys[::incy] = y0
    lda = m + 1
    a = fortran(A, lda)
    q.check(ctx, L.qmckl_dgemv(ctx, trans.encode(), m, n, 1.5, q.ptr(a), lda, q.ptr(xs), incx, -0.5, q.ptr(ys), incy))
    opA = A.T if trans == "T" else A
    ref = 1.5 * opA @ x - 0.5 * y0
    assert np.abs(ys[::incy] - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * max(m, n)
    gaps = np.ones(ys.size, bool)
    gaps[::incy] = False
    assert np.all(ys[gaps] == 3.5)
    # dger: A += alpha x y^T with x of length m, y of length n
    xm, yn = rng.standard_normal(m), rng.standard_normal(n)
    q.check(ctx, L.qmckl_dger(ctx, m, n, 0.25, q.ptr(xm), 1, q.ptr(yn), 1, q.ptr(a), lda))
    refA = A + 0.25 * np.outer(xm, yn)
    assert np.abs(unfortran(a, m, n, lda) - refA).max() <= 1e-14 * np.abs(refA).max()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 40, 150, 200])
def test_adjugate(ctx, n):
    """Orders 1-5 by cofactors, 6-128 through the inverse in LDS, above in global memory."""
    L = q.lib()
    # above order 6 the reference test's matrix, shifted and scaled so that det(A) stays finite
    # (condition numbers 5-200)
    A = ref_matrix(n) if n <= 6 else (ref_matrix(n) + np.diag(np.full(n, 0.5 * n))) / (0.8 * n)
    lda, ldb = n + 1, n + 2
    a, b = fortran(A, lda), fortran(np.zeros((n, n)), ldb)
    d = ctypes.c_double(0.0)
    q.check(ctx, L.qmckl_adjugate(ctx, n, q.ptr(a), lda, q.ptr(b), ldb, ctypes.byref(d)))
    B = unfortran(b, n, n, ldb)
    golden = {3: 47.62200000000001, 4: 397.2747999999998, 5: 4096.899599999997}
    if n in golden:
        assert abs(d.value - golden[n]) <= 1e-13 * abs(golden[n])
    det = np.linalg.det(A)
    assert abs(d.value - det) <= 1e-10 * abs(det)
    # A adj(A) = det(A) I, as the reference's test checks it
    assert np.abs(A @ B - d.value * np.eye(n)).max() <= 1e-10 * abs(d.value)
    b2 = fortran(np.zeros((n, n)), ldb)
    q.check(ctx, L.qmckl_adjugate_safe(ctx, n, q.ptr(a), a.size, lda, q.ptr(b2), b2.size, ldb, ctypes.byref(d)))
    assert np.array_equal(b2, b)


@pytest.mark.gpu
def test_adjugate_singular(ctx):
    """Orders up to 5 by cofactors: a singular matrix has det 0 and its adjugate; above, the
    inverse-based path refuses it."""
    L = q.lib()
    A = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]])
    a, b = fortran(A, 3), np.zeros(9)
    d = ctypes.c_double(1.0)
    q.check(ctx, L.qmckl_adjugate(ctx, 3, q.ptr(a), 3, q.ptr(b), 3, ctypes.byref(d)))
    cof = np.array([[(-1) ** (i + j) * np.linalg.det(np.delete(np.delete(A, i, 0), j, 1)) for j in range(3)]
                    for i in range(3)])
    assert abs(d.value) < 1e-12
    assert np.abs(unfortran(b, 3, 3, 3) - cof.T).max() < 1e-12
    S = np.ones((8, 8))
    a8, b8 = fortran(S, 8), np.full(64, 2.0)
    assert L.qmckl_adjugate(ctx, 8, q.ptr(a8), 8, q.ptr(b8), 8, ctypes.byref(d)) == q.QMCKL_FAILURE
    assert d.value == 0.0 and np.all(b8 == 2.0)
