"""CPU check of the blocked Householder tridiagonalisation that hetrd_lower runs on the GPU
(kylin-v_amd/csrc/kv_kernels.hip: trd_column_kernel, trd_matvec_kernel; holstein/svd.h:
hetrd_lower), restated step for step in numpy and compared with LAPACK zhetrd('L') from
scipy, an independent implementation: d, e, tau and the reflector tails must agree to
rounding.  The restatement keeps the kernels' split of the work:

* per column j (local index i in a panel of nb), the column step finishes the previous
  column's companion, w' = tau (y - V q1 - W q2) and alpha = -tau (w'^H v) / 2, brings
  column j up to date with the panel, using v(j) = 1 to fold the previous column's part
  into -v conj(w'(j)) - w' - 2 Re(alpha) v, and reflects it with larfg's formulas;
* the matvec step forms y = A v over the trailing matrix as it stood at the panel start,
  plus q1 = W^H v and q2 = V^H v;
* after each panel, A -= V W^H + W V^H on the trailing matrix.

A mistake in any of those (a dropped panel term, the wrong sign of alpha, conj on the wrong
factor, a missed row) changes d, e or tau well above rounding.
"""
import numpy as np
import pytest

la = pytest.importorskip("scipy.linalg.lapack")


def _larfg(alpha, x):
    xn2 = float(np.vdot(x, x).real)
    if xn2 == 0.0 and alpha.imag == 0.0:
        return alpha.real, 0j, np.complex128(1.0)
    beta = -np.copysign(np.sqrt(alpha.real ** 2 + alpha.imag ** 2 + xn2), alpha.real)
    tau = complex((beta - alpha.real) / beta, -alpha.imag / beta)
    return beta, tau, 1.0 / (alpha - beta)


def trd_blocked(G, nb, slip=None):
    """Lower tridiagonalisation of the Hermitian G (only its lower triangle is read).
    Returns (A, d, e, tau) in zhetrd('L') storage.  slip plants a mistake (tests only)."""
    n = G.shape[0]
    A = np.tril(G) + np.tril(G, -1).conj().T  # both triangles, as the device keeps it
    A[np.diag_indices(n)] = A.diagonal().real
    d = np.zeros(n)
    e = np.zeros(max(n - 1, 0))
    tau = np.zeros(max(n - 1, 0), complex)
    for p in range(0, n, nb):
        pw = min(nb, n - p)
        V = np.zeros((n, nb), complex)
        W = np.zeros((n, nb), complex)
        y = np.zeros(n, complex)
        q1 = np.zeros(nb, complex)
        q2 = np.zeros(nb, complex)

        def column(j, i, prev, col):
            r = slice(j, n)
            k = i - 1 if prev else 0
            if prev:
                v = V[r, i - 1]
                wp = tau[j - 1] * (y[r] - V[r, :k] @ q1[:k] - W[r, :k] @ q2[:k])
                alpha = -0.5 * tau[j - 1] * np.vdot(wp, v)
                W[r, i - 1] = wp + alpha * v
            if not col:
                return
            if slip == "panel_conj":
                c = A[r, j] - np.conj(V[r, :k]) @ W[j, :k] - W[r, :k] @ np.conj(V[j, :k])
            else:
                c = A[r, j] - V[r, :k] @ np.conj(W[j, :k]) - W[r, :k] @ np.conj(V[j, :k])
            if prev:
                c = c - v * np.conj(wp[0]) - wp - (0.0 if slip == "alpha_term" else 2.0 * alpha.real) * v
            A[r, j] = c
            d[j] = c[0].real
            if j <= n - 2:
                beta, t, z = _larfg(c[1], c[2:])
                e[j], tau[j] = beta, t
                vt = c[2:] * z if t != 0 else c[2:]
                V[j + 1, i] = 1.0
                V[j + 2:, i] = vt
                A[j + 1, j] = beta
                A[j + 2:, j] = vt

        for i in range(pw):
            j = p + i
            column(j, i, i > 0, True)
            if j <= n - 2:
                s = j + 1
                v = V[s:, i]
                y[:] = 0.0
                y[s:] = A[s:, s:].conj().T @ v  # column dot products
                q1[:i] = W[s:, :i].conj().T @ v
                q2[:i] = V[s:, :i].conj().T @ v
        t = p + pw
        if t > n - 1:
            break
        column(t, pw, True, False)
        A[t:, t:] -= V[t:, :pw] @ W[t:, :pw].conj().T + W[t:, :pw] @ V[t:, :pw].conj().T
    return A, d, e, tau


def _gram(rng, n, rank):
    M = rng.standard_normal((n, rank)) + 1j * rng.standard_normal((n, rank))
    return M @ M.conj().T


@pytest.mark.parametrize("n,nb", [(1, 4), (2, 4), (3, 2), (7, 3), (33, 8), (64, 32), (100, 32)])
def test_blocked_reduction_equals_lapack(n, nb):
    rng = np.random.default_rng(n * 31 + nb)
    G = _gram(rng, n, n + 3)
    A, d, e, tau = trd_blocked(np.tril(G), nb)
    c, d0, e0, t0, info = la.zhetrd(G, lower=1)
    assert info == 0
    scale = np.abs(G).max()
    assert np.abs(d - d0).max() <= 1e-13 * n * scale
    if n > 1:
        assert np.abs(e - e0).max() <= 1e-13 * n * scale
        assert np.abs(tau - t0[:n - 1]).max() <= 1e-12 * n
        assert np.abs(np.tril(A, -2) - np.tril(c, -2)).max(initial=0.0) <= 1e-12 * n


def test_rank_deficient_and_padded_sectors():
    """A rank-5 Gram matrix, alone and padded as diag(G, -I): eigenvalues of the tridiagonal
    form equal G's, and the padding stays exactly decoupled (e = 0, tau = 0 there)."""
    rng = np.random.default_rng(3)
    G = _gram(rng, 40, 5)
    P = np.zeros((48, 48), complex)
    P[:40, :40] = G
    P[40:, 40:] = -np.eye(8)
    for X in (G, P):
        _, d, e, tau = trd_blocked(np.tril(X), 8)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.abs(np.linalg.eigvalsh(T) - np.linalg.eigvalsh(X)).max() <= 1e-12 * np.abs(X).max()
    assert np.all(e[39:] == 0.0) and np.all(tau[39:] == 0.0)


@pytest.mark.parametrize("slip", ["alpha_term", "panel_conj"])
def test_planted_mistake_is_caught(slip):
    """The same restatement with a plausible slip (the 2 Re(alpha) v term of the column update
    dropped, or conj on the wrong factor of the panel term) must disagree with LAPACK far
    above rounding, so the agreement above is not vacuous."""
    rng = np.random.default_rng(9)
    G = _gram(rng, 20, 23)
    c, d0, e0, t0, info = la.zhetrd(G, lower=1)
    _, d, e, _ = trd_blocked(np.tril(G), 4)
    assert max(np.abs(d - d0).max(), np.abs(e - e0).max()) < 1e-11
    _, d, e, _ = trd_blocked(np.tril(G), 4, slip=slip)
    assert max(np.abs(d - d0).max(), np.abs(e - e0).max()) > 1e-3


def trd_fused(G, slip=None):
    """trd_small_kernel's order: unblocked, each step's update A -= v w^H + w v^H applied to
    the next column on the fly and to A(j+2:, j+2:) in the same pass that forms the next
    y = A v.  Returns (A, d, e, tau) in zhetrd('L') storage."""
    n = G.shape[0]
    A = np.tril(G) + np.tril(G, -1).conj().T
    A[np.diag_indices(n)] = A.diagonal().real
    d = np.zeros(n)
    e = np.zeros(max(n - 1, 0))
    tau = np.zeros(max(n - 1, 0), complex)

    def reflect(j, col):
        beta, t, z = _larfg(col[j + 1], col[j + 2:])
        e[j], tau[j] = beta, t
        v = np.zeros(n, complex)
        v[j + 1] = 1.0
        v[j + 2:] = col[j + 2:] * z
        A[j + 1, j] = beta
        A[j + 2:, j] = v[j + 2:]
        return v

    d[0] = A[0, 0].real
    if n == 1:
        return A, d, e, tau
    v = reflect(0, A[:, 0].copy())
    y = np.zeros(n, complex)
    y[1:] = A[1:, 1:] @ v[1:]
    for j in range(n - 1):
        r = slice(j + 1, n)
        w = np.zeros(n, complex)
        w[r] = tau[j] * y[r]
        alpha = -0.5 * tau[j] * np.vdot(w[r], v[r])
        if slip == "alpha_sign":
            alpha = -alpha
        w[r] += alpha * v[r]
        c1 = j + 1
        col = np.zeros(n, complex)
        col[c1:] = A[c1:, c1] - v[c1:] * np.conj(w[c1]) - w[c1:] * np.conj(v[c1])
        d[c1] = col[c1].real
        if c1 + 1 >= n:
            break
        v1 = reflect(c1, col)
        s = slice(c1 + 1, n)
        A[s, s] -= np.outer(v[s], w[s].conj()) + np.outer(w[s], v[s].conj())
        y = np.zeros(n, complex)
        y[s] = A[s, s] @ v1[s]
        v = v1
    return A, d, e, tau


@pytest.mark.parametrize("n", [1, 2, 3, 9, 64, 130])
def test_fused_unblocked_reduction_equals_lapack(n):
    """trd_small_kernel's restatement gives zhetrd('L')'s d, e, tau and reflector tails."""
    rng = np.random.default_rng(n + 5)
    G = _gram(rng, n, n + 2)
    A, d, e, tau = trd_fused(np.tril(G))
    c, d0, e0, t0, info = la.zhetrd(G, lower=1)
    assert info == 0
    scale = np.abs(G).max()
    assert np.abs(d - d0).max() <= 1e-13 * n * scale
    if n > 1:
        assert np.abs(e - e0).max() <= 1e-13 * n * scale
        assert np.abs(tau - t0[:n - 1]).max() <= 1e-12 * n
        assert np.abs(np.tril(A, -2) - np.tril(c, -2)).max(initial=0.0) <= 1e-12 * n


def test_fused_reduction_planted_sign_is_caught():
    rng = np.random.default_rng(11)
    G = _gram(rng, 20, 23)
    c, d0, e0, t0, info = la.zhetrd(G, lower=1)
    _, d, e, _ = trd_fused(np.tril(G), slip="alpha_sign")
    assert max(np.abs(d - d0).max(), np.abs(e - e0).max()) > 1e-3
