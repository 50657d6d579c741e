"""CPU check of the blocking used by the device back-transformation of the Gram
eigenvectors (gram_eigvecs in the bond SVD and hh_tfactor_kernel): reflectors of LAPACK's
lower-triangle Hermitian tridiagonalisation written out with their unit entries, blocks of
bw reflectors with the forward triangular factor T_b(0:c, c) = -tau_c T_b(0:c, 0:c)
(V_b^H V_b)(0:c, c), applied from the last block as Z <- Z - (V_b T_b)(V_b^H Z) over the
rows from the block's first reflector.  The vectors so formed must be orthonormal
eigenvectors of the matrix (LAPACK zhetrd through scipy; numpy's eigh on the tridiagonal)."""
import numpy as np
import pytest
from scipy.linalg import lapack


def back_transform(c, tau, Zt, bw):
    n = c.shape[0]
    nref = n - 1
    nblk = (nref + bw - 1) // bw
    V = np.zeros((n, nblk * bw), complex)
    for i in range(nref):
        V[i + 1, i] = 1
        V[i + 2:, i] = c[i + 2:, i]
    Z = Zt.astype(complex)
    T = np.zeros((nblk, bw, bw), complex)
    for b in range(nblk):
        i0 = b * bw
        w = min(bw, nref - i0)
        S = V[:, i0:i0 + bw].conj().T @ V[:, i0:i0 + bw]
        for cc in range(w):
            t = tau[i0 + cc]
            T[b, :cc, cc] = -t * (T[b, :cc, :cc] @ S[:cc, cc])
            T[b, cc, cc] = t
    for b in range(nblk - 1, -1, -1):
        Vb = V[:, b * bw:(b + 1) * bw]
        P = Vb @ T[b]
        r0 = b * bw + 1
        Z[r0:] -= P[r0:] @ (Vb[r0:].conj().T @ Z[r0:])
    return Z


@pytest.mark.parametrize("n,lo,bw", [(5, 0, 4), (37, 30, 8), (300, 200, 128), (300, 0, 128), (257, 1, 128),
                                     (129, 64, 128)])
def test_blocked_back_transformation(n, lo, bw):
    rng = np.random.default_rng(n + lo)
    M = rng.standard_normal((2 * n, n)) + 1j * rng.standard_normal((2 * n, n))
    A = M.conj().T @ M
    c, d, e, tau, info = lapack.zhetrd(A, lower=1)
    assert info == 0
    w, Zt = np.linalg.eigh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))
    Z = back_transform(c, tau, Zt[:, lo:], bw)
    scale = np.abs(w).max()
    assert np.abs(A @ Z - Z * w[lo:]).max() <= 1e-13 * scale
    assert np.abs(Z.conj().T @ Z - np.eye(n - lo)).max() <= 1e-13


def test_identity_reflectors_are_skipped():
    """A matrix that is already tridiagonal in places gives tau = 0 reflectors (identity);
    the recurrence must leave their rows and columns of T zero."""
    n = 40
    rng = np.random.default_rng(1)
    B = np.zeros((n, n), complex)
    B[:20, :20] = rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))
    B[20:, 20:] = np.diag(rng.standard_normal(20))
    A = B.conj().T @ B
    c, d, e, tau, info = lapack.zhetrd(A, lower=1)
    assert info == 0 and np.any(tau == 0)
    w, Zt = np.linalg.eigh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))
    Z = back_transform(c, tau, Zt, 16)
    assert np.abs(A @ Z - Z * w).max() <= 1e-13 * np.abs(w).max()
