"""Index algebra of the bond SVD's Householder reduction (qr_reduce / qr_expand in
kylin-v_amd/csrc/holstein/svd.h), restated on numpy buffers.

A sector matrix M (nr x nc) is stored row-major; rocSOLVER sees that buffer as the
column-major nc x nr matrix M^T.  zgeqrf (nc > nr) leaves R above the diagonal and
reflectors below, zgelqf (nr > nc) leaves L below and reflectors above; zungqr / zunglq then
overwrite the buffer with the explicit orthonormal factor.  The restatement fills the
reflector positions with noise, reads the triangle the way tri_copy_kernel does
(element (i, j) at i * nc + j) and the orthonormal factor row-major, and checks what the
mathematics fixes: M is the product of the two pieces, and the singular triples of the small
triangle carried back through the orthonormal factor are those of M.
"""
import numpy as np
import pytest


def _reduce(M, rng):
    """(T, Qf, colside) as the device holds them after qr_reduce."""
    nr, nc = M.shape
    colside = nc > nr
    cm = M.reshape(nr, nc).T  # the column-major nc x nr view: M^T
    buf = np.empty(nc * nr, dtype=complex)  # buf[j * nc + i] = column-major element (i, j)
    full = np.zeros((nc, nr), dtype=complex)  # the column-major buffer as a matrix
    if colside:  # M^T = Q R: R on and above the diagonal, reflectors below
        Q, R = np.linalg.qr(cm)
        full[:nr, :nr] = R
        refl = np.tril(np.ones((nc, nr), bool), -1)
    else:  # M^T = L Q, from the QR of (M^T)^H: L on and below the diagonal, reflectors above
        Qh, Rh = np.linalg.qr(cm.conj().T)
        Q = Qh.conj().T  # nc x nr, orthonormal rows
        full[:, :nc] = Rh.conj().T
        refl = np.triu(np.ones((nc, nr), bool), 1)
    noise = rng.standard_normal((nc, nr)) + 1j * rng.standard_normal((nc, nr))
    full = np.where(refl, noise, full)  # reflector storage is not part of the factor
    buf[:] = full.T.reshape(-1)  # column-major storage
    m = min(nr, nc)
    # tri_copy_kernel: dst[i][j] = buf[i * nc + j], lower (j <= i) for colside, upper otherwise
    T = np.zeros((m, m), dtype=complex)
    for i in range(m):
        for j in range(m):
            if (j <= i) if colside else (j >= i):
                T[i, j] = buf[i * nc + j]
    # zungqr / zunglq: the buffer holds the explicit factor, column-major nc x nr
    buf[:] = Q.T.reshape(-1)
    Qf = buf.reshape(nr, nc)  # read row-major (nr x nc, ld nc)
    return T, Qf, colside


@pytest.mark.parametrize("shape", [(12, 40), (40, 12), (7, 15), (15, 7)])
def test_householder_reduction_reconstructs_the_sector(shape):
    rng = np.random.default_rng(sum(shape))
    nr, nc = shape
    M = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    T, Qf, colside = _reduce(M, rng)
    m = min(nr, nc)
    if colside:
        assert np.allclose(np.triu(T, 1), 0)
        np.testing.assert_allclose(T @ Qf, M, atol=1e-12)  # M = R^T Q^T
        np.testing.assert_allclose(Qf @ Qf.conj().T, np.eye(m), atol=1e-12)
    else:
        assert np.allclose(np.tril(T, -1), 0)
        np.testing.assert_allclose(Qf @ T, M, atol=1e-12)  # M = Q^T L^T
        np.testing.assert_allclose(Qf.conj().T @ Qf, np.eye(m), atol=1e-12)
    # qr_expand: the small problem's factors carried back give M's singular triples
    u, s, vh = np.linalg.svd(T)
    if colside:
        vh_full = vh @ Qf  # V^H <- V'^H Q^T
        np.testing.assert_allclose((u * s) @ vh_full, M, atol=1e-12)
        np.testing.assert_allclose(vh_full @ vh_full.conj().T, np.eye(m), atol=1e-12)
    else:
        u_full = Qf @ u  # U <- Q^T U'
        np.testing.assert_allclose((u_full * s) @ vh, M, atol=1e-12)
        np.testing.assert_allclose(u_full.conj().T @ u_full, np.eye(m), atol=1e-12)
    np.testing.assert_allclose(s, np.linalg.svd(M, compute_uv=False), rtol=1e-12)


def test_planted_transpose_fails():
    """Reading the triangle column-major instead (a transposed operand) breaks the product."""
    rng = np.random.default_rng(3)
    M = rng.standard_normal((9, 30)) + 1j * rng.standard_normal((9, 30))
    T, Qf, _ = _reduce(M, rng)
    assert not np.allclose(T.T @ Qf, M, atol=1e-6)
