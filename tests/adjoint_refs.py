"""References for the one-site TDVP gradient: the oracle's FP64 forward under central
finite differences, and the exact-dynamics closed form from exact diagonalisation."""
import numpy as np

from oracle import blocksparse as B
from oracle import model as M
from oracle import tdvp as T


def theta_split(theta, N):
    return np.asarray(theta[:N], dtype=float), float(theta[N]), float(theta[N + 1])


def oracle_autocorr(theta, N, L, psi0_blocks, dt, steps):
    """|<psi0|psi(T)>|^2 after `steps` one-site TDVP steps of the oracle (no truncation,
    Krylov tolerance 1e-12)."""
    g, J, w = theta_split(theta, N)
    lat = M.holstein(N, [w], g.reshape(N, 1), [L], J, "merged")
    psi0 = [B.Tensor((1, 1, -1), {tuple(q): np.asarray(b) for q, b in blk.items()})
            for blk in psi0_blocks]
    eng = T.TDVP(psi0, M.build_mpo(lat), cutoff=-1.0, maxdim=1 << 30, dt=dt,
                 tol_krylov=1e-12, max_iter=30)
    for _ in range(steps):
        eng.one_site()
    return abs(T.overlap_states(psi0, eng.mps)) ** 2


def oracle_fd_gradient(theta, N, L, psi0_blocks, dt, steps, h=1e-5):
    grad = np.zeros(len(theta))
    for p in range(len(theta)):
        tp, tm = np.array(theta, float), np.array(theta, float)
        tp[p] += h
        tm[p] -= h
        grad[p] = (oracle_autocorr(tp, N, L, psi0_blocks, dt, steps)
                   - oracle_autocorr(tm, N, L, psi0_blocks, dt, steps)) / (2 * h)
    return grad


def ed_gradient(theta, N, L, v0, T_total):
    """d|C(T)|^2/dtheta for C(T) = <v0|exp(-iHT)|v0> exactly:
    dC/dtheta_p = -i sum_jk conj(a_j) (H_p)_jk a_k F(l_j, l_k) in the eigenbasis of H,
    F(x, y) = int_0^T e^{-i x (T-t)} e^{-i y t} dt, with H affine (here linear) in theta."""
    g, J, w = theta_split(theta, N)
    H = M.ed_hamiltonian(N, [w], g.reshape(N, 1), [L], J)
    lam, V = np.linalg.eigh(H)
    a = V.conj().T @ v0
    C = np.vdot(v0, V @ (np.exp(-1j * lam * T_total) * a))
    dl = lam[:, None] - lam[None, :]
    ej = np.exp(-1j * lam * T_total)
    with np.errstate(divide="ignore", invalid="ignore"):
        F = (ej[None, :] - ej[:, None]) / (1j * dl)  # row j (bra), column k (ket)
    same = np.abs(dl) < 1e-12
    F[same] = (T_total * np.broadcast_to(ej[:, None], F.shape))[same]
    grad = np.zeros(len(theta))
    P = len(theta)
    zero = np.zeros(P)
    for p in range(P):
        e = zero.copy()
        e[p] = 1.0
        ge, Je, we = theta_split(e, N)
        Hp = M.ed_hamiltonian(N, [we], ge.reshape(N, 1), [L], Je)
        Hpe = V.conj().T @ Hp @ V
        dC = -1j * np.sum(np.conj(a)[:, None] * Hpe * a[None, :] * F)
        grad[p] = 2.0 * np.real(np.conj(C) * dC)
    return grad, C
