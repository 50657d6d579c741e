"""Effective-Hamiltonian matvecs, environment updates, Lanczos exponential, sector SVD
and the TDVP sweep: a plain numpy (complex128) restatement of the reference's hot path.

TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import it; the product path never does.

Reference semantics followed (file:line in YiHeXu517/Kylin-V):
  env_mps_mpo            holstein/product.h:203-319
  env_mps2_mpo1 / mpo2   holstein/product.h:321-477
  apply_double           holstein/product.h:478-486
  env_sweep (3 forms)    holstein/product.h:488-538
  apply_zero / single    holstein/product.h:540-553
  local_krylov           holstein/exponential.h:92-190
  svd_impl               holstein/svd.h:296-440
  TDVP one_site/two_site holstein/tdvp.h:78-227, total_start init :297-316
  SA-1TDVP sweep         holstein/tdvp.h:230-294 (enrichment SVD in oracle/sa.py)
  Operator::join         holstein/operator.h:178-198
  overlap(State, State)  holstein/state.h:206-227
"""
from __future__ import annotations

import numpy as np

from .blocksparse import Tensor, norm, overlap, product, transpose

__all__ = [
    "env_mps_mpo", "env_mps2_mpo1", "env_mps2_mpo2", "apply_double", "apply_single",
    "apply_zero", "env_sweep", "env_sweep_braket", "local_krylov", "KrylovExit",
    "svd_impl", "TDVP", "join", "overlap_states", "herm_eig_small",
]


def _apply_mpo_pairs(lm: Tensor, mpo: Tensor, lm_axes, spec, qn_fn, dirs):
    """Generic MPO fold used by the env_mps*_mpo* family.

    Pairs blocks of ``lm`` and ``mpo`` whose legs ``lm_axes`` equal the mpo legs (1, 2)
    in shape and qn, and sums einsum(spec, lm_block, mpo_block) into the output block
    with qn = qn_fn(lm_qn, mpo_qn).  Duplicate output (shape, qn) blocks are summed, which
    is what the reference's later product<> does with them (product.h:250-275).
    """
    out = Tensor(dirs)
    i0, i1 = lm_axes
    for ql, bl in lm.blocks.items():
        for qw, bw in mpo.blocks.items():
            if (ql[i0] == qw[1] and ql[i1] == qw[2]
                    and bl.shape[i0] == bw.shape[1] and bl.shape[i1] == bw.shape[2]):
                out.push_back_ch(qn_fn(ql, qw), np.einsum(spec, bl, bw, optimize=True))
    return out


def env_mps_mpo(env: Tensor, mps: Tensor, mpo: Tensor) -> Tensor:
    """lmo[a',s',wR,b] = sum_{wL,s} W[s',wL,s,wR] (env.mps)[a',wL,s,b]  (product.h:203-275)."""
    lm = product(env, mps, 1)
    return _apply_mpo_pairs(
        lm, mpo, (1, 2), "awsb,pwsr->aprb",
        lambda ql, qw: (ql[0], qw[0], qw[3], ql[3]),
        (lm.dirs[0], mpo.dirs[0], mpo.dirs[3], lm.dirs[3]))


def env_mps2_mpo1(env: Tensor, mps2: Tensor, mpo: Tensor) -> Tensor:
    """(a',w,s1,s2,b) . W1 -> (a',s1',wM,s2,b)  (product.h:321-397)."""
    lm = product(env, mps2, 1)
    return _apply_mpo_pairs(
        lm, mpo, (1, 2), "awstb,pwsr->aprtb",
        lambda ql, qw: (ql[0], qw[0], qw[3], ql[3], ql[4]),
        (lm.dirs[0], mpo.dirs[0], mpo.dirs[3], lm.dirs[3], lm.dirs[4]))


def env_mps2_mpo2(lm2o1: Tensor, mpo: Tensor) -> Tensor:
    """(a',s1',wM,s2,b) . W2 -> (a',s1',s2',wR,b)  (product.h:399-477)."""
    return _apply_mpo_pairs(
        lm2o1, mpo, (2, 3), "apwtb,qwtr->apqrb",
        lambda ql, qw: (ql[0], ql[1], qw[0], qw[3], ql[4]),
        (lm2o1.dirs[0], lm2o1.dirs[1], mpo.dirs[0], mpo.dirs[3], lm2o1.dirs[4]))


def apply_double(envL, x, mpo1, mpo2, envR) -> Tensor:
    """y = L.W1.W2.x.R for a two-site tensor (product.h:478-486)."""
    lm2o1 = env_mps2_mpo1(envL, x, mpo1)
    lm2o2 = env_mps2_mpo2(lm2o1, mpo2)
    return product(lm2o2, envR, 2, "N", "T")


def apply_single(envL, x, mpo, envR) -> Tensor:
    """y = L.W.x.R for a one-site tensor (product.h:547-553)."""
    lmo = env_mps_mpo(envL, x, mpo)
    return product(lmo, envR, 2, "N", "T")


def apply_zero(envL, c, envR) -> Tensor:
    """y = L.C.R for a bond matrix (product.h:540-545)."""
    lc = product(envL, c, 1)
    return product(lc, envR, 2, "N", "T")


def env_sweep(env, mps, mpo, left2right: bool) -> Tensor:
    """Grow an environment by one site (product.h:488-503).

    L2R: L' = A^dag . (L . A . W).  R2L: the same after transposing A by (2,1,0) and
    W by (0,3,2,1).
    """
    if left2right:
        lmo = env_mps_mpo(env, mps, mpo)
        return product(mps, lmo, 2, "C", "N")
    return env_sweep(env, transpose(mps, (2, 1, 0)), transpose(mpo, (0, 3, 2, 1)), True)


def env_sweep_braket(env, bra, mpo, ket, left2right: bool = True) -> Tensor:
    """Bra != ket form used by Operator::join (product.h:505-521)."""
    if left2right:
        lmo = env_mps_mpo(env, ket, mpo)
        return product(bra, lmo, 2, "C", "N")
    return env_sweep_braket(env, transpose(bra, (2, 1, 0)), transpose(mpo, (0, 3, 2, 1)),
                            transpose(ket, (2, 1, 0)), True)


def herm_eig_small(h: np.ndarray):
    """Eigen-decomposition of the Hermitian matrix whose UPPER triangle is stored in h.

    The reference calls LAPACKE_zheevd('V','U') (util/allocation.h:223-230), which reads
    only the upper triangle and the real part of the diagonal.
    """
    n = h.shape[0]
    a = np.triu(h, 1)
    a = a + a.conj().T + np.diag(np.real(np.diag(h)))
    w, v = np.linalg.eigh(a)
    return w, v


class KrylovExit:
    __slots__ = ("converged", "err", "iters", "loss_norm")

    def __init__(self, converged, err, iters, loss_norm):
        self.converged, self.err, self.iters, self.loss_norm = converged, err, iters, loss_norm

    def __repr__(self):
        return (f"KrylovExit(converged={self.converged}, err={self.err:.3e}, "
                f"iters={self.iters}, loss_norm={self.loss_norm:.3e})")


def local_krylov(envl, x: Tensor, mpo1, mpo2, envr, dt: float, tol: float = 1e-8,
                 max_iter: int = 15, reverse: bool = False, nsite: int | None = None):
    """Lanczos exponential x <- exp(-/+ i (dt/2) H_eff) x/|x|  (exponential.h:92-190).

    Mirrors the reference step by step: the input is normalised in place (:97-98); each
    iteration forms h = (dt/2) H_eff v_k (:113-125); beta = <h|v_{k-1}> with the
    conjugate on h exactly as overlap(v_{k-1}, h) computes it (:126-132); then
    alpha = <h|v_k> after subtracting beta v_{k-1} (:134-137); err *= |h| and the loop
    stops when err < tol (:139-147), else v_{k+1} = h/|h| while k < max_iter-1 (:109,
    :156-159).  T = diag(alpha) + upper beta, zheevd('U'), c = V e^{-/+i Lambda} V^dag e_0
    (:169-183) and x <- sum_i c_i v_i without renormalisation (:184-186).
    Returns (x_new, KrylovExit).
    """
    rank = x.rank if nsite is None else nsite + 2
    nrms = norm(x)
    v0 = x.scaled(1.0 / nrms)
    vm = [v0]
    alp, bet_up = [], []
    k, err = 0, 1.0
    converged = False
    iters = None
    while k < max_iter - 1 and err > tol:
        if rank == 2:
            hv = apply_zero(envl, vm[k], envr)
        elif rank == 3:
            hv = apply_single(envl, vm[k], mpo1, envr)
        else:
            hv = apply_double(envl, vm[k], mpo1, mpo2, envr)
        hv.scale_(dt * 0.5)
        if k > 0:
            betk = overlap(vm[k - 1], hv)
            bet_up.append(betk)
            hv.add_(vm[k - 1], -betk)
        alpk = overlap(vm[k], hv)
        alp.append(alpk)
        hv.add_(vm[k], -alpk)
        nmhv = norm(hv)
        err *= nmhv
        if err < tol:
            converged, iters = True, k + 1
            break
        hv.scale_(1.0 / nmhv)
        vm.append(hv)
        k += 1
    if iters is None:
        iters = len(alp)
    n = len(alp)
    hm = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        hm[i, i] = alp[i]
        if i < n - 1:
            hm[i, i + 1] = bet_up[i]
    w, v = herm_eig_small(hm)
    img = 1j if reverse else -1j
    c = v @ (np.exp(img * w) * np.conj(v[0, :]))
    out = vm[0].scaled(c[0])
    for i in range(1, n):
        out.add_(vm[i], c[i])
    return out, KrylovExit(converged, err, iters, 1.0 - norm(out))


def _sector_layout(A: Tensor, r1: int):
    """Group blocks into row / column charge sectors (fill_block, svd.h:110-205).

    Row sector charge = sum_{l<r1} dir_l qn_l, column = sum_{l>=r1} dir_l qn_l.  Sub-blocks
    are ordered by ascending qn (the reference uses first-appearance order; the order
    only permutes rows/columns inside a sector matrix and changes no result).
    """
    rows, cols = {}, {}
    for q, b in A.sorted_items():
        qr = sum(A.dirs[l] * q[l] for l in range(r1))
        qc = sum(A.dirs[l] * q[l] for l in range(r1, A.rank))
        rsub, csub = q[:r1], q[r1:]
        rs = rows.setdefault(qr, {})
        if rsub not in rs:
            rs[rsub] = b.shape[:r1]
        cs = cols.setdefault(qc, {})
        if csub not in cs:
            cs[csub] = b.shape[r1:]

    def with_pos(d):
        out = {}
        for key, subs in d.items():
            pos, lst = 0, []
            for sq in sorted(subs):
                sh = subs[sq]
                sz = int(np.prod(sh))
                lst.append((sq, sh, pos, sz))
                pos += sz
            out[key] = (lst, pos)
        return out

    return with_pos(rows), with_pos(cols)


def svd_impl(A: Tensor, r1: int, cutoff: float, maxdim: int, left2right: bool,
             left_dir: int = -1, return_svals: bool = False):
    """Sector SVD with global truncation (svd.h:296-440).

    For each row sector qr paired with the column sector qc = -qr: dense matrix, thin
    SVD; S goes into V^dag rows when left2right else into U columns (:330-345).  All
    singular values are sorted descending in one list (:351-353) and, walking from the
    smallest, entry i is dropped while i >= maxdim or s_i <= cutoff, never index 0
    (:361-388).  The new bond leg has dir left_dir on the left tensor and -left_dir on
    the right, qn = -left_dir * (row sector charge) (:262, :285, :406-408).
    Ties in s are broken by (sector order, index), a stable order.
    """
    rows, cols = _sector_layout(A, r1)
    pairs = []  # (qr, qc, U, s, Vh)
    collect = []  # (s, order_key, pair_index)
    for qr in sorted(rows):
        qc = -qr
        if qc not in cols:
            continue
        rlist, nr = rows[qr]
        clist, ncl = cols[qc]
        m = np.zeros((nr, ncl), dtype=np.complex128)
        for (rq, rsh, rpos, rsz) in rlist:
            for (cq, csh, cpos, csz) in clist:
                b = A.blocks.get(tuple(rq) + tuple(cq))
                if b is not None:
                    m[rpos:rpos + rsz, cpos:cpos + csz] = b.reshape(rsz, csz)
        u, s, vh = np.linalg.svd(m, full_matrices=False)
        if left2right:
            vh = s[:, None] * vh
        else:
            u = u * s[None, :]
        pi = len(pairs)
        pairs.append([qr, qc, u, s, vh, len(s)])
        for i, sv in enumerate(s):
            collect.append((float(sv), pi, i))
    order = sorted(range(len(collect)), key=lambda j: (-collect[j][0], collect[j][1], collect[j][2]))
    ranked = [collect[j] for j in order]
    kept = [p[5] for p in pairs]
    for i in range(len(ranked) - 1, 0, -1):
        sv, pi, _ = ranked[i]
        if i >= maxdim or sv <= cutoff:
            kept[pi] -= 1
        else:
            break
    ldirs = A.dirs[:r1] + (left_dir,)
    rdirs = (-left_dir,) + A.dirs[r1:]
    lef, rig = Tensor(ldirs), Tensor(rdirs)
    for pi, (qr, qc, u, s, vh, _) in enumerate(pairs):
        nk = kept[pi]
        if nk <= 0:
            continue
        rlist, _ = rows[qr]
        clist, _ = cols[qc]
        for (rq, rsh, rpos, rsz) in rlist:
            blk = u[rpos:rpos + rsz, :nk].reshape(tuple(rsh) + (nk,))
            lef.push_back_ch(tuple(rq) + (-left_dir * qr,), blk)
        for (cq, csh, cpos, csz) in clist:
            blk = vh[:nk, cpos:cpos + csz].reshape((nk,) + tuple(csh))
            rig.push_back_ch((left_dir * qc,) + tuple(cq), blk)
    if return_svals:
        svals = {pairs[pi][0]: pairs[pi][3][:kept[pi]] for pi in range(len(pairs)) if kept[pi] > 0}
        return lef, rig, svals
    return lef, rig


def boundary_envs(mps):
    """Left boundary block qn (0,0,0) dirs (+1,-1,-1); right boundary (Q,0,Q) dirs
    (-1,+1,+1) with Q the right-leg qn of the last site (tdvp.h:302-311)."""
    envl = Tensor((1, -1, -1))
    envl.push_back_ch((0, 0, 0), np.ones((1, 1, 1), dtype=np.complex128))
    last = mps[-1]
    q = next(iter(last.sorted_items()))[0][2]
    envr = Tensor((-1, 1, 1))
    envr.push_back_ch((q, 0, q), np.ones((1, 1, 1), dtype=np.complex128))
    return envl, envr


class TDVP:
    """Sweep driver (holstein/tdvp.h:34-386).  ``mps`` and ``mpo`` are lists of Tensors."""

    def __init__(self, mps, mpo, cutoff, maxdim, dt, tol_krylov=1e-8, max_iter=15):
        self.mps = [t.copy() for t in mps]
        self.mpo = mpo
        self.cutoff, self.maxdim, self.dt = cutoff, maxdim, dt
        self.tol, self.max_iter = tol_krylov, max_iter
        self.krylov_log = []
        self.kept_log = []
        self.svs = []  # (kept count, SVS record) per two-site SVD (holstein/svd.h:390-400)
        self.svs_values = []  # the kept values behind each record, descending
        envl, envr = boundary_envs(self.mps)
        self.envL = [envl]
        self.envR = [envr]
        for i in range(len(self.mps) - 1, 0, -1):  # tdvp.h:313-316
            self.envR.append(env_sweep(self.envR[-1], self.mps[i], self.mpo[i], False))

    def _kry(self, x, w1, w2, reverse, nsite):
        out, ex = local_krylov(self.envL[-1], x, w1, w2, self.envR[-1], self.dt, self.tol,
                               self.max_iter, reverse, nsite)
        self.krylov_log.append(ex.iters)
        return out

    def two_site(self):
        """One time step dt (tdvp.h:144-227)."""
        ns = len(self.mps)
        self.envR.pop()
        for i in range(ns - 1):
            mm = product(self.mps[i], self.mps[i + 1], 1)
            mm = self._kry(mm, self.mpo[i], self.mpo[i + 1], False, 2)
            cur, nex, sv = svd_impl(mm, 2, self.cutoff, self.maxdim, True, return_svals=True)
            self._log_svs(sv)
            self.kept_log.append(sum(d for d in cur.legs()[2].values()))
            self.mps[i] = cur
            if i < ns - 2:
                self.envL.append(env_sweep(self.envL[-1], self.mps[i], self.mpo[i], True))
                nex = self._kry(nex, self.mpo[i + 1], None, True, 1)
                self.mps[i + 1] = nex
                self.envR.pop()
            else:
                self.mps[i + 1] = nex
        for i in range(ns - 1, 0, -1):
            mm = product(self.mps[i - 1], self.mps[i], 1)
            mm = self._kry(mm, self.mpo[i - 1], self.mpo[i], False, 2)
            nex, cur, sv = svd_impl(mm, 2, self.cutoff, self.maxdim, False, return_svals=True)
            self._log_svs(sv)
            self.kept_log.append(sum(d for d in nex.legs()[2].values()))
            self.mps[i] = cur
            if i > 1:
                self.envR.append(env_sweep(self.envR[-1], self.mps[i], self.mpo[i], False))
                nex = self._kry(nex, self.mpo[i - 1], None, True, 1)
                self.mps[i - 1] = nex
                self.envL.pop()
            else:
                self.mps[i - 1] = nex
        self.envR.append(env_sweep(self.envR[-1], self.mps[1], self.mpo[1], False))

    def _log_svs(self, sv):
        from .sa import svs_record
        kept = sorted((float(x) for v in sv.values() for x in v), reverse=True)
        self.svs.append((len(kept), svs_record(kept)))
        self.svs_values.append(kept)

    def one_site(self, betas=None, seed: int = 0):
        """One time step dt of one-site TDVP (tdvp.h:78-141); with ``betas`` (one per bond,
        the inverse fitted widths) the SA-1TDVP step (tdvp.h:230-294): every SVD is the
        enrichment SVD of bond i (sweeping right) or i-1 (sweeping left), numbered 0, 1, ...
        in sweep order as the ``call`` of its draws under ``seed``."""
        from .sa import sa_svd_impl
        ns = len(self.mps)
        calls = [0]

        def split(t, r1, l2r, bond):
            if betas is None:
                return svd_impl(t, r1, self.cutoff, self.maxdim, l2r)
            lef, rig, _ = sa_svd_impl(t, r1, self.cutoff, self.maxdim, l2r, betas[bond], seed, calls[0])
            calls[0] += 1
            return lef, rig

        for i in range(ns):
            self.mps[i] = self._kry(self.mps[i], self.mpo[i], None, False, 1)
            if i < ns - 1:
                cur, nex = split(self.mps[i], 2, True, i)
                self.mps[i] = cur
                self.envL.append(env_sweep(self.envL[-1], self.mps[i], self.mpo[i], True))
                nex = self._kry(nex, None, None, True, 0)
                self.mps[i + 1] = product(nex, self.mps[i + 1], 1)
                self.envR.pop()
        for i in range(ns - 1, -1, -1):
            self.mps[i] = self._kry(self.mps[i], self.mpo[i], None, False, 1)
            if i > 0:
                nex, cur = split(self.mps[i], 1, False, i - 1)
                self.mps[i] = cur
                self.envR.append(env_sweep(self.envR[-1], self.mps[i], self.mpo[i], False))
                nex = self._kry(nex, None, None, True, 0)
                self.mps[i - 1] = product(self.mps[i - 1], nex, 1)
                self.envL.pop()

    def bond_dims(self):
        return [sum(self.mps[i].legs()[2].values()) for i in range(len(self.mps) - 1)]


def join(mpo, bra, ket) -> complex:
    """<bra|O|ket> by a left-to-right environment pass (operator.h:178-198)."""
    envl = Tensor((1, -1, -1))
    envl.push_back_ch((0, 0, 0), np.ones((1, 1, 1), dtype=np.complex128))
    for i in range(len(mpo)):
        envl = env_sweep_braket(envl, bra[i], mpo[i], ket[i], True)
    if not envl.blocks:
        return 0j
    q = next(iter(envl.sorted_items()))[0][0]
    envr = Tensor((1, 1, 1))
    envr.push_back_ch((q, 0, q), np.ones((1, 1, 1), dtype=np.complex128))
    return overlap(envl, envr)


def overlap_states(s1, s2) -> complex:
    """<s1|s2> (state.h:206-227)."""
    envl = Tensor((1, -1))
    envl.push_back_ch((0, 0), np.ones((1, 1), dtype=np.complex128))
    for i in range(len(s1)):
        lm = product(envl, s2[i], 1)
        envl = product(s1[i], lm, 2, "C", "N")
    if not envl.blocks:
        return 0j
    q = next(iter(envl.sorted_items()))[0][0]
    envr = Tensor((-1, 1))
    envr.push_back_ch((q, q), np.ones((1, 1), dtype=np.complex128))
    return overlap(envl, envr)
