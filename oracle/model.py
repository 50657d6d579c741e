"""Holstein exciton-phonon chains as U(1) MPO/MPS in three site layouts, plus exact
diagonalisation for tiny chains.

TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import it.

H = J sum_i (|i><i+1| + h.c.) + sum_{i,m} w_m b+_{im} b_{im}
    + sum_{i,m} g_{im} n_i (b_{im} + b+_{im})               (no zero-point term)

Layouts (charge conventions of the reference: MPS legs (a,s,b) dirs (+1,+1,-1) so
q_b = q_a + q_s, holstein/state.h:35-39; MPO legs (s',wL,s,wR) dirs (+1,+1,-1,-1) so
s'+wL = s+wR, holstein/operator.h:35-39, holstein/local.h:47):

* "pp"     projected purification as in the reference (holstein/local.h:3-86,
           holstein/lattice.h:83-339, holstein/input.h:104-151): exciton site (dim 2),
           then per mode a physical site and an auxiliary site (dim L each).  Every basis
           state is its own 1-dim sector; charge = N_ex + sum(n_phys + n_aux); b+ is
           b+_phys (x) aux-lower with value 1 on the aux (local.h:65-86); the aux starts
           full (lattice.h:113-118).  Coupling g n (b_phys a+_aux + b+_phys a_aux)
           (lattice.h:254-266 with val = sqrt(2) g and the sqrt(0.5) prefactor).
* "merged" one MPS site per Holstein site, s = (n_ex, k_1..k_M); charge = n_ex, so two
           sectors of dim prod(L_m).
* "split"  exciton site (sectors n=0,1 of dim 1) then one site per mode (single
           charge-0 sector of dim L); charge = n_ex.

The MPO is built by a finite-state automaton: channels 'start' and 'done' (charge 0)
plus one open channel per term spanning the bond, whose charge is the charge already
raised by that term.  Channels of equal charge form one MPO-leg sector.
"""
from __future__ import annotations

import numpy as np

from .blocksparse import Tensor

__all__ = ["SiteType", "Lattice", "holstein", "build_mpo", "product_state",
           "density_mpo", "hopping_mpo", "identity_mpo", "mpo_to_dense",
           "mps_to_dense", "ed_hamiltonian", "ed_initial_state", "chain8_params", "chain32_params"]


class SiteType:
    """Local Hilbert space with charge sectors and named dense operators."""

    def __init__(self, name, charges, ops):
        self.name = name
        self.charges = np.asarray(charges, dtype=np.int64)  # charge of each basis state
        self.ops = ops  # name -> (d x d) ndarray
        self.dim = len(self.charges)
        # sectors: ascending charge; basis states of a sector keep their order
        self.sectors = {}
        for idx, q in enumerate(self.charges):
            self.sectors.setdefault(int(q), []).append(idx)

    def op(self, name):
        return self.ops[name]

    def op_charge(self, name):
        o = self.ops[name]
        nz = np.argwhere(np.abs(o) > 0)
        dq = {int(self.charges[i] - self.charges[j]) for i, j in nz}
        if len(dq) > 1:
            raise ValueError(f"operator {name} does not have a definite charge")
        return dq.pop() if dq else 0


def _boson_ops(L):
    b = np.diag(np.sqrt(np.arange(1, L, dtype=np.float64)), 1)
    return b, b.T.copy()


def exciton_site():
    cdag = np.array([[0, 0], [1, 0]], dtype=np.complex128)
    c = cdag.T.copy()
    n = np.diag([0.0, 1.0]).astype(np.complex128)
    return SiteType("e", [0, 1], {"I": np.eye(2, dtype=np.complex128), "cdag": cdag, "c": c, "n": n})


def phonon_site(L):
    """Phonon mode with charge 0 on every level (split layout)."""
    b, bd = _boson_ops(L)
    return SiteType("p", [0] * L, {"I": np.eye(L, dtype=np.complex128),
                                   "nb": (bd @ b).astype(np.complex128),
                                   "x": (b + bd).astype(np.complex128)})


def pp_phys_site(L):
    """Physical phonon site of the projected purification: level k has charge k."""
    b, bd = _boson_ops(L)
    return SiteType("pp", list(range(L)), {"I": np.eye(L, dtype=np.complex128),
                                           "bdag": bd.astype(np.complex128),
                                           "b": b.astype(np.complex128),
                                           "nb": (bd @ b).astype(np.complex128)})


def pp_aux_site(L):
    """Auxiliary site: level k has charge k; raise/lower carry value 1 (local.h:73-86)."""
    up = np.diag(np.ones(L - 1), -1).astype(np.complex128)
    return SiteType("pa", list(range(L)), {"I": np.eye(L, dtype=np.complex128),
                                           "araise": up, "alower": up.T.copy()})


def merged_site(Ls):
    """Exciton (x) phonon modes; basis index = n * prod(L) + (k_1, ..., k_M) row-major."""
    P = int(np.prod(Ls))
    ex = exciton_site()
    I_p = np.eye(P, dtype=np.complex128)
    ops = {"I": np.eye(2 * P, dtype=np.complex128),
           "cdag": np.kron(ex.op("cdag"), I_p), "c": np.kron(ex.op("c"), I_p),
           "n": np.kron(ex.op("n"), I_p)}
    for m, L in enumerate(Ls):
        b, bd = _boson_ops(L)
        left = np.eye(int(np.prod(Ls[:m])))
        right = np.eye(int(np.prod(Ls[m + 1:])))
        nb = np.kron(np.kron(left, bd @ b), right)
        x = np.kron(np.kron(left, b + bd), right)
        ops[f"nb{m}"] = np.kron(np.eye(2), nb).astype(np.complex128)
        ops[f"nx{m}"] = np.kron(ex.op("n"), x).astype(np.complex128)
    return SiteType("m", [0] * P + [1] * P, ops)


class Lattice:
    """A chain of SiteTypes plus a term list and the initial product state."""

    def __init__(self, sites, terms, init, layout, exciton_sites, meta):
        self.sites = sites  # list[SiteType]
        self.terms = terms  # list of (coef, [(site, opname), ...]) with ascending sites
        self.init = init  # basis index per site
        self.layout = layout
        self.exciton_sites = exciton_sites  # MPS site holding exciton number of Holstein site i
        self.meta = meta


def holstein(N, omegas, gs, Ls, J, layout="merged"):
    """Build a Holstein chain.

    omegas, Ls: per-mode lists (same for all sites); gs: array (N, M) of couplings.
    """
    omegas = list(omegas)
    Ls = list(Ls)
    M = len(omegas)
    gs = np.asarray(gs, dtype=np.float64)
    gs = np.full((N, M), float(gs)) if gs.ndim == 0 else gs.reshape(N, M)
    sites, terms, init, exs = [], [], [], []
    if layout == "merged":
        st = merged_site(Ls)
        for i in range(N):
            sites.append(st)
            exs.append(i)
            init.append(st.dim // 2 if i == 0 else 0)
            for m in range(M):
                terms.append((omegas[m], [(i, f"nb{m}")]))
                terms.append((gs[i, m], [(i, f"nx{m}")]))
        for i in range(N - 1):
            terms.append((J, [(i, "cdag"), (i + 1, "c")]))
            terms.append((J, [(i, "c"), (i + 1, "cdag")]))
    elif layout == "split":
        ex = exciton_site()
        phs = [phonon_site(L) for L in Ls]
        stride = 1 + M
        for i in range(N):
            sites.append(ex)
            exs.append(stride * i)
            init.append(1 if i == 0 else 0)
            for m in range(M):
                sites.append(phs[m])
                init.append(0)
                terms.append((omegas[m], [(stride * i + 1 + m, "nb")]))
                terms.append((gs[i, m], [(stride * i, "n"), (stride * i + 1 + m, "x")]))
        for i in range(N - 1):
            terms.append((J, [(stride * i, "cdag"), (stride * (i + 1), "c")]))
            terms.append((J, [(stride * i, "c"), (stride * (i + 1), "cdag")]))
    elif layout == "pp":
        ex = exciton_site()
        ph = [pp_phys_site(L) for L in Ls]
        ax = [pp_aux_site(L) for L in Ls]
        stride = 1 + 2 * M
        for i in range(N):
            base = stride * i
            sites.append(ex)
            exs.append(base)
            init.append(1 if i == 0 else 0)
            for m in range(M):
                sites.append(ph[m])
                sites.append(ax[m])
                init.append(0)
                init.append(Ls[m] - 1)
                p, a = base + 1 + 2 * m, base + 2 + 2 * m
                terms.append((omegas[m], [(p, "nb")]))
                terms.append((gs[i, m], [(base, "n"), (p, "bdag"), (a, "alower")]))
                terms.append((gs[i, m], [(base, "n"), (p, "b"), (a, "araise")]))
        for i in range(N - 1):
            terms.append((J, [(stride * i, "cdag"), (stride * (i + 1), "c")]))
            terms.append((J, [(stride * i, "c"), (stride * (i + 1), "cdag")]))
    else:
        raise ValueError(f"unknown layout {layout}")
    return Lattice(sites, terms, init, layout, exs,
                   dict(N=N, omegas=omegas, gs=np.array(gs), Ls=Ls, J=J))


def _channels(lat: Lattice):
    """Channel list per bond: bond k sits between site k-1 and k (k = 0..ns)."""
    ns = len(lat.sites)
    bonds = []
    for k in range(ns + 1):
        ch = []
        if k < ns:
            ch.append(("start", 0))
        if k > 0:
            ch.append(("done", 0))
        for t, (coef, ops) in enumerate(lat.terms):
            first, last = ops[0][0], ops[-1][0]
            if len(ops) > 1 and first < k <= last:
                q = sum(lat.sites[s].op_charge(o) for s, o in ops if s < k)
                ch.append((("t", t), q))
        bonds.append(ch)
    return bonds


def build_mpo(lat: Lattice, terms=None) -> list:
    """MPO tensors (s', wL, s, wR) from the term list by the channel automaton."""
    if terms is not None:
        lat = Lattice(lat.sites, terms, lat.init, lat.layout, lat.exciton_sites, lat.meta)
    ns = len(lat.sites)
    bonds = _channels(lat)
    mpo = []
    for k in range(ns):
        st = lat.sites[k]
        chL, chR = bonds[k], bonds[k + 1]
        idxL = {c[0]: j for j, c in enumerate(chL)}
        idxR = {c[0]: j for j, c in enumerate(chR)}
        d = st.dim
        W = np.zeros((d, len(chL), d, len(chR)), dtype=np.complex128)
        if "start" in idxL and "start" in idxR:
            W[:, idxL["start"], :, idxR["start"]] += st.op("I")
        if "done" in idxL and "done" in idxR:
            W[:, idxL["done"], :, idxR["done"]] += st.op("I")
        for t, (coef, ops) in enumerate(lat.terms):
            sites_ops = dict(ops)
            first, last = ops[0][0], ops[-1][0]
            if not (first <= k <= last):
                continue
            o = st.op(sites_ops.get(k, "I"))
            if first == last == k:
                W[:, idxL["start"], :, idxR["done"]] += coef * o
            elif k == first:
                W[:, idxL["start"], :, idxR[("t", t)]] += coef * o
            elif k == last:
                W[:, idxL[("t", t)], :, idxR["done"]] += o
            else:
                W[:, idxL[("t", t)], :, idxR[("t", t)]] += o
        mpo.append(_dense_to_blocks4(W, st, [c[1] for c in chL], [c[1] for c in chR]))
    return mpo


def _leg_sectors(charges):
    sec = {}
    for idx, q in enumerate(charges):
        sec.setdefault(int(q), []).append(idx)
    return sec


def _dense_to_blocks4(W, st: SiteType, qL, qR):
    """Cut a dense (s', wL, s, wR) array into charge blocks, dropping all-zero blocks."""
    t = Tensor((1, 1, -1, -1))
    ss = st.sectors
    sL, sR = _leg_sectors(qL), _leg_sectors(qR)
    for qp, ip in ss.items():
        for ql, il in sL.items():
            for qs, is_ in ss.items():
                for qr, ir in sR.items():
                    blk = W[np.ix_(ip, il, is_, ir)]
                    if np.any(blk != 0):
                        if qp + ql != qs + qr:
                            raise ValueError("MPO entry breaks charge conservation")
                        t.push_back_ch((qp, ql, qs, qr), blk)
    return t


def product_state(lat: Lattice) -> list:
    """Product MPS of lat.init with bond qn = running charge (lattice.h:83-124)."""
    mps, q = [], 0
    for st, idx in zip(lat.sites, lat.init):
        qs = int(st.charges[idx])
        t = Tensor((1, 1, -1))
        blk = np.zeros((1, len(st.sectors[qs]), 1), dtype=np.complex128)
        blk[0, st.sectors[qs].index(idx), 0] = 1.0
        t.push_back_ch((q, qs, q + qs), blk)
        mps.append(t)
        q += qs
    return mps


def density_mpo(lat: Lattice, i: int) -> list:
    """n on the exciton of Holstein site i, identity elsewhere."""
    return build_mpo(lat, [(1.0, [(lat.exciton_sites[i], "n")])])


def hopping_mpo(lat: Lattice, i: int) -> list:
    """J (|i><i+1| + h.c.) for one bond (the bond-energy observable)."""
    J = lat.meta["J"]
    a, b = lat.exciton_sites[i], lat.exciton_sites[i + 1]
    return build_mpo(lat, [(J, [(a, "cdag"), (b, "c")]), (J, [(a, "c"), (b, "cdag")])])


def identity_mpo(lat: Lattice) -> list:
    ns = len(lat.sites)
    return build_mpo(lat, [(1.0, [(0, "I")])]) if ns else []


def _dense_mpo_site(t: Tensor, st: SiteType):
    legs = t.legs()
    out_w = []
    for ax in (1, 3):
        offs, pos = {}, 0
        for q in sorted(legs[ax]):
            offs[q] = pos
            pos += legs[ax][q]
        out_w.append((offs, pos))
    d = st.dim
    W = np.zeros((d, out_w[0][1], d, out_w[1][1]), dtype=np.complex128)
    for q, b in t.blocks.items():
        ip, is_ = st.sectors[q[0]], st.sectors[q[2]]
        ol, orr = out_w[0][0][q[1]], out_w[1][0][q[3]]
        W[np.ix_(ip, range(ol, ol + b.shape[1]), is_, range(orr, orr + b.shape[3]))] += b
    return W


def mpo_to_dense(lat: Lattice, mpo) -> np.ndarray:
    """Full operator matrix (small chains only), basis = kron of site bases."""
    W0 = _dense_mpo_site(mpo[0], lat.sites[0])
    M = W0[:, 0, :, :]  # (s', s, wR)
    for k in range(1, len(mpo)):
        W = _dense_mpo_site(mpo[k], lat.sites[k])
        M = np.einsum("abw,cwdr->acbdr", M, W)
        da, dc = M.shape[0], M.shape[1]
        M = M.reshape(da * dc, M.shape[2] * M.shape[3], M.shape[4])
    return M[:, :, 0]


def mps_to_dense(lat: Lattice, mps) -> np.ndarray:
    """State vector in the kron basis (small chains only)."""
    v = None
    for k, t in enumerate(mps):
        st = lat.sites[k]
        legs = t.legs()
        offs = []
        for ax in (0, 2):
            o, pos = {}, 0
            for q in sorted(legs[ax]):
                o[q] = pos
                pos += legs[ax][q]
            offs.append((o, pos))
        A = np.zeros((offs[0][1], st.dim, offs[1][1]), dtype=np.complex128)
        for q, b in t.blocks.items():
            a0, b0 = offs[0][0][q[0]], offs[1][0][q[2]]
            A[a0:a0 + b.shape[0], st.sectors[q[1]], b0:b0 + b.shape[2]] += b
        v = A[0] if v is None else np.einsum("pa,asb->psb", v, A).reshape(-1, A.shape[2])
    return v[:, 0]


def ed_hamiltonian(N, omegas, gs, Ls, J):
    """Dense H in the merged kron basis (site i basis = n * prod(L) + phonon index)."""
    lat = holstein(N, omegas, gs, Ls, J, "merged")
    st = lat.sites[0]
    d = st.dim
    dim = d ** N
    H = np.zeros((dim, dim), dtype=np.complex128)

    def embed(ops):
        m = np.eye(1, dtype=np.complex128)
        for k in range(N):
            m = np.kron(m, ops.get(k, np.eye(d)))
        return m

    for coef, ops in lat.terms:
        H += coef * embed({s: st.op(o) for s, o in ops})
    return H


def ed_initial_state(N, Ls):
    P = int(np.prod(Ls))
    d = 2 * P
    v = np.zeros(d ** N, dtype=np.complex128)
    idx = 0
    for k in range(N):
        idx = idx * d + (P if k == 0 else 0)
    v[idx] = 1.0
    return v


def chain8_params():
    """8 sites, exciton (x) one mode at 8 levels, J = w = g = 1."""
    return dict(N=8, omegas=[1.0], gs=np.ones((8, 1)), Ls=[8], J=1.0)


def chain32_params(seed=2024):
    """32 sites, 16 levels, J=1, w=0.5, g_i ~ U[0.8, 1.2] from numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.8, 1.2, size=32)
    return dict(N=32, omegas=[0.5], gs=g.reshape(32, 1), Ls=[16], J=1.0)
