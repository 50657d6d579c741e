"""Holstein exciton-phonon chain as a U(1) MPO + initial product MPS for the engine.

H = J sum_i (|i><i+1| + h.c.) + sum_{i,m} w_m b+_{im} b_{im} + sum_{i,m} g_{im} n_i (b_{im} + b+_{im})

Charge conventions are the reference's (MPS (a,s,b) dirs (+1,+1,-1), holstein/state.h:35-39;
MPO (s',wL,s,wR) dirs (+1,+1,-1,-1), holstein/operator.h:35-39).  Site layouts:

  "merged"  one site per Holstein site, s = (n_ex, k_1..k_M); charge = n_ex
  "pp"      the reference's projected purification (holstein/local.h:3-86,
            holstein/lattice.h:83-339): exciton site, then per mode a physical and an
            auxiliary site, each basis state its own sector, aux starting full
  "split"   exciton site then one charge-0 site per mode

The MPO is a channel automaton: 'start'/'done' channels plus one open channel per term
crossing the bond, carrying the charge that term has raised so far; channels of equal
charge form one sector of the MPO leg.  This stands in for kln-holstein's MPO sums
(holstein/lattice.h:183-339), which build the same operator by MPO products and SVD
compression.
"""
from __future__ import annotations

import numpy as np


def _boson(L):
    b = np.diag(np.sqrt(np.arange(1.0, L)), 1)
    return b, b.T.copy()


class Site:
    def __init__(self, charges, ops):
        self.q = np.asarray(charges, dtype=np.int64)
        self.ops = {k: np.asarray(v, dtype=np.complex128) for k, v in ops.items()}
        self.d = len(self.q)
        self.sec = {}
        for i, c in enumerate(self.q):
            self.sec.setdefault(int(c), []).append(i)

    def dq(self, name):
        o = self.ops[name]
        ii, jj = np.nonzero(o)
        vals = {int(self.q[i] - self.q[j]) for i, j in zip(ii, jj)}
        if len(vals) > 1:
            raise ValueError(f"operator {name} has no definite charge")
        return vals.pop() if vals else 0


def _exciton():
    cd = np.array([[0, 0], [1, 0]], dtype=float)
    return Site([0, 1], {"I": np.eye(2), "cdag": cd, "c": cd.T, "n": np.diag([0.0, 1.0])})


def chain(N, omegas, gs, Ls, J, layout="merged"):
    """Returns (sites, terms, init, exciton_site_index) for a Holstein chain."""
    omegas, Ls = list(omegas), list(Ls)
    M = len(omegas)
    g = np.asarray(gs, dtype=float)
    g = np.full((N, M), float(g)) if g.ndim == 0 else g.reshape(N, M)
    sites, terms, init, ex = [], [], [], []

    def hop(a, b):
        terms.append((J, [(a, "cdag"), (b, "c")]))
        terms.append((J, [(a, "c"), (b, "cdag")]))

    if layout == "merged":
        P = int(np.prod(Ls))
        e = _exciton()
        ops = {"I": np.eye(2 * P), "cdag": np.kron(e.ops["cdag"], np.eye(P)),
               "c": np.kron(e.ops["c"], np.eye(P)), "n": np.kron(e.ops["n"], np.eye(P))}
        for m, L in enumerate(Ls):
            b, bd = _boson(L)
            left, right = np.eye(int(np.prod(Ls[:m]))), np.eye(int(np.prod(Ls[m + 1:])))
            ops[f"nb{m}"] = np.kron(np.eye(2), np.kron(np.kron(left, bd @ b), right))
            ops[f"nx{m}"] = np.kron(e.ops["n"], np.kron(np.kron(left, b + bd), right))
        st = Site([0] * P + [1] * P, ops)
        for i in range(N):
            sites.append(st)
            ex.append(i)
            init.append(P if i == 0 else 0)
            for m in range(M):
                terms.append((omegas[m], [(i, f"nb{m}")]))
                terms.append((g[i, m], [(i, f"nx{m}")]))
        for i in range(N - 1):
            hop(i, i + 1)
    elif layout in ("split", "pp"):
        e = _exciton()
        stride = 1 + (2 * M if layout == "pp" else M)
        for i in range(N):
            base = stride * i
            sites.append(e)
            ex.append(base)
            init.append(1 if i == 0 else 0)
            for m, L in enumerate(Ls):
                b, bd = _boson(L)
                if layout == "split":
                    sites.append(Site([0] * L, {"I": np.eye(L), "nb": bd @ b, "x": b + bd}))
                    init.append(0)
                    p = base + 1 + m
                    terms.append((omegas[m], [(p, "nb")]))
                    terms.append((g[i, m], [(base, "n"), (p, "x")]))
                else:
                    up = np.diag(np.ones(L - 1), -1)
                    sites.append(Site(list(range(L)), {"I": np.eye(L), "bdag": bd, "b": b,
                                                       "nb": bd @ b}))
                    sites.append(Site(list(range(L)), {"I": np.eye(L), "araise": up,
                                                       "alower": up.T}))
                    init += [0, L - 1]
                    p, a = base + 1 + 2 * m, base + 2 + 2 * m
                    terms.append((omegas[m], [(p, "nb")]))
                    terms.append((g[i, m], [(base, "n"), (p, "bdag"), (a, "alower")]))
                    terms.append((g[i, m], [(base, "n"), (p, "b"), (a, "araise")]))
        for i in range(N - 1):
            hop(stride * i, stride * (i + 1))
    else:
        raise ValueError(f"unknown layout {layout}")
    return sites, terms, init, ex


def mpo_blocks(sites, terms):
    """MPO tensors as lists of {qn: block} dicts with dirs (+1,+1,-1,-1)."""
    ns = len(sites)
    bonds = []
    for k in range(ns + 1):
        ch = []
        if k < ns:
            ch.append(("start", 0))
        if k > 0:
            ch.append(("done", 0))
        for t, (_, ops) in enumerate(terms):
            if len(ops) > 1 and ops[0][0] < k <= ops[-1][0]:
                ch.append((t, sum(sites[s].dq(o) for s, o in ops if s < k)))
        bonds.append(ch)
    out = []
    for k, st in enumerate(sites):
        cl, cr = bonds[k], bonds[k + 1]
        il = {c[0]: j for j, c in enumerate(cl)}
        ir = {c[0]: j for j, c in enumerate(cr)}
        W = np.zeros((st.d, len(cl), st.d, len(cr)), dtype=np.complex128)
        for name in ("start", "done"):
            if name in il and name in ir:
                W[:, il[name], :, ir[name]] += st.ops["I"]
        for t, (coef, ops) in enumerate(terms):
            first, last = ops[0][0], ops[-1][0]
            if not first <= k <= last:
                continue
            o = st.ops[dict(ops).get(k, "I")]
            src = il["start"] if k == first else il[t]
            dst = ir["done"] if k == last else ir[t]
            W[:, src, :, dst] += (coef if k == first else 1.0) * o
        qL = {}
        for j, c in enumerate(cl):
            qL.setdefault(c[1], []).append(j)
        qR = {}
        for j, c in enumerate(cr):
            qR.setdefault(c[1], []).append(j)
        blocks = {}
        for qp, ip in st.sec.items():
            for ql, jl in qL.items():
                for qs, is_ in st.sec.items():
                    for qr, jr in qR.items():
                        blk = W[np.ix_(ip, jl, is_, jr)]
                        if np.any(blk != 0):
                            assert qp + ql == qs + qr
                            blocks[(qp, ql, qs, qr)] = blk
        out.append(blocks)
    return out


def product_state_blocks(sites, init):
    """Product MPS with bond qn = running charge (holstein/lattice.h:83-124)."""
    out, q = [], 0
    for st, idx in zip(sites, init):
        qs = int(st.q[idx])
        blk = np.zeros((1, len(st.sec[qs]), 1), dtype=np.complex128)
        blk[0, st.sec[qs].index(idx), 0] = 1.0
        out.append({(q, qs, q + qs): blk})
        q += qs
    return out


class HolsteinChain:
    """Device-side MPO, initial MPS and observable MPOs of one chain."""

    def __init__(self, N, omegas, gs, Ls, J, layout="merged", ctx=None):
        self._setup(N, J, layout, chain(N, omegas, gs, Ls, J, layout), ctx)

    @classmethod
    def from_deck(cls, deck, exciton=0, ctx=None):
        """The chain of a reference input deck (kln-holstein's -i file, read as
        holstein/input.h:85-155 reads it), in the projected-purification layout the
        reference builds, with the exciton in electronic state `exciton` (0-based).  The
        deck's hopping amplitudes may differ from bond to bond, so hopping() is unavailable."""
        from . import deck as KD
        cfg = KD.read_raw(deck)
        obj = cls.__new__(cls)
        obj._setup(cfg["num_ele"], None, "pp", KD.lattice_from_deck(cfg, exciton), ctx)
        obj.deck = cfg
        return obj

    def _setup(self, N, J, layout, lattice, ctx):
        from . import Tensor, default_context
        self.ctx = ctx or default_context()
        self.N, self.layout = N, layout
        self.sites, self.terms, self.init, self.ex = lattice
        self.J = J
        self._T = Tensor
        self.mpo = self._up(mpo_blocks(self.sites, self.terms), (1, 1, -1, -1))
        self.psi0 = self._up(product_state_blocks(self.sites, self.init), (1, 1, -1))
        self._density = {}
        self._hopping = {}

    def _up(self, blist, dirs):
        return [self._T.from_blocks(dirs, b, self.ctx) for b in blist]

    def density(self, i):
        """n on the exciton of Holstein site i (the reference's Op/Density_j)."""
        if i not in self._density:
            self._density[i] = self._up(mpo_blocks(self.sites, [(1.0, [(self.ex[i], "n")])]),
                                        (1, 1, -1, -1))
        return self._density[i]

    def hopping(self, i):
        """J (|i><i+1| + h.c.): the bond-energy observable."""
        if self.J is None:
            raise ValueError("hopping(i) needs a uniform J (a chain built from a deck has none)")
        if i not in self._hopping:
            a, b = self.ex[i], self.ex[i + 1]
            terms = [(self.J, [(a, "cdag"), (b, "c")]), (self.J, [(a, "c"), (b, "cdag")])]
            self._hopping[i] = self._up(mpo_blocks(self.sites, terms), (1, 1, -1, -1))
        return self._hopping[i]

    def bond_energies(self, mps):
        """E_{i,i+1} = <psi| J (|i><i+1| + h.c.) |psi> for every bond (Operator::join over the
        hopping MPO of each bond, holstein/operator.h:178-198)."""
        from . import join
        return np.array([join(self.hopping(i), mps, mps).real for i in range(self.N - 1)])


def synthetic_state_blocks(sites, D, seed=0, decay=1.0, Q=1):
    """Seeded random right-canonical MPS of total charge Q with every bond filled up to D.

    Used as the benchmark input at the steady-state bond dimension (a run from the product
    state needs hundreds of steps to fill the bonds).  Bond charges q of the left part lie
    in [0, Q]; D is split evenly over the charge sectors and capped by the dimension
    available.  Columns are weighted by exp(-decay * phonon number) so the state sits at
    low phonon occupation like the physical one, then orthonormalised by QR.
    """
    rng = np.random.default_rng(seed)
    ns = len(sites)
    out = [None] * ns
    right = {Q: 1}
    for i in reversed(range(ns)):
        st = sites[i]
        nph = np.zeros(st.d)
        for name, o in st.ops.items():
            if name.startswith("nb"):
                nph += np.real(np.diag(o))
        w = np.exp(-decay * nph)
        cols = {}  # qa -> list of (qs, qb, ncols)
        for qb in sorted(right):
            for qs in sorted(st.sec):
                qa = qb - qs
                if 0 <= qa <= Q:
                    cols.setdefault(qa, []).append((qs, qb))
        if i == 0:
            cols = {0: cols[0]}
            share = {0: 1}
        else:
            share = {}
            per = max(1, D // len(cols))
            for qa, lst in cols.items():
                avail = sum(len(st.sec[qs]) * right[qb] for qs, qb in lst)
                share[qa] = min(per, avail)
        blocks, left = {}, {}
        for qa, lst in cols.items():
            widths = [len(st.sec[qs]) * right[qb] for qs, qb in lst]
            C = sum(widths)
            X = rng.standard_normal((share[qa], C)) + 1j * rng.standard_normal((share[qa], C))
            pos = 0
            for (qs, qb), wd in zip(lst, widths):
                wcol = np.repeat(w[st.sec[qs]], right[qb])
                X[:, pos:pos + wd] *= wcol[None, :]
                pos += wd
            q, _ = np.linalg.qr(X.conj().T)
            rows = q.conj().T  # orthonormal rows
            pos = 0
            for (qs, qb), wd in zip(lst, widths):
                blocks[(qa, qs, qb)] = rows[:, pos:pos + wd].reshape(share[qa], len(st.sec[qs]), right[qb])
                pos += wd
            left[qa] = share[qa]
        out[i] = blocks
        right = left
    return out


def _bond_dims(blocks, axis):
    out = {}
    for q, b in blocks.items():
        out[q[axis]] = b.shape[axis]
    return out


def canonical_blocks(mps, rel_cutoff=1e-13, maxdim=None):
    """Right-canonical form of an MPS given as {qn: block} dicts, with every bond cut to its
    Schmidt rank (singular values below rel_cutoff * the bond's largest are dropped) and to
    at most maxdim values per bond (the largest over all sectors).  A left-to-right SVD
    sweep cuts the bonds, a right-to-left sweep leaves sites 1..N-1 right-isometric and the
    norm on site 0, which is then normalised."""
    mps = [dict(b) for b in mps]
    ns = len(mps)
    for i in range(ns - 1):  # left to right: rows (a, s) grouped by right charge
        rows = {}
        for (qa, qs, qb), b in mps[i].items():
            rows.setdefault(qb, []).append((qa, qs))
        svd = {}
        for qb, keys in rows.items():
            keys.sort()
            mats = [mps[i][k + (qb,)].reshape(-1, mps[i][k + (qb,)].shape[2]) for k in keys]
            svd[qb] = (keys, mats) + tuple(np.linalg.svd(np.vstack(mats), full_matrices=False))
        alls = np.sort(np.concatenate([v[3] for v in svd.values()]))[::-1]
        thr = rel_cutoff * max(alls[0], 1e-300)
        if maxdim is not None and len(alls) > maxdim:
            thr = max(thr, alls[maxdim])
        new_i, carry = {}, {}
        for qb, (keys, mats, u, s, vh) in svd.items():
            k = int(np.sum(s > thr))
            if k == 0:
                continue
            pos = 0
            for key, m in zip(keys, mats):
                sh = mps[i][key + (qb,)].shape
                new_i[key + (qb,)] = u[pos:pos + m.shape[0], :k].reshape(sh[0], sh[1], k)
                pos += m.shape[0]
            carry[qb] = s[:k, None] * vh[:k]
        mps[i] = new_i
        nxt = {}
        for (qa, qs, qb), b in mps[i + 1].items():
            if qa in carry:
                nxt[(qa, qs, qb)] = np.tensordot(carry[qa], b, axes=(1, 0))
        mps[i + 1] = nxt
    for i in range(ns - 1, 0, -1):  # right to left: columns (s, b) grouped by left charge
        cols = {}
        for (qa, qs, qb), b in mps[i].items():
            cols.setdefault(qa, []).append((qs, qb))
        new_i, carry = {}, {}
        for qa, keys in cols.items():
            keys.sort()
            mats = [mps[i][(qa,) + k].reshape(mps[i][(qa,) + k].shape[0], -1) for k in keys]
            u, s, vh = np.linalg.svd(np.hstack(mats), full_matrices=False)
            k = int(np.sum(s > rel_cutoff * max(s[0], 1e-300)))
            pos = 0
            for key, m in zip(keys, mats):
                sh = mps[i][(qa,) + key].shape
                new_i[(qa,) + key] = vh[:k, pos:pos + m.shape[1]].reshape(k, sh[1], sh[2])
                pos += m.shape[1]
            carry[qa] = u[:, :k] * s[None, :k]
        mps[i] = new_i
        prv = {}
        for (qa, qs, qb), b in mps[i - 1].items():
            if qb in carry:
                prv[(qa, qs, qb)] = np.tensordot(b, carry[qb], axes=(2, 0))
        mps[i - 1] = prv
    nrm = np.sqrt(sum(np.vdot(b, b).real for b in mps[0].values()))
    mps[0] = {q: b / nrm for q, b in mps[0].items()}
    return mps


def expanded_state_blocks(sites, init, D, seed=0, eps=0.05, decay=1.0):
    """The product state of `init` expanded to bond dimension D with full-rank Schmidt
    spectra: the direct sum of the product MPS and eps times a seeded random MPS of the same
    charge with bonds up to D, brought to right-canonical form with at most D values kept
    per bond.  Every bond then carries
    the product state's Schmidt value (~1) plus a seeded tail of weight ~eps, which a run at
    fixed bond dimension (one-site TDVP) needs to be well conditioned."""
    P = product_state_blocks(sites, init)
    Q = max(q[2] for q in P[-1])
    R = synthetic_state_blocks(sites, D, seed=seed, decay=decay, Q=Q)
    ns = len(sites)
    out = []
    for i in range(ns):
        pl, pr = _bond_dims(P[i], 0), _bond_dims(P[i], 2)
        rl, rr = _bond_dims(R[i], 0), _bond_dims(R[i], 2)
        first, last = i == 0, i == ns - 1
        blocks = {}
        keys = set(P[i]) | set(R[i])
        for (qa, qs, qb) in keys:
            ds = len(sites[i].sec[qs])
            da = 1 if first else pl.get(qa, 0) + rl.get(qa, 0)
            db = 1 if last else pr.get(qb, 0) + rr.get(qb, 0)
            blk = np.zeros((da, ds, db), dtype=np.complex128)
            p, r = P[i].get((qa, qs, qb)), R[i].get((qa, qs, qb))
            if first:
                if p is not None:
                    blk[:, :, :p.shape[2]] = p
                if r is not None:
                    blk[:, :, pr.get(qb, 0):] = eps * r
            elif last:
                if p is not None:
                    blk[:p.shape[0]] = p
                if r is not None:
                    blk[pl.get(qa, 0):] = r
            else:
                if p is not None:
                    blk[:p.shape[0], :, :p.shape[2]] = p
                if r is not None:
                    blk[pl.get(qa, 0):, :, pr.get(qb, 0):] = r
            blocks[(qa, qs, qb)] = blk
        out.append(blocks)
    return canonical_blocks(out, maxdim=D)


def centre_bond(mps, mpo, mid, Q=1):
    """Environments and site tensors of a right-canonical device MPS with its centre moved
    to the bond (mid, mid+1), as a two-site sweep reaches that bond
    (holstein/tdvp.h:157-175): sites 0..mid-1 left-canonical by exact SVDs
    (svd_impl with no truncation, holstein/svd.h:296-440), the left environment L_mid
    built over them, and the right environments over sites mid+2.. and mid+1..
    Returns (envL, A_mid, A_mid1, envR2, envR1) as device tensors; envR2 is the right
    environment the two-site H_eff at (mid, mid+1) uses, envR1 the one the backward
    one-site step at site mid+1 uses."""
    from . import Tensor, env_sweep, product, svd_impl
    ctx = mps[0].ctx
    A = list(mps)
    envL = Tensor.from_blocks((1, -1, -1), {(0, 0, 0): np.ones((1, 1, 1))}, ctx)
    for i in range(mid):
        cur, nex = svd_impl(A[i], 2, -1.0, 10 ** 6, True)
        A[i] = cur
        A[i + 1] = product(nex, A[i + 1], 1)
        envL = env_sweep(envL, A[i], mpo[i], True)
    envR2 = Tensor.from_blocks((-1, 1, 1), {(Q, 0, Q): np.ones((1, 1, 1))}, ctx)
    for i in range(len(A) - 1, mid + 1, -1):
        envR2 = env_sweep(envR2, A[i], mpo[i], False)
    envR1 = env_sweep(envR2, A[mid + 1], mpo[mid + 1], False)
    return envL, A[mid], A[mid + 1], envR2, envR1
