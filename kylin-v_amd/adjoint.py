"""Gradient of a one-site TDVP run with respect to the Holstein couplings.

Forward: one-site TDVP at fixed bond dimension (the sweep of holstein/tdvp.h:78-141: per
site a forward Lanczos step exp(-i tau H1), an SVD that moves the centre, the environment
push and a backward zero-site step exp(+i tau H0), tau = dt/2; mirrored right to left),
then C(T) = <psi0|psi(T)> (holstein/state.h:206-227) and the objective |C(T)|^2.

Backward: reverse-mode over the device tensor algebra of this package.  Every primitive
has a hand-written vector-Jacobian product; cotangents use the convention
df = Re sum conj(zbar) dz, so a complex cotangent has the primal's legs and directions.

  * block-sparse product (N,N) and (C,N): zbar -> abar, bbar are again block-sparse
    products with one operand conjugated (the reference's 'C' rule supplies the conj);
  * transpose: the inverse transpose;
  * H_eff matvec y = H(L, W, R) u (one- and zero-site): L-bar, W-bar and R-bar are the
    four-leg and three-leg contractions of ybar with u along the same chain;
  * Lanczos step y = exp(-i tau H) x: xbar = exp(+i tau H) ybar, and for the operator
    dH the exact Frechet form  df = Re int_0^1 (-i tau) <w(s)| dH |u(s)> ds  with
    u(s) = exp(-i s tau H) x and w(s) = exp(+i (1-s) tau H) ybar, integrated by
    Gauss-Legendre on [0, 1]; each node is one matvec adjoint with ybar = i tau w_k w(s_k);
  * centre move y = Q C (SVD gauge, no truncation at fixed bond dimension): the objective
    is invariant under Q -> Q U, C -> U^H C, which gives
    ybar = (1 - Q Q^H) Qbar C^{-H} + Q Cbar with C^{-H} = S^{-2} C (left-to-right) and
    ybar = C^{-H} Qbar (1 - Q^H Q) + Cbar Q with C^{-H} = C S^{-2} (right-to-left);
    S^2 is the diagonal of C C^H (C^H C).  Bonds need full-rank Schmidt spectra.

Environments are recomputed from per-step checkpoints (the MPS and the right
environments at the start of every step) while the reverse pass walks the steps
backwards.  The couplings enter through the MPO, which is affine in each of them, so
dtheta_p = sum_i Re <Wbar_i, dW_i/dtheta_p>.
"""
from __future__ import annotations

import numpy as np

from . import model as KM

BIG = 1 << 30


def gauss_legendre01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


class DeviceOps:
    """Device tensor algebra the adjoint is written in (the package's C ABI)."""

    def __init__(self, ctx, krylov_tol=1e-12, krylov_maxiter=30):
        from . import Tensor
        import kylin_v_amd as kv  # noqa: F401  (the package this module lives in)
        self.kv = kv
        self.ctx = ctx
        self.Tensor = Tensor
        self.tol, self.maxiter = krylov_tol, krylov_maxiter

    def product(self, a, b, nc, o1="N", o2="N"):
        return self.kv.product(a, b, nc, o1, o2)

    def transpose(self, a, axes):
        return self.kv.transpose(a, axes)

    def copy(self, a):
        return self.kv.transpose(a, list(range(a.rank)))

    def add(self, x, y, alpha=1.0):
        """y + alpha x (new tensor, union of blocks)."""
        return self.kv.axpy(x, y, alpha)

    def scale(self, x, alpha):
        return self.kv.scale_(self.copy(x), alpha)

    def inner(self, a, b):
        """sum conj(a) b."""
        return self.kv.overlap(b, a)

    def norm(self, a):
        return a.norm()

    def expm(self, L, W, R, x, tau, nsite):
        """exp(-i tau H) x for real tau of either sign; norm-preserving, any input norm."""
        nrm = x.norm()
        if nrm == 0.0 or tau == 0.0:
            return self.copy(x)
        c = self.copy(x)
        out, _ = self.kv.local_krylov(L, c, W, None, R, 2.0 * abs(tau), self.tol, self.maxiter,
                                      reverse=tau < 0, nsite=nsite)
        return self.kv.scale_(out, nrm)

    def split(self, y, l2r):
        """(Q, C): y = Q C (l2r, Q left isometry) or y = C Q (r2l, Q right isometry)."""
        if l2r:
            q, c = self.kv.svd_impl(y, 2, -1.0, BIG, True)
        else:
            c, q = self.kv.svd_impl(y, 1, -1.0, BIG, False)
        return q, c

    def inv_sq_svals(self, C, l2r):
        """Diagonal S^-2 on the centre bond, from C C^H (l2r) or C^H C (r2l)."""
        G = self.product(C, C, 1, "N", "C") if l2r else self.product(C, C, 1, "C", "N")
        dirs, blocks = G.to_blocks()
        out = {}
        for q, b in blocks.items():
            d = np.real(np.diag(b))
            if np.any(d <= 0):
                raise ValueError("centre matrix is rank deficient: the one-site TDVP adjoint "
                                 "needs full-rank Schmidt spectra")
            out[q] = np.diag(1.0 / d).astype(np.complex128)
        return self.Tensor.from_blocks(dirs, out, self.ctx)

    def scalar_block(self, t):
        """Value of the single 1x1 block of a closed overlap environment."""
        dirs, blocks = t.to_blocks()
        (q, b), = blocks.items()
        return complex(b.reshape(-1)[0]), dirs, q

    def from_blocks(self, dirs, blocks):
        return self.Tensor.from_blocks(dirs, blocks, self.ctx)

    def boundary_left(self):
        return self.from_blocks((1, -1, -1), {(0, 0, 0): np.ones((1, 1, 1), complex)})

    def boundary_right(self, Q):
        return self.from_blocks((-1, 1, 1), {(Q, 0, Q): np.ones((1, 1, 1), complex)})

    def overlap_boundary(self):
        return self.from_blocks((1, -1), {(0, 0): np.ones((1, 1), complex)})


class _Tape:
    """Reverse-mode record: (outputs, vjp) in execution order.  Cotangents are keyed by
    the id of the primal tensor object; the tape keeps every primal alive."""

    def __init__(self, ops, record=True):
        self.ops, self.record, self.nodes = ops, record, []

    def rec(self, outs, vjp):
        if self.record:
            self.nodes.append((outs, vjp))

    def backward(self, cot):
        ops = self.ops
        for outs, vjp in reversed(self.nodes):
            bars = [cot.pop(id(o), None) for o in outs]
            if all(b is None for b in bars):
                continue
            for var, g in vjp(*bars):
                if g is None or var is None:
                    continue
                k = id(var)
                cot[k] = g if k not in cot else ops.add(g, cot[k])
        self.nodes = []
        return cot


class OneSiteAdjoint:
    """|C(T)|^2 of a one-site TDVP run and its gradient with respect to theta =
    (g_1..g_N, J, omega) of a merged-layout Holstein chain with one phonon mode.

    psi0_blocks: list of {qn: block} site tensors (a right-canonical MPS with full-rank
    Schmidt spectra at the fixed bond dimension; the run keeps those dimensions).
    """

    def __init__(self, N, L, dt, steps, psi0_blocks, ops, nodes=6):
        self.N, self.L, self.dt, self.steps = N, L, dt, steps
        self.ops = ops
        self.sites = None
        self.psi0 = [ops.from_blocks((1, 1, -1), b) for b in psi0_blocks]
        self.Q = max(q[2] for q in psi0_blocks[-1])
        self.gl = gauss_legendre01(nodes)

    # -- model -------------------------------------------------------------------------
    def _terms(self, theta):
        g, J, w = theta[:self.N], theta[self.N], theta[self.N + 1]
        sites, terms, init, ex = KM.chain(self.N, [w], np.asarray(g).reshape(self.N, 1),
                                          [self.L], J, "merged")
        return sites, terms

    def set_params(self, theta):
        theta = np.asarray(theta, dtype=float)
        self.theta = theta
        sites, terms = self._terms(theta)
        self.W = [self.ops.from_blocks((1, 1, -1, -1), b) for b in KM.mpo_blocks(sites, terms)]

    def param_derivatives(self):
        """dW_i/dtheta_p as device tensors: list over sites of {p: tensor}.  The MPO is
        affine in every coupling, so W(e_p) - W(0) with the same channel layout is exact."""
        P = self.N + 2
        sites, terms0 = self._terms(np.zeros(P))
        base = KM.mpo_blocks(sites, terms0)
        out = [dict() for _ in range(self.N)]
        for p in range(P):
            th = np.zeros(P)
            th[p] = 1.0
            _, terms = self._terms(th)
            Wp = KM.mpo_blocks(sites, terms)
            for i in range(self.N):
                d = {}
                for q, b in Wp[i].items():
                    diff = b - base[i].get(q, 0.0)
                    if np.any(diff != 0):
                        d[q] = diff
                if d:
                    out[i][p] = self.ops.from_blocks((1, 1, -1, -1), d)
        return out

    # -- recorded primitives -----------------------------------------------------------
    def _prod(self, tape, a, b, nc, o1="N", o2="N"):
        ops = self.ops
        z = ops.product(a, b, nc, o1, o2)

        def vjp(zb):
            na, nb = a.rank, b.rank
            if (o1, o2) == ("N", "N"):
                return [(a, ops.product(zb, b, nb - nc, "N", "C")),
                        (b, ops.product(a, zb, na - nc, "C", "N"))]
            if (o1, o2) == ("C", "N"):
                return [(a, ops.product(b, zb, nb - nc, "N", "C")),
                        (b, ops.product(a, zb, na - nc, "N", "N"))]
            raise NotImplementedError((o1, o2))

        tape.rec([z], vjp)
        return z

    def _tr(self, tape, a, axes):
        ops = self.ops
        z = ops.transpose(a, axes)
        inv = [0] * len(axes)
        for k, ax in enumerate(axes):
            inv[ax] = k
        tape.rec([z], lambda zb: [(a, ops.transpose(zb, inv))])
        return z

    def _env_left(self, tape, L, A, W):
        t1 = self._prod(tape, L, A, 1)                       # [a',wl,s,b]
        t1t = self._tr(tape, t1, (0, 3, 1, 2))               # [a',b,wl,s]
        Wt = self._tr(tape, W, (1, 2, 0, 3))                 # [wl,s,s',wr]
        t2 = self._prod(tape, t1t, Wt, 2)                    # [a',b,s',wr]
        t2t = self._tr(tape, t2, (0, 2, 3, 1))               # [a',s',wr,b]
        return self._prod(tape, A, t2t, 2, "C", "N")         # [b',wr,b]

    def _env_right(self, tape, R, A, W):
        Rt = self._tr(tape, R, (2, 0, 1))                    # [b,b',wr]
        t1 = self._prod(tape, A, Rt, 1)                      # [a,s,b',wr]
        t1t = self._tr(tape, t1, (0, 2, 1, 3))               # [a,b',s,wr]
        Wt = self._tr(tape, W, (2, 3, 0, 1))                 # [s,wr,s',wl]
        t2 = self._prod(tape, t1t, Wt, 2)                    # [a,b',s',wl]
        t2t = self._tr(tape, t2, (2, 1, 3, 0))               # [s',b',wl,a]
        At = self._tr(tape, A, (1, 2, 0))                    # [s',b',a']
        return self._prod(tape, At, t2t, 2, "C", "N")        # [a',wl,a]

    def _matvec_vjp(self, L, W, R, u, yb, nsite):
        """Cotangents of L, W, R for y = H(L, W, R) u given ybar (u held fixed)."""
        ops = self.ops
        out = []
        Rt = ops.transpose(R, (1, 2, 0))                     # [w(r),b,b']
        if nsite == 1:
            t1 = ops.product(L, u, 1)                        # [a',wl,s,b]
            t1t = ops.transpose(t1, (0, 3, 1, 2))            # [a',b,wl,s]
            Wt = ops.transpose(W, (1, 2, 0, 3))              # [wl,s,s',wr]
            t2 = ops.product(t1t, Wt, 2)                     # [a',b,s',wr]
            t2t = ops.transpose(t2, (0, 2, 3, 1))            # [a',s',wr,b]
            t2tb = ops.product(yb, Rt, 1, "N", "C")          # [a',s',wr,b]
            Rtb = ops.product(t2t, yb, 2, "C", "N")          # [wr,b,b']
            t2b = ops.transpose(t2tb, (0, 3, 1, 2))          # [a',b,s',wr]
            t1tb = ops.product(t2b, Wt, 2, "N", "C")         # [a',b,wl,s]
            Wtb = ops.product(t1t, t2b, 2, "C", "N")         # [wl,s,s',wr]
            t1b = ops.transpose(t1tb, (0, 2, 3, 1))          # [a',wl,s,b]
            out.append((W, ops.transpose(Wtb, (2, 0, 1, 3))))
            out.append((L, ops.product(t1b, u, 2, "N", "C")))
        else:
            t1 = ops.product(L, u, 1)                        # [a',w,b]
            t1b = ops.product(yb, Rt, 1, "N", "C")           # [a',w,b]
            Rtb = ops.product(t1, yb, 1, "C", "N")           # [w,b,b']
            out.append((L, ops.product(t1b, u, 1, "N", "C")))
        out.append((R, ops.transpose(Rtb, (2, 0, 1))))
        return out

    def _expm(self, tape, L, W, R, x, tau, nsite):
        ops = self.ops
        y = ops.expm(L, W, R, x, tau, nsite)

        def vjp(yb):
            res = [(x, ops.expm(L, W, R, yb, -tau, nsite))]
            for s, wgt in zip(*self.gl):
                u = ops.expm(L, W, R, x, tau * s, nsite)
                wv = ops.expm(L, W, R, yb, -tau * (1.0 - s), nsite)
                vb = ops.scale(wv, 1j * tau * wgt)
                res += self._matvec_vjp(L, W, R, u, vb, nsite)
            return res

        tape.rec([y], vjp)
        return y

    def _split(self, tape, y, l2r):
        ops = self.ops
        Q, C = ops.split(y, l2r)

        def vjp(Qb, Cb):
            yb = None
            if Qb is not None:
                Sm2 = ops.inv_sq_svals(C, l2r)
                if l2r:
                    QhQb = ops.product(Q, Qb, 2, "C", "N")           # [k,k']
                    Qperp = ops.add(ops.product(Q, QhQb, 1), Qb, -1.0)
                    yb = ops.product(Qperp, ops.product(Sm2, C, 1), 1)
                else:
                    QbQh = ops.product(Qb, Q, 2, "N", "C")           # [k,k']
                    Qperp = ops.add(ops.product(QbQh, Q, 1), Qb, -1.0)
                    yb = ops.product(ops.product(C, Sm2, 1), Qperp, 1)
            if Cb is not None:
                t = ops.product(Q, Cb, 1) if l2r else ops.product(Cb, Q, 1)
                yb = t if yb is None else ops.add(t, yb)
            return [(y, yb)]

        tape.rec([Q, C], vjp)
        return Q, C

    # -- sweeps ------------------------------------------------------------------------
    def _right_envs(self, tape, mps):
        envR = {self.N: self.ops.boundary_right(self.Q)}
        for i in range(self.N - 1, 0, -1):
            envR[i] = self._env_right(tape, envR[i + 1], mps[i], self.W[i])
        return envR

    def _step(self, tape, mps, envR):
        N, W, tau = self.N, self.W, 0.5 * self.dt
        mps, envR = list(mps), dict(envR)
        envL = {-1: self.Lb}
        for i in range(N):
            y = self._expm(tape, envL[i - 1], W[i], envR[i + 1], mps[i], tau, 1)
            if i < N - 1:
                Q, C = self._split(tape, y, True)
                mps[i] = Q
                envL[i] = self._env_left(tape, envL[i - 1], Q, W[i])
                C2 = self._expm(tape, envL[i], None, envR[i + 1], C, -tau, 0)
                mps[i + 1] = self._prod(tape, C2, mps[i + 1], 1)
            else:
                mps[i] = y
        for i in range(N - 1, -1, -1):
            y = self._expm(tape, envL[i - 1], W[i], envR[i + 1], mps[i], tau, 1)
            if i > 0:
                Q, C = self._split(tape, y, False)
                mps[i] = Q
                envR[i] = self._env_right(tape, envR[i + 1], Q, W[i])
                C2 = self._expm(tape, envL[i - 1], None, envR[i], C, -tau, 0)
                mps[i - 1] = self._prod(tape, mps[i - 1], C2, 1)
            else:
                mps[i] = y
        return mps, envR

    def _autocorr(self, tape, mps):
        E = self.ops.overlap_boundary()
        for i in range(self.N):
            lm = self._prod(tape, E, mps[i], 1)
            E = self._prod(tape, self.psi0[i], lm, 2, "C", "N")
        return E

    # -- public ------------------------------------------------------------------------
    def forward(self, theta, on_step=None):
        """Runs the forward pass, keeps per-step checkpoints; returns C(T)."""
        self.set_params(theta)
        ops = self.ops
        self.Lb = ops.boundary_left()
        self._t0 = _Tape(ops, record=True)
        mps = [ops.copy(t) for t in self.psi0]
        envR = self._right_envs(self._t0, mps)
        self._env0 = envR
        self._ckpt = []
        off = _Tape(ops, record=False)
        for k in range(self.steps):
            self._ckpt.append((mps, envR))
            mps, envR = self._step(off, mps, envR)
            if on_step:
                on_step("forward", k)
        self._final = mps
        self._tO = _Tape(ops, record=True)
        self._E = self._autocorr(self._tO, mps)
        c, _, _ = ops.scalar_block(self._E)
        self.C = c
        return c

    def backward(self, on_step=None):
        """Gradient of |C(T)|^2 with respect to theta (after forward)."""
        ops = self.ops
        c, dirs, q = ops.scalar_block(self._E)
        seed = ops.from_blocks(dirs, {q: np.full((1, 1), 2.0 * c, dtype=complex)})
        cot = self._tO.backward({id(self._E): seed})
        mbar = [cot.get(id(t)) for t in self._final]
        rbar = {}
        Wbar = [None] * self.N

        def take_w(cot):
            for i, w in enumerate(self.W):
                g = cot.pop(id(w), None)
                if g is not None:
                    Wbar[i] = g if Wbar[i] is None else ops.add(g, Wbar[i])

        for k in reversed(range(self.steps)):
            mps_in, envR_in = self._ckpt[k]
            tape = _Tape(ops, record=True)
            mps_out, envR_out = self._step(tape, mps_in, envR_in)
            cot = {}
            for t, g in zip(mps_out, mbar):
                if g is not None:
                    cot[id(t)] = g
            for i, g in rbar.items():
                if g is not None and i in envR_out and i < self.N:
                    cot[id(envR_out[i])] = g
            cot = tape.backward(cot)
            del tape, mps_out, envR_out
            mbar = [cot.get(id(t)) for t in mps_in]
            rbar = {i: cot.get(id(envR_in[i])) for i in envR_in if i < self.N}
            take_w(cot)
            self._ckpt[k] = None
            if on_step:
                on_step("backward", k)
        envR0 = self._env0
        cot = {id(envR0[i]): g for i, g in rbar.items() if g is not None and i in envR0}
        cot = self._t0.backward(cot)
        take_w(cot)
        dW = self.param_derivatives()
        grad = np.zeros(self.N + 2)
        for i in range(self.N):
            if Wbar[i] is None:
                continue
            for p, d in dW[i].items():
                grad[p] += ops.inner(Wbar[i], d).real
        return grad

    def value_and_grad(self, theta, on_step=None):
        c = self.forward(theta, on_step)
        return abs(c) ** 2, self.backward(on_step)


def _autograd_function():
    import torch

    class TdvpAutocorr(torch.autograd.Function):
        """|C(T)|^2 of a one-site TDVP run as a function of theta = (g_1..g_N, J, omega)
        (float64 tensor).  forward runs the device forward pass and keeps the per-step
        checkpoints on the engine; backward runs the hand-written adjoint."""

        @staticmethod
        def forward(ctx, theta, engine):
            c = engine.forward(theta.detach().cpu().numpy().astype(np.float64))
            ctx.engine = engine
            return torch.tensor(abs(c) ** 2, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            grad = torch.as_tensor(ctx.engine.backward(), dtype=torch.float64)
            return gout * grad, None

    return TdvpAutocorr


def __getattr__(name):
    if name == "TdvpAutocorr":
        fn = _autograd_function()
        globals()["TdvpAutocorr"] = fn
        return fn
    raise AttributeError(name)
