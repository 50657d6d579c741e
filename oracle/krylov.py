"""Oracle (test infrastructure only): numpy restatement of the reference's global Krylov
stage and the MPS-level operations it uses.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU baseline import this package; the product never does.

References (YiHeXu517/Kylin-V):
  stack (direct sum along legs)        holstein/tensor.h:262-319, holstein/block.h:181-232
  State += / * / norm / canonicalize   holstein/state.h:52-80, :166-185, :203
  apply_operator (zip-up MPO x MPS)    holstein/operator.h:231-272
  apply_ortho_site, env helpers        holstein/krylov.h:12-125
  apply_ortho (one- and two-vector)    holstein/krylov.h:126-205
  global_main                          holstein/krylov.h:206-284
  Krylov stage driver (H scaled by dt) holstein/tdvp.h:318-327
"""
from __future__ import annotations

import numpy as np

from .blocksparse import Tensor, norm, overlap, product, transpose
from .tdvp import apply_double, env_mps_mpo, env_sweep_braket, herm_eig_small, join, svd_impl


def stack(t1: Tensor, t2: Tensor, axes) -> Tensor:
    """Direct sum of t1 and t2 along `axes` (tensor.h:262-319): on a stacked leg the sector
    of charge q gets dim1(q) + dim2(q), t1's blocks sit at offset 0 and t2's after t1's
    part of that sector; other legs must agree."""
    axes = list(axes)
    dim1 = [dict() for _ in range(t1.rank)]
    dim2 = [dict() for _ in range(t1.rank)]
    for q, b in t1.blocks.items():
        for a in range(t1.rank):
            dim1[a][q[a]] = b.shape[a]
    for q, b in t2.blocks.items():
        for a in range(t2.rank):
            dim2[a][q[a]] = b.shape[a]
    tot = [dict() for _ in range(t1.rank)]
    for a in axes:
        for q in set(dim1[a]) | set(dim2[a]):
            tot[a][q] = dim1[a].get(q, 0) + dim2[a].get(q, 0)
    res = Tensor(t1.dirs)

    def place(q, b, second):
        shape = list(b.shape)
        pos = [0] * len(shape)
        for a in axes:
            shape[a] = tot[a][q[a]]
            if second:
                pos[a] = dim1[a].get(q[a], 0)
        out = res.blocks.get(q)
        if out is None:
            out = np.zeros(shape, dtype=np.complex128)
            res.blocks[q] = out
        sl = tuple(slice(p, p + n) for p, n in zip(pos, b.shape))
        out[sl] += b

    for q, b in t1.blocks.items():
        place(q, b, False)
    for q, b in t2.blocks.items():
        place(q, b, True)
    return res


def state_add(s1, s2):
    """s1 += s2 (state.h:52-63): first site stacked on its right bond, middle sites on both
    bonds, last site on its left bond."""
    ns = len(s1)
    out = [stack(s1[0], s2[0], [2])]
    for i in range(1, ns - 1):
        out.append(stack(s1[i], s2[i], [0, 2]))
    out.append(stack(s1[ns - 1], s2[ns - 1], [0]))
    return out


def state_scale(s, c):
    """State * c scales the first site (state.h:66-80)."""
    return [s[0].scaled(c)] + [t.copy() for t in s[1:]]


def state_norm(s) -> float:
    """norm(State) is the norm of the first site (state.h:203); the state is kept with its
    orthogonality centre there."""
    return norm(s[0])


def canonicalize(s, tol=0.0, maxdim=1000):
    """state.h:166-185: left-to-right svd_impl<2,1>, then right-to-left svd_impl<1,2>, both
    truncating with (tol, maxdim); the result is right-canonical with the centre on site 0."""
    s = [t.copy() for t in s]
    ns = len(s)
    for i in range(ns - 1):
        cur, nex = svd_impl(s[i], 2, tol, maxdim, True)
        s[i] = cur
        s[i + 1] = product(nex, s[i + 1], 1)
    for i in range(ns - 1, 0, -1):
        nex, cur = svd_impl(s[i], 1, tol, maxdim, False)
        s[i] = cur
        s[i - 1] = product(s[i - 1], nex, 1)
    return s


def apply_operator(op, wf, tol, maxdim):
    """Zip-up MPO x MPS (operator.h:231-272): env_mps_mpo with the running left factor,
    svd_impl<2,2> left-to-right (the right factor becomes the next left env), the last
    site closed with the (lastB, lastw, lastB) boundary, then a right-to-left svd_impl<1,2>
    sweep."""
    ns = len(op)
    envl = Tensor((1, -1, -1))
    envl.push_back_ch((0, 0, 0), np.ones((1, 1, 1), dtype=np.complex128))
    last_b = next(iter(wf[ns - 1].blocks))[2]
    last_w = next(iter(op[ns - 1].blocks))[3]
    res = [None] * ns
    for i in range(ns):
        lmo = env_mps_mpo(envl, wf[i], op[i])
        if i == ns - 1:
            envr = Tensor((-1, 1, 1))
            envr.push_back_ch((last_b, last_w, last_b), np.ones((1, 1, 1), dtype=np.complex128))
            res[i] = product(lmo, envr, 2, "N", "T")
        else:
            cur, nex = svd_impl(lmo, 2, tol, maxdim, True)
            res[i] = cur
            envl = nex
    for i in range(ns - 1, 0, -1):
        nex, cur = svd_impl(res[i], 1, tol, maxdim, False)
        res[i] = cur
        res[i - 1] = product(res[i - 1], nex, 1)
    return res


def apply_ortho_site(Nh, La, Lb):
    """krylov.h:12-45: Nh minus its components along La and Lb, solving the 2x2 Gram system
    (only La when the system is singular, nothing when La vanishes too)."""
    LNa, LNb = overlap(La, Nh), overlap(Lb, Nh)
    aa, bb = norm(La) ** 2, norm(Lb) ** 2
    ab = overlap(La, Lb)
    ba = np.conj(ab)
    cc = ab * ba - aa * bb
    res = Nh.copy()
    if abs(cc) < 1e-17:
        if abs(aa) < 1e-17:
            return res
        res.add_(La, -1.0 * LNa / aa)
    else:
        res.add_(La, (bb * LNa - ab * LNb) / cc)
        res.add_(Lb, (aa * LNb - ba * LNa) / cc)
    return res


def _env_state_right(envr, s1, s2):
    s1t, s2t = transpose(s1, (2, 1, 0)), transpose(s2, (2, 1, 0))
    lm = product(envr, s2t, 1)
    return product(s1t, lm, 2, "C", "N")


def _env_state_left(envl, s1, s2):
    lm = product(envl, s2, 1)
    return product(s1, lm, 2, "C", "N")


def _init_envs(H, s1, s2):
    """init_env_op / init_env_state (krylov.h:47-88): left boundaries, right stacks down to
    site 2 (the two-site sweep starts at bond 0-1)."""
    ns = len(s1)
    tot = next(iter(s1[ns - 1].blocks))[2]
    L3 = Tensor((1, -1, -1))
    L3.push_back_ch((0, 0, 0), np.ones((1, 1, 1), dtype=np.complex128))
    R3 = Tensor((-1, 1, 1))
    R3.push_back_ch((tot, 0, tot), np.ones((1, 1, 1), dtype=np.complex128))
    L2 = Tensor((1, -1))
    L2.push_back_ch((0, 0), np.ones((1, 1), dtype=np.complex128))
    R2 = Tensor((-1, 1))
    R2.push_back_ch((tot, tot), np.ones((1, 1), dtype=np.complex128))
    EL3, ER3 = [L3], [R3]
    if H is not None:
        for i in range(ns - 1, 1, -1):
            ER3.append(env_sweep_braket(ER3[-1], s1[i], H[i], s2[i], False))
    EL2, ER2 = [L2], [R2]
    for i in range(ns - 1, 1, -1):
        ER2.append(_env_state_right(ER2[-1], s1[i], s2[i]))
    return EL3, ER3, EL2, ER2


def apply_double_state(envl, M, envr):
    """krylov.h:117-122: the two-site tensor M between overlap environments."""
    lm = product(envl, M, 1)
    return product(lm, envr, 1, "N", "T")


def apply_ortho1(H, Va, tol, maxdim):
    """First Krylov vector (krylov.h:126-134): H Va by zip-up, minus its Va component, then
    canonicalised without truncation (canonicalize() defaults: tol 0, maxdim 1000)."""
    gss = apply_operator(H, Va, tol, maxdim)
    ovp = overlap_states_full(gss, Va)
    gss = state_add(gss, state_scale(Va, -ovp))
    return canonicalize(gss)


def overlap_states_full(s1, s2) -> complex:
    """overlap(State, State) = <s2|s1> as the reference's State overlap computes it
    (state.h:206-227, conjugating the second argument)."""
    from .tdvp import overlap_states
    return overlap_states(s2, s1)


def apply_ortho2(H, Va, Vb, tol, maxdim):
    """Later Krylov vectors (krylov.h:136-204): start from H Va by zip-up, then one
    left-to-right and one right-to-left two-site sweep that replaces each bond's tensor
    by H_eff applied to Va's two-site tensor, orthogonalised against Va's and Vb's
    two-site tensors (apply_ortho_site), and split by svd_impl<2,2> (tol, maxdim)."""
    gss = apply_operator(H, Va, tol, maxdim)
    ELH, ERH, ELa, ERa = _init_envs(H, gss, Va)
    _, _, ELb, ERb = _init_envs(None, gss, Vb)
    ns = len(H)
    for i in range(ns - 1):
        Ta = product(Va[i], Va[i + 1], 1)
        Tb = product(Vb[i], Vb[i + 1], 1)
        Nh = apply_double(ELH[-1], Ta, H[i], H[i + 1], ERH[-1])
        La = apply_double_state(ELa[-1], Ta, ERa[-1])
        Lb = apply_double_state(ELb[-1], Tb, ERb[-1])
        Mgs = apply_ortho_site(Nh, La, Lb)
        lef, rig = svd_impl(Mgs, 2, tol, maxdim, True)
        gss[i], gss[i + 1] = lef, rig
        if i < ns - 2:
            ELH.append(env_sweep_braket(ELH[-1], gss[i], H[i], Va[i], True))
            ERH.pop()
            ELa.append(_env_state_left(ELa[-1], gss[i], Va[i]))
            ERa.pop()
            ELb.append(_env_state_left(ELb[-1], gss[i], Vb[i]))
            ERb.pop()
    for i in range(ns - 1, 0, -1):
        Ta = product(Va[i - 1], Va[i], 1)
        Tb = product(Vb[i - 1], Vb[i], 1)
        Nh = apply_double(ELH[-1], Ta, H[i - 1], H[i], ERH[-1])
        La = apply_double_state(ELa[-1], Ta, ERa[-1])
        Lb = apply_double_state(ELb[-1], Tb, ERb[-1])
        Mgs = apply_ortho_site(Nh, La, Lb)
        lef, rig = svd_impl(Mgs, 2, tol, maxdim, False)
        gss[i - 1], gss[i] = lef, rig
        if i > 1:
            ERH.append(env_sweep_braket(ERH[-1], gss[i], H[i], Va[i], False))
            ELH.pop()
            ERa.append(_env_state_right(ERa[-1], gss[i], Va[i]))
            ELa.pop()
            ERb.append(_env_state_right(ERb[-1], gss[i], Vb[i]))
            ELb.pop()
    return gss


def scale_mpo(H, c):
    """Operator *= c (holstein/operator.h): the coupling scale goes on the first site."""
    return [H[0].scaled(c)] + [w.copy() for w in H[1:]]


def global_main(Hc, mps, tol_svd, maxdim, krylov_tol=1e-8, max_iter=15):
    """krylov.h:206-284 with Hc = H * dt: Krylov vectors as MPS (normalised after each
    apply_ortho), stop when the product of norms drops below krylov_tol or a norm exceeds
    1; T from <V_i|Hc|V_i> and <V_i|Hc|V_{i+1}> (Operator::join); c = V e^{-i Lambda} V^H e0;
    psi = sum c_i V_i with canonicalize(tol_svd, maxdim) after every addition.
    Returns (new MPS, iterations used)."""
    nrm = state_norm(mps)
    V = [state_scale(mps, 1.0 / nrm)]
    k, err = 0, 1.0
    while k < max_iter - 1 and err > krylov_tol:
        if k > 0:
            hv = apply_ortho2(Hc, V[k], V[k - 1], tol_svd, maxdim)
        else:
            hv = apply_ortho1(Hc, V[k], tol_svd, maxdim)
        nmhv = state_norm(hv)
        err *= nmhv
        if err < krylov_tol or nmhv > 1.0:
            break
        if k == max_iter - 1:
            break
        V.append(state_scale(hv, 1.0 / nmhv))
        k += 1
    n = len(V)
    hm = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        hm[i, i] = join(Hc, V[i], V[i])
        if i < n - 1:
            hm[i, i + 1] = join(Hc, V[i], V[i + 1])
    w, v = herm_eig_small(hm)
    c = v @ (np.exp(-1j * w) * np.conj(v[0, :]))
    out = state_scale(V[0], c[0])
    for i in range(1, n):
        out = canonicalize(state_add(out, state_scale(V[i], c[i])), tol_svd, maxdim)
    return out, n
