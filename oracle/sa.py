"""Stochastic-adaptive one-site TDVP (SA-1TDVP): the enrichment SVD, the quasi-Gaussian
model of the singular-value tails, the SVS singular-value log and the width fit.

TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import it; the product path never does.

Reference semantics followed (file:line in YiHeXu517/Kylin-V):
  quasi_gauss, inv_quasi_gauss   util/allocation.h:233-258
  gs_ortho                       util/allocation.h:260-327
  sa_svd_impl                    holstein/svd.h:443-614 (count rule :501-533, placement :541-594)
  SVS record of a two-site SVD   holstein/svd.h:358, :390-400; header holstein/tdvp.h:71-74
  width fit (the beta file)      bin/nls.py:1-58, read back by holstein/tdvp.h:343-358
  sa_site                        holstein/tdvp.h:230-294

Seeded redesign.  The reference draws every random number from a fresh
std::random_device, so no two of its runs agree.  Here every draw is a function of
(seed, call, stream, index) through the SplitMix64 finaliser (Steele, Lea and Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): ``call`` numbers the
enrichment SVDs of one sweep (0, 1, ... in sweep order), stream 0 gives the uniform
[0, 1) numbers of the count rule (index = draw number), stream 1 + v the entries of the
v-th new vector of that call (index = position along the vector, uniform [-1, 1), real
as the reference's AverRan values assigned to a complex entry).  Sectors receive new
vectors in ascending charge of their row sector (the reference visits them in the order
their charges first appear in the tensor's block list, an artefact of block insertion).
"""
from __future__ import annotations

import math

import numpy as np

from .blocksparse import Tensor
from .tdvp import _sector_layout

__all__ = [
    "draw_u01", "quasi_gauss", "inv_quasi_gauss", "extension_count", "sa_svd_impl",
    "svs_record", "svs_text", "parse_svs", "fit_widths", "widths_text", "betas_from_widths",
]

_M64 = (1 << 64) - 1
_PI_REF = 3.1415926  # the constant util/allocation.h:236,238 uses for pi


def _mix(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def draw_u01(seed: int, call: int, stream: int, index: int) -> float:
    """Uniform [0, 1) from the top 53 bits of mix(mix(mix(seed ^ call) ^ stream) ^ index)."""
    h = _mix((seed ^ call) & _M64)
    h = _mix(h ^ (stream & _M64))
    h = _mix(h ^ (index & _M64))
    return (h >> 11) * (1.0 / 9007199254740992.0)


def _draw_vector(seed: int, call: int, stream: int, n: int) -> np.ndarray:
    return np.array([2.0 * draw_u01(seed, call, stream, e) - 1.0 for e in range(n)])


def quasi_gauss(x: float, bet: float) -> float:
    """util/allocation.h:233-240, with its truncated pi."""
    r = math.exp(-0.25 * bet * (x * x)) * abs(0.25 * x / bet)
    r += (0.5 / bet - 0.25 * (x * x)) * math.erfc(0.5 * math.sqrt(bet) * abs(x)) * 0.5 * math.sqrt(_PI_REF / bet)
    r *= math.exp(-0.25 * bet * (x * x)) * 4.0 * bet * math.sqrt(bet / _PI_REF)
    return r


def inv_quasi_gauss(y: float, bet: float) -> float:
    """Bisection on [0, 2] to width 1e-8 for quasi_gauss(x) = y (util/allocation.h:241-258).
    Where no root lies in the bracket the interval closes on 2."""
    up, dn, eps = 2.0, 0.0, 1e-8
    res = 0.0
    while up - dn > eps:
        res = 0.5 * (up + dn)
        if (quasi_gauss(res, bet) - y) * (quasi_gauss(dn, bet) - y) < 0:
            up = res
        else:
            dn = res
    return res


def extension_count(svals_desc, maxdim: int, thres: float, beta: float, rds_of) -> int:
    """Number of states the enrichment adds (holstein/svd.h:501-533).

    The tail of the descending spectrum is continued by s_new = 10^(+-r) s_last^2 / s_prev
    (the sign alternating with the position), r = rds_of(j) for the j-th draw, clamped so
    it never rises, until a value falls to the cutoff or the spectrum reaches maxdim;
    nothing is added unless 5 < len(svals) < maxdim.
    """
    n = len(svals_desc)
    if not (5 < n < maxdim):
        return 0
    tmps = [float(svals_desc[n - 2]), float(svals_desc[n - 1])]
    add = 0
    for j, i in enumerate(range(n, maxdim)):
        rds = rds_of(j)
        last, prev = tmps[-1], tmps[-2]
        if (i + n) % 2 == 0:
            new = 10.0 ** rds * (last * last) / prev
        else:
            new = 10.0 ** (-1.0 * rds) * (last * last) / prev
        if new <= thres:
            break
        tmps.append(new)
        if tmps[-1] > tmps[-2]:
            tmps[-1] = tmps[-2]
        add = len(tmps) - 2
    return add


def _gs_extend(vecs: np.ndarray, new: list) -> np.ndarray:
    """gs_ortho (util/allocation.h:260-327): each new vector normalised, made orthogonal to
    every earlier one by modified Gram-Schmidt (zdotc / zaxpy), normalised again.
    vecs holds the vectors as columns."""
    out = [vecs[:, j].copy() for j in range(vecs.shape[1])]
    for r in new:
        v = r.astype(np.complex128)
        v = v / np.linalg.norm(v)
        for u in out:
            v = v - np.vdot(u, v) * u
        v = v / np.linalg.norm(v)
        out.append(v)
    return np.stack(out, axis=1)


def sa_svd_impl(A: Tensor, r1: int, cutoff: float, maxdim: int, left2right: bool, beta: float,
                seed: int, call: int, left_dir: int = -1):
    """sa_svd_impl (holstein/svd.h:443-614): full SVD of every sector (nothing truncated),
    S absorbed as in svd_impl, then extension_count new states placed into the sectors
    with room on the isometric side (rows for left2right, columns otherwise) in ascending
    row charge: random orthonormal vectors on the isometric side, zeros on the other, so
    the product of the two factors is still A.  Returns (left, right, added)."""
    rows, cols = _sector_layout(A, r1)
    pairs = []
    allv = []
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
        pairs.append([qr, qc, u, vh])
        allv.extend(float(x) for x in s)
    allv.sort(reverse=True)
    total_add = extension_count(allv, maxdim, cutoff, beta,
                                lambda j: inv_quasi_gauss(draw_u01(seed, call, 0, j), beta))
    v = 0
    for p in pairs:
        if total_add == 0:
            break
        u, vh = p[2], p[3]
        k = u.shape[1]
        room = (u.shape[0] - k) if left2right else (vh.shape[1] - k)
        if room <= 0:
            continue
        a = min(total_add, room)
        if left2right:
            new = [_draw_vector(seed, call, 1 + v + i, u.shape[0]) for i in range(a)]
            p[2] = _gs_extend(u, new)
            p[3] = np.vstack([vh, np.zeros((a, vh.shape[1]), dtype=np.complex128)])
        else:
            new = [_draw_vector(seed, call, 1 + v + i, vh.shape[1]) for i in range(a)]
            p[3] = _gs_extend(vh.T, new).T
            p[2] = np.hstack([u, np.zeros((u.shape[0], a), dtype=np.complex128)])
        total_add -= a
        v += a
    added = v
    ldirs = A.dirs[:r1] + (left_dir,)
    rdirs = (-left_dir,) + A.dirs[r1:]
    lef, rig = Tensor(ldirs), Tensor(rdirs)
    for qr, qc, u, vh in pairs:
        nk = u.shape[1]
        rlist, _ = rows[qr]
        clist, _ = cols[qc]
        for (rq, rsh, rpos, rsz) in rlist:
            lef.push_back_ch(tuple(rq) + (-left_dir * qr,), u[rpos:rpos + rsz, :].reshape(tuple(rsh) + (nk,)))
        for (cq, csh, cpos, csz) in clist:
            rig.push_back_ch((left_dir * qc,) + tuple(cq), vh[:, cpos:cpos + csz].reshape((nk,) + tuple(csh)))
    return lef, rig, added


def svs_record(kept_desc) -> list:
    """log10(s_i s_{i-2} / s_{i-1}^2) for i >= 2 over the kept values of a two-site SVD,
    descending (holstein/svd.h:390-400)."""
    s = [float(x) for x in kept_desc]
    out = []
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(2, len(s)):
            out.append(float(np.log10(np.float64(s[i] * s[i - 2]) / np.float64(s[i - 1] * s[i - 1]))))
    return out


def svs_text(nsites: int, records) -> str:
    """The SVS file: 'Total Site = N' (holstein/tdvp.h:71-74), then per two-site SVD the
    number of kept values and its record, one value per line, left-aligned in 20
    columns, 11 digits (holstein/svd.h:392-398)."""
    out = ["Total Site = %d\n" % nsites]
    for count, rec in records:
        out.append("%d\n" % count)
        out.extend("%-20.11e\n" % x for x in rec)
    return "".join(out)


def parse_svs(text: str):
    """Per-bond lists of SVS values (bin/nls.py:19-48).  Bonds 0..N-2 of the first sweep
    open the lists; later records append to list (record number mod 2N-2), folded back
    from the right half, where list N-1 does not exist and its values go to list N-2."""
    lines = text.splitlines(keepends=True)
    pos = [0]

    def readline():
        if pos[0] < len(lines):
            pos[0] += 1
            return lines[pos[0] - 1]
        return ""

    ss = readline()
    nsite = int(ss.split()[3])
    crl = nsite * 2 - 2
    cur = 0
    lists = []
    while ss:
        ss = readline()
        if 0 < len(ss) < 10:
            cur += 1
            nval = max(int(ss) - 2, 0)
            if cur <= crl:
                vals = [float(readline()) for _ in range(nval)]
                if cur < nsite:
                    lists.append(vals)
                else:
                    lists[crl - cur] = lists[crl - cur] + vals
            else:
                b = cur % crl
                if b > nsite - 1:
                    b = crl - b
                for _ in range(nval):
                    x = float(readline())
                    if b < len(lists):
                        lists[b].append(x)
                    else:
                        lists[b - 1].append(x)
    return lists


def fit_model(x, bet):
    """The fitted density (bin/nls.py:10-14)."""
    from scipy.special import erfc
    x = np.asarray(x, dtype=np.float64)
    r = np.exp(-0.25 * bet * np.power(x, 2)) * np.abs(x / 4.0 / bet)
    r += (np.ones(x.shape) / 2.0 / bet - np.power(x, 2.0) / 4.0) * erfc(np.sqrt(bet) * np.abs(x) / 2.0) \
        * 0.5 * np.sqrt(np.pi / bet)
    return r * np.exp(-bet * np.power(x, 2.0) / 4.0) * bet ** 2


def fit_widths(text: str) -> list:
    """bin/nls.py:49-58: per bond with more than two values, a 50-bin histogram and a
    least-squares fit of fit_model(left bin edges) to the bin fractions over bet > 0 from
    bet = 1 (scipy's curve_fit, as the reference); 1.0 otherwise."""
    from scipy.optimize import curve_fit
    out = []
    for vals in parse_svs(text):
        if len(vals) > 2:
            ni, edges = np.histogram(vals, 50)
            res = curve_fit(fit_model, edges[:-1], ni / len(vals), bounds=(0, np.inf))
            out.append(float(res[0][0]))
        else:
            out.append(1.0)
    return out


def widths_text(widths) -> str:
    """The beta file: one '{:21.15f}' line per bond (bin/nls.py:50-58)."""
    return "".join("{:21.15f}\n".format(w) for w in widths)


def betas_from_widths(text: str) -> list:
    """betaSet: 1 / value of every beta-file line longer than 3 characters
    (holstein/tdvp.h:347-357)."""
    return [1.0 / float(l) for l in text.splitlines() if len(l) > 3]
