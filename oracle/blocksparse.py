"""U(1) block-sparse tensors and contractions: a plain numpy (complex128) restatement.

TEST INFRASTRUCTURE ONLY. This module is the oracle that the HIP product path is
checked against; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg may import it. Nothing in kylin-v_amd/ uses it.

Reference semantics followed (file:line in YiHeXu517/Kylin-V):
  * Block = dense row-major array + per-leg qn          holstein/block.h:23-162
  * Tensor = dirs (+1 In / -1 Out) + list of blocks     holstein/tensor.h:16-212
  * push_back_ch: a block with equal (shape, qn) is summed into the existing one
                                                        holstein/tensor.h:173-186
  * check_sym: every block has sum(dir*qn) == 0        holstein/tensor.h:201-211
  * norm / overlap                                      holstein/tensor.h:214-241
  * transpose                                           holstein/tensor.h:244-260
  * product<N1,N2,nc>(m1, m2, T1, T2)                   holstein/product.h:61-201

Blocks are kept in a dict keyed by the qn tuple.  In a consistent tensor each leg's qn
fixes that leg's block dimension, so "same (shape, qn)" and "same qn" coincide; a
shape clash under one qn raises instead of silently making a second block.
"""
from __future__ import annotations

import numpy as np

__all__ = ["Tensor", "norm", "overlap", "transpose", "product", "to_dense_legs"]


class Tensor:
    """Block-sparse tensor. ``blocks`` maps qn tuple -> complex128 ndarray."""

    def __init__(self, dirs, blocks=None):
        self.dirs = tuple(int(d) for d in dirs)
        self.blocks: dict[tuple, np.ndarray] = {}
        if blocks:
            for q, b in blocks.items():
                self.push_back_ch(q, b)

    @property
    def rank(self):
        return len(self.dirs)

    def push_back_ch(self, qn, blk):
        """Sum into an existing block of the same qn, else append (tensor.h:173-186)."""
        qn = tuple(int(q) for q in qn)
        blk = np.asarray(blk, dtype=np.complex128)
        if blk.ndim != self.rank or len(qn) != self.rank:
            raise ValueError("block rank does not match tensor rank")
        old = self.blocks.get(qn)
        if old is None:
            self.blocks[qn] = blk.copy()
        else:
            if old.shape != blk.shape:
                raise ValueError(f"shape clash under qn {qn}: {old.shape} vs {blk.shape}")
            old += blk

    def copy(self):
        t = Tensor(self.dirs)
        t.blocks = {q: b.copy() for q, b in self.blocks.items()}
        return t

    def scaled(self, c):
        t = Tensor(self.dirs)
        t.blocks = {q: b * c for q, b in self.blocks.items()}
        return t

    def scale_(self, c):
        for q in self.blocks:
            self.blocks[q] *= c
        return self

    def add_(self, other, c=1.0):
        """self += c*other with push_back_ch semantics (tensor.h:88-99)."""
        for q, b in other.blocks.items():
            self.push_back_ch(q, c * b)
        return self

    def legs(self):
        """Per-leg {qn: dim} (check_legs, tensor.h:189-198)."""
        out = [dict() for _ in range(self.rank)]
        for q, b in self.blocks.items():
            for ax in range(self.rank):
                d = out[ax].setdefault(q[ax], b.shape[ax])
                if d != b.shape[ax]:
                    raise ValueError("inconsistent leg dimension")
        return out

    def check_sym(self):
        """True when every block conserves charge (tensor.h:201-211)."""
        return all(sum(d * x for d, x in zip(self.dirs, q)) == 0 for q in self.blocks)

    def sorted_items(self):
        return sorted(self.blocks.items(), key=lambda kv: kv[0])

    def nnz(self):
        return sum(b.size for b in self.blocks.values())


def norm(t: Tensor) -> float:
    """sqrt(sum over blocks |block|^2)  (tensor.h:214-223)."""
    return float(np.sqrt(sum(np.vdot(b, b).real for b in t.blocks.values())))


def overlap(ori: Tensor, oth: Tensor) -> complex:
    """sum over matching (shape, qn) blocks of zdotc(oth, ori) = <oth|ori>  (tensor.h:225-241)."""
    acc = 0j
    for q, b in ori.blocks.items():
        o = oth.blocks.get(q)
        if o is not None and o.shape == b.shape:
            acc += np.vdot(o.ravel(), b.ravel())
    return complex(acc)


def transpose(t: Tensor, axes) -> Tensor:
    """Permute legs of every block and the dirs (tensor.h:244-260, block.h:167-179)."""
    axes = tuple(axes)
    out = Tensor(tuple(t.dirs[a] for a in axes))
    for q, b in t.blocks.items():
        out.blocks[tuple(q[a] for a in axes)] = np.ascontiguousarray(np.transpose(b, axes))
    return out


def product(m1: Tensor, m2: Tensor, nc: int, t1: str = "N", t2: str = "N") -> Tensor:
    """Block-sparse tensordot following product<N1,N2,nc> (holstein/product.h:61-201).

    t1 == 'N': contract the LAST nc legs of m1;  'T'/'C': contract its FIRST nc legs
    ('C' also conjugates the data and negates the dirs of the kept legs, :78-81).
    t2 == 'N': contract the FIRST nc legs of m2; 'T'/'C': contract its LAST nc legs (:92-95).
    Result legs: kept legs of m1, then kept legs of m2.  Blocks pair when every
    contracted leg has equal shape and qn (:115-130); each pair's GEMM result is summed
    into the unique (shape, qn) output block (:167-198).  Mismatched contracted-leg
    dirs raise (the reference prints and exits, :102-110).
    """
    n1, n2 = m1.rank, m2.rank
    if t1 == "N":
        iax1, kept1 = n1 - nc, list(range(0, n1 - nc))
    else:
        iax1, kept1 = 0, list(range(nc, n1))
    if t2 == "N":
        iax2, kept2 = 0, list(range(nc, n2))
    else:
        iax2, kept2 = n2 - nc, list(range(0, n2 - nc))
    fac1 = -1 if t1 == "C" else 1
    fac2 = -1 if t2 == "C" else 1
    for ax in range(nc):
        if m1.dirs[iax1 + ax] * fac1 != -1 * m2.dirs[iax2 + ax] * fac2:
            raise ValueError("Contracted legs' directions mismatch")
    rd = [m1.dirs[a] * fac1 for a in kept1] + [m2.dirs[a] * fac2 for a in kept2]
    res = Tensor(rd)
    ca1 = list(range(iax1, iax1 + nc))
    ca2 = list(range(iax2, iax2 + nc))
    for q1, b1 in m1.blocks.items():
        for q2, b2 in m2.blocks.items():
            ok = True
            for ax in range(nc):
                if q1[iax1 + ax] != q2[iax2 + ax] or b1.shape[iax1 + ax] != b2.shape[iax2 + ax]:
                    ok = False
                    break
            if not ok:
                continue
            a = np.conj(b1) if t1 == "C" else b1
            b = np.conj(b2) if t2 == "C" else b2
            c = np.tensordot(a, b, axes=(ca1, ca2))
            qn = tuple(q1[a_] for a_ in kept1) + tuple(q2[a_] for a_ in kept2)
            res.push_back_ch(qn, c)
    return res


def to_dense_legs(t: Tensor, legs=None):
    """Expand to a dense array.  Each leg's sectors are laid out in ascending qn order.

    Returns (dense, offsets) where offsets[ax] maps qn -> start index on that leg.
    Used by tests to pin block-sparse contractions against plain einsum.
    """
    legs = legs if legs is not None else t.legs()
    offs, dims = [], []
    for ax in range(t.rank):
        o, pos = {}, 0
        for q in sorted(legs[ax]):
            o[q] = pos
            pos += legs[ax][q]
        offs.append(o)
        dims.append(pos)
    dense = np.zeros(dims, dtype=np.complex128)
    for q, b in t.blocks.items():
        sl = tuple(slice(offs[ax][q[ax]], offs[ax][q[ax]] + b.shape[ax]) for ax in range(t.rank))
        dense[sl] += b
    return dense, offs
