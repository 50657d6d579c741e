"""Python host layer of the MI355X Holstein TDVP engine.

Thin ctypes binding of the C ABI in include/kylinv.h (libkylinv.so, built in-tree by
build.py), with names that mirror the reference's hot-path API (YiHeXu517/Kylin-V):

  Tensor            Block/Tensor  (holstein/block.h, holstein/tensor.h)
  product           product<N1,N2,nc>          (holstein/product.h:61-201)
  apply_double/single/zero                      (holstein/product.h:478-553)
  env_sweep                                      (holstein/product.h:488-521)
  local_krylov                                   (holstein/exponential.h:92-190)
  svd_impl                                       (holstein/svd.h:296-440)
  TDVP.two_site / one_site / total_start         (holstein/tdvp.h:78-385)
  join / overlap_states                          (holstein/operator.h:178-198,
                                                  holstein/state.h:206-227)
  State.load / Operator.load / save              (holstein/state.h:107-157,
                                                  holstein/operator.h:94-143)

There is no CPU fallback: every call goes through the HIP library, and importing this
package without the built library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkylinv.so")

KV_FP64, KV_FP32, KV_BF16X3 = 0, 1, 2
_PREC = {"fp64": KV_FP64, "fp32": KV_FP32, "bf16x3": KV_BF16X3,
         KV_FP64: KV_FP64, KV_FP32: KV_FP32, KV_BF16X3: KV_BF16X3}


class KVError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kylinv error {code}: {msg}")
        self.code = code


class _KrylovExit(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("error", ctypes.c_double), ("loss_norm", ctypes.c_double)]


class _TdvpConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("cutoff", ctypes.c_double), ("maxdim", ctypes.c_int),
                ("krylov_tol", ctypes.c_double), ("krylov_maxiter", ctypes.c_int),
                ("precision", ctypes.c_int)]


_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
_LP = ctypes.POINTER(ctypes.c_int64)

# exported symbol -> (restype, argtypes); tests check that the library exports them all
SIGNATURES = {
    "kv_ctx_create": (_I, [_I, _PP]),
    "kv_ctx_destroy": (_I, [_P]),
    "kv_last_error": (ctypes.c_char_p, [_P]),
    "kv_sync": (_I, [_P]),
    "kv_ctx_trim": (_I, [_P]),
    "kv_tensor_create": (_I, [_P, _I, _IP, _I, _IP, _IP, _PP]),
    "kv_tensor_destroy": (_I, [_P]),
    "kv_tensor_rank": (_I, [_P, _IP, _IP, _LP]),
    "kv_tensor_dirs": (_I, [_P, _IP]),
    "kv_tensor_blocks": (_I, [_P, _IP, _IP, _LP]),
    "kv_tensor_upload": (_I, [_P, _DP]),
    "kv_tensor_download": (_I, [_P, _DP]),
    "kv_tensor_norm": (_I, [_P, _DP]),
    "kv_tensor_overlap": (_I, [_P, _P, _DP]),
    "kv_tensor_overlap_noalloc": (_I, [_P, _P, _DP]),
    "kv_product": (_I, [_P, _P, _P, _I, ctypes.c_char, ctypes.c_char, _I, _PP]),
    "kv_heff_profile": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _I, _DP]),
    "kv_tensor_transpose": (_I, [_P, _P, _IP, _PP]),
    "kv_tensor_axpy": (_I, [_P, _P, _P, _D, _D, _PP]),
    "kv_tensor_scale": (_I, [_P, _P, _D, _D]),
    "kv_heff_apply": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _PP]),
    "kv_heff_plan_create": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _PP]),
    "kv_heff_plan_apply": (_I, [_P, _P, _D, _PP]),
    "kv_heff_plan_destroy": (_I, [_P]),
    "kv_env_update": (_I, [_P, _P, _P, _P, _P, _I, _I, _PP]),
    "kv_krylov_step": (_I, [_P, _I, _P, _P, _P, _P, _P, _D, _I, _D, _I, _I,
                            ctypes.POINTER(_KrylovExit)]),
    "kv_svd_truncate": (_I, [_P, _P, _I, _D, _I, _I, _PP, _PP, _IP]),
    "kv_tdvp_create": (_I, [_P, _I, _PP, _PP, ctypes.POINTER(_TdvpConfig), _PP]),
    "kv_tdvp_destroy": (_I, [_P]),
    "kv_tdvp_two_site": (_I, [_P]),
    "kv_tdvp_one_site": (_I, [_P]),
    "kv_tdvp_sa_one_site": (_I, [_P, _DP, _I, ctypes.c_uint64]),
    "kv_tdvp_svs_open": (_I, [_P, ctypes.c_char_p]),
    "kv_sa_fit_widths": (_I, [ctypes.c_char_p, _DP, _I, _IP]),
    "kv_tdvp_global_krylov": (_I, [_P, _IP]),
    "kv_tdvp_get_site": (_I, [_P, _I, _PP]),
    "kv_tdvp_bond_dims": (_I, [_P, _IP]),
    "kv_tdvp_stats": (_I, [_P, _DP, _I]),
    "kv_tdvp_reset_stats": (_I, [_P]),
    "kv_join": (_I, [_P, _I, _PP, _PP, _PP, _DP]),
    "kv_overlap_states": (_I, [_P, _I, _PP, _PP, _DP]),
    "kv_mps_load": (_I, [_P, ctypes.c_char_p, _I, _IP, _PP]),
    "kv_mpo_load": (_I, [_P, ctypes.c_char_p, _I, _IP, _PP]),
    "kv_mps_save": (_I, [_P, ctypes.c_char_p, _I, _PP]),
    "kv_mpo_save": (_I, [_P, ctypes.c_char_p, _I, _PP]),
}

_lib = None


def library(path: str | None = None):
    """Load libkylinv.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build the HIP library first (__graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Context:
    """One HIP device + stream (kv_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = library()
        h = ctypes.c_void_p()
        rc = self.lib.kv_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise KVError(rc, f"kv_ctx_create(device={device}) failed")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise KVError(rc, self.lib.kv_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.kv_sync(self.h))

    def trim(self):
        """Release cached device buffers (kv_ctx_trim)."""
        self.check(self.lib.kv_ctx_trim(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.kv_ctx_destroy(self.h)
            self.h = None


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("KV_DEVICE", "0")))
    return _default_ctx


def _iarr(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
    return a, a.ctypes.data_as(_IP)


class Tensor:
    """Device-resident U(1) block-sparse tensor (kv_tensor handle)."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.h = ctx, handle

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.kv_tensor_destroy(self.h)
        except Exception:
            pass
        self.h = None

    @classmethod
    def from_blocks(cls, dirs, blocks, ctx: Context | None = None):
        """blocks: dict qn-tuple -> complex ndarray (row-major)."""
        ctx = ctx or default_context()
        dirs = tuple(int(d) for d in dirs)
        rank = len(dirs)
        keys = sorted(blocks)
        qn = np.array([list(k) for k in keys], dtype=np.int32).reshape(-1, rank)
        shape = np.array([list(np.shape(blocks[k])) for k in keys], dtype=np.int32).reshape(-1, rank)
        d_a, d_p = _iarr(dirs)
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.kv_tensor_create(ctx.h, rank, d_p, len(keys), qn.ctypes.data_as(_IP),
                                           shape.ctypes.data_as(_IP), ctypes.byref(h)))
        t = cls(ctx, h)
        if keys:
            data = np.concatenate([np.asarray(blocks[k], dtype=np.complex128).ravel() for k in keys])
            t.upload(data)
        return t

    def structure(self):
        rank, nb, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        self.ctx.check(self.ctx.lib.kv_tensor_rank(self.h, ctypes.byref(rank), ctypes.byref(nb),
                                                   ctypes.byref(ne)))
        r, n = rank.value, nb.value
        dirs = np.zeros(r, dtype=np.int32)
        qn = np.zeros((max(n, 1), r), dtype=np.int32)
        sh = np.zeros((max(n, 1), r), dtype=np.int32)
        off = np.zeros(max(n, 1), dtype=np.int64)
        self.ctx.check(self.ctx.lib.kv_tensor_dirs(self.h, dirs.ctypes.data_as(_IP)))
        self.ctx.check(self.ctx.lib.kv_tensor_blocks(self.h, qn.ctypes.data_as(_IP),
                                                     sh.ctypes.data_as(_IP),
                                                     off.ctypes.data_as(_LP)))
        return tuple(int(d) for d in dirs), qn[:n], sh[:n], off[:n], int(ne.value)

    @property
    def dirs(self):
        return self.structure()[0]

    @property
    def rank(self):
        return len(self.dirs)

    def upload(self, flat_complex):
        a = np.ascontiguousarray(np.asarray(flat_complex, dtype=np.complex128))
        self.ctx.check(self.ctx.lib.kv_tensor_upload(self.h, a.view(np.float64).ctypes.data_as(_DP)))

    def download(self):
        ne = self.structure()[4]
        a = np.zeros(ne, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.kv_tensor_download(self.h, a.view(np.float64).ctypes.data_as(_DP)))
        return a

    def to_blocks(self):
        dirs, qn, sh, off, ne = self.structure()
        flat = self.download()
        out = {}
        for q, s, o in zip(qn, sh, off):
            n = int(np.prod(s))
            out[tuple(int(x) for x in q)] = flat[o:o + n].reshape(tuple(int(x) for x in s))
        return dirs, out

    def norm(self):
        v = ctypes.c_double()
        self.ctx.check(self.ctx.lib.kv_tensor_norm(self.h, ctypes.byref(v)))
        return v.value

    def bond_dim(self, axis):
        dirs, qn, sh, off, ne = self.structure()
        seen = {}
        for q, s in zip(qn, sh):
            seen[int(q[axis])] = int(s[axis])
        return sum(seen.values())


def overlap(ori: Tensor, oth: Tensor) -> complex:
    """<oth|ori> over blocks of equal qn (holstein/tensor.h:225-241)."""
    v = (ctypes.c_double * 2)()
    ori.ctx.check(ori.ctx.lib.kv_tensor_overlap(ori.h, oth.h, v))
    return complex(v[0], v[1])


def overlap_noalloc(ori: Tensor, oth: Tensor) -> complex:
    """overlap() without any device allocation (kv_tensor_overlap_noalloc): the device's
    static scratch holds the pair list and partial sums, so calls on one device must not
    overlap in time."""
    v = (ctypes.c_double * 2)()
    ori.ctx.check(ori.ctx.lib.kv_tensor_overlap_noalloc(ori.h, oth.h, v))
    return complex(v[0], v[1])


def _new(ctx):
    return ctypes.c_void_p()


def product(m1: Tensor, m2: Tensor, nc: int, t1: str = "N", t2: str = "N") -> Tensor:
    ctx = m1.ctx
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kv_product(ctx.h, m1.h, m2.h, nc, t1.encode(), t2.encode(), KV_FP64,
                                 ctypes.byref(h)))
    return Tensor(ctx, h)


def transpose(t: Tensor, axes) -> Tensor:
    """transpose(t, axes) (holstein/tensor.h:243-260): out leg a is t leg axes[a]."""
    ctx = t.ctx
    h = ctypes.c_void_p()
    keep, ax = _iarr(list(axes))
    ctx.check(ctx.lib.kv_tensor_transpose(ctx.h, t.h, ax, ctypes.byref(h)))
    del keep
    return Tensor(ctx, h)


def axpy(x: Tensor, y: Tensor, alpha=1.0) -> Tensor:
    """New tensor y + alpha x over the union of both block sets (operator+=,
    holstein/tensor.h:88-105)."""
    ctx = y.ctx
    a = complex(alpha)
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kv_tensor_axpy(ctx.h, x.h, y.h, a.real, a.imag, ctypes.byref(h)))
    return Tensor(ctx, h)


def scale_(t: Tensor, alpha) -> Tensor:
    """t *= alpha in place (holstein/tensor.h:107-123); returns t."""
    a = complex(alpha)
    t.ctx.check(t.ctx.lib.kv_tensor_scale(t.ctx.h, t.h, a.real, a.imag))
    return t


def _heff(nsite, L, W1, W2, R, x, precision):
    ctx = x.ctx
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kv_heff_apply(ctx.h, nsite, L.h, W1.h if W1 is not None else None,
                                    W2.h if W2 is not None else None, R.h, x.h,
                                    _PREC[precision], ctypes.byref(h)))
    return Tensor(ctx, h)


class HeffPlan:
    """H_eff built once per bond (kv_heff_plan_create): apply() is y = scale * H_eff x with
    the pairing lists kept on the device, as a Lanczos loop calls it.  The environments
    must stay alive as long as the plan."""

    def __init__(self, nsite, envL, x, mpo1, mpo2, envR, precision="fp64"):
        ctx = x.ctx
        self.ctx = ctx
        self._keep = (envL, envR)
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.kv_heff_plan_create(ctx.h, nsite, envL.h, mpo1.h if mpo1 is not None else None,
                                              mpo2.h if mpo2 is not None else None, envR.h, x.h,
                                              _PREC[precision], ctypes.byref(h)))
        self.h = h

    def apply(self, x, scale=1.0):
        y = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.kv_heff_plan_apply(self.h, x.h, float(scale), ctypes.byref(y)))
        return Tensor(self.ctx, y)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.kv_heff_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def heff_profile(nsite, envL, x, mpo1, mpo2, envR, precision="fp64", reps=5):
    """Times `reps` applications of one H_eff (kv_heff_profile); returns a dict of per-apply
    seconds, flop and launch counts."""
    ctx = x.ctx
    out = (ctypes.c_double * 8)()
    ctx.check(ctx.lib.kv_heff_profile(ctx.h, nsite, envL.h, mpo1.h if mpo1 is not None else None,
                                      mpo2.h if mpo2 is not None else None, envR.h, x.h,
                                      _PREC[precision], reps, out))
    keys = ("apply_s", "flop", "gemm_s", "gemm_flop", "gemm_launches", "wfold_s", "wfold_flop",
            "wfold_launches")
    return dict(zip(keys, list(out)))


def apply_double(envL, x, mpo1, mpo2, envR, precision="fp64"):
    return _heff(2, envL, mpo1, mpo2, envR, x, precision)


def apply_single(envL, x, mpo, envR, precision="fp64"):
    return _heff(1, envL, mpo, None, envR, x, precision)


def apply_zero(envL, c, envR, precision="fp64"):
    return _heff(0, envL, None, None, envR, c, precision)


def env_sweep(env, mps, mpo, left2right: bool, bra=None, precision="fp64") -> Tensor:
    """env_sweep(env, A, W, l2r); with bra the bra/ket form (holstein/product.h:488-521)."""
    ctx = env.ctx
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kv_env_update(ctx.h, env.h, bra.h if bra is not None else None, mpo.h,
                                    mps.h, 1 if left2right else 0, _PREC[precision],
                                    ctypes.byref(h)))
    return Tensor(ctx, h)


def local_krylov(envl, x: Tensor, mpo1, mpo2, envr, dt, tol=1e-8, max_iter=15, reverse=False,
                 nsite=None, precision="fp64"):
    """Returns (x_new, exit dict); x is replaced in place on the device handle as well."""
    ctx = x.ctx
    if nsite is None:
        nsite = x.rank - 2
    ex = _KrylovExit()
    ctx.check(ctx.lib.kv_krylov_step(ctx.h, nsite, envl.h, mpo1.h if mpo1 is not None else None,
                                     mpo2.h if mpo2 is not None else None, envr.h, x.h, dt,
                                     1 if reverse else 0, tol, max_iter, _PREC[precision],
                                     ctypes.byref(ex)))
    return x, dict(converged=bool(ex.converged), iters=ex.iterations, err=ex.error,
                   loss_norm=ex.loss_norm)


def svd_impl(A: Tensor, r1: int, cutoff: float, maxdim: int, left2right: bool):
    ctx = A.ctx
    l, r, k = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
    ctx.check(ctx.lib.kv_svd_truncate(ctx.h, A.h, r1, cutoff, maxdim, 1 if left2right else 0,
                                      ctypes.byref(l), ctypes.byref(r), ctypes.byref(k)))
    return Tensor(ctx, l), Tensor(ctx, r)


def _handles(ts):
    arr = (ctypes.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = t.h
    return arr


def join(mpo, bra, ket) -> complex:
    """<bra|O|ket> (holstein/operator.h:178-198)."""
    ctx = ket[0].ctx
    v = (ctypes.c_double * 2)()
    ctx.check(ctx.lib.kv_join(ctx.h, len(mpo), _handles(mpo), _handles(bra), _handles(ket), v))
    return complex(v[0], v[1])


def overlap_states(s1, s2) -> complex:
    """<s1|s2> (holstein/state.h:206-227)."""
    ctx = s1[0].ctx
    v = (ctypes.c_double * 2)()
    ctx.check(ctx.lib.kv_overlap_states(ctx.h, len(s1), _handles(s1), _handles(s2), v))
    return complex(v[0], v[1])


def _load(kind, path, ctx=None):
    ctx = ctx or default_context()
    n = ctypes.c_int()
    fn = ctx.lib.kv_mps_load if kind == "mps" else ctx.lib.kv_mpo_load
    ctx.check(fn(ctx.h, path.encode(), 0, ctypes.byref(n), None))
    arr = (ctypes.c_void_p * n.value)()
    ctx.check(fn(ctx.h, path.encode(), n.value, ctypes.byref(n), arr))
    return [Tensor(ctx, ctypes.c_void_p(arr[i])) for i in range(n.value)]


def load_mps(path, ctx=None):
    """State::load (holstein/state.h:130-157)."""
    return _load("mps", path, ctx)


def load_mpo(path, ctx=None):
    """Operator::load (holstein/operator.h:117-143)."""
    return _load("mpo", path, ctx)


def save_mps(path, sites):
    ctx = sites[0].ctx
    ctx.check(ctx.lib.kv_mps_save(ctx.h, path.encode(), len(sites), _handles(sites)))


def save_mpo(path, sites):
    ctx = sites[0].ctx
    ctx.check(ctx.lib.kv_mpo_save(ctx.h, path.encode(), len(sites), _handles(sites)))


STAT_NAMES = ["Evl-2", "Evl-1", "Evl-0", "SVD", "Update", "Sweep", "heff_gpu_s", "heff_flop",
              "env_gpu_s", "env_flop", "matvecs", "heff_gemm_gpu_s", "heff_gemm_flop",
              "heff_gemm_launches", "heff_wfold_gpu_s", "heff_wfold_flop", "heff_wfold_launches",
              "heff_gemm_bytes", "heff_gemm32_gpu_s", "heff_gemm32_flop", "heff_gemm32_launches",
              "heff_gemm32_bytes", "trd_gpu_s", "trd_bytes", "trd_columns", "blk_gemm_gpu_s",
              "blk_gemm_flop", "blk_gemm_launches", "blk_gemm_bytes", "blk_orth_gpu_s", "blk_orth_calls",
              "blk_eig_gpu_s", "blk_eig_calls", "blk_sectors", "blk_fallbacks", "sa_added"]


class TDVP:
    """Sweep driver (holstein/tdvp.h:34-386) on one GPU."""

    def __init__(self, mps, mpo, cutoff, maxdim, dt, tol_krylov=1e-8, max_iter=15,
                 precision="fp64"):
        self.ctx = mps[0].ctx
        cfg = _TdvpConfig(dt, cutoff, maxdim, tol_krylov, max_iter, _PREC[precision])
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.kv_tdvp_create(self.ctx.h, len(mps), _handles(mps),
                                                   _handles(mpo), ctypes.byref(cfg),
                                                   ctypes.byref(h)))
        self.h = h
        self.ns = len(mps)

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.kv_tdvp_destroy(self.h)
        except Exception:
            pass
        self.h = None

    def two_site(self):
        self.ctx.check(self.ctx.lib.kv_tdvp_two_site(self.h))

    def one_site(self):
        self.ctx.check(self.ctx.lib.kv_tdvp_one_site(self.h))

    def sa_one_site(self, betas, seed: int):
        """One SA-1TDVP step (TDVP::sa_site, holstein/tdvp.h:230-294) with betaSet =
        ``betas`` (1 / fitted width per bond) and the draws of seed ``seed``."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        self.ctx.check(self.ctx.lib.kv_tdvp_sa_one_site(self.h, b.ctypes.data_as(_DP), len(b),
                                                        ctypes.c_uint64(seed)))

    def svs_open(self, path):
        """Start the SVS log at ``path``; every later two-site SVD appends its record."""
        self.ctx.check(self.ctx.lib.kv_tdvp_svs_open(self.h, os.fsencode(path)))

    @property
    def mps(self):
        out = []
        for i in range(self.ns):
            h = ctypes.c_void_p()
            self.ctx.check(self.ctx.lib.kv_tdvp_get_site(self.h, i, ctypes.byref(h)))
            out.append(Tensor(self.ctx, h))
        return out

    def bond_dims(self):
        a = np.zeros(self.ns - 1, dtype=np.int32)
        self.ctx.check(self.ctx.lib.kv_tdvp_bond_dims(self.h, a.ctypes.data_as(_IP)))
        return [int(x) for x in a]

    def stats(self):
        a = np.zeros(len(STAT_NAMES))
        self.ctx.check(self.ctx.lib.kv_tdvp_stats(self.h, a.ctypes.data_as(_DP), len(a)))
        return dict(zip(STAT_NAMES, a.tolist()))

    def reset_stats(self):
        self.ctx.check(self.ctx.lib.kv_tdvp_reset_stats(self.h))

    def global_krylov(self):
        """One global Krylov step (global_main, holstein/krylov.h:206-284); returns the
        number of Krylov vectors."""
        n = ctypes.c_int()
        self.ctx.check(self.ctx.lib.kv_tdvp_global_krylov(self.h, ctypes.byref(n)))
        return n.value

    def total_start(self, stage, on_step=None, svs_path=None, seed: int = 1):
        """Stage list [nK, n2, nSA, n1] (parse_stage, holstein/tdvp.h:19-32, stages as
        total_start, holstein/tdvp.h:318-383).  The SA-1TDVP stage fits the widths to the
        SVS log the two-site sweeps wrote to ``svs_path`` (required when nSA > 0) and runs
        sweep k under seed + k."""
        nK, n2, nSA, n1 = stage
        if nSA and not svs_path:
            raise KVError(-1, "the SA-1TDVP stage needs the SVS log of the two-site sweeps")
        if svs_path:
            self.svs_open(svs_path)
        k = 0
        for _ in range(nK):
            self.global_krylov()
            if on_step:
                on_step(k, self)
            k += 1
        for _ in range(n2):
            self.two_site()
            if on_step:
                on_step(k, self)
            k += 1
        if nSA:
            betas = [1.0 / float("%21.15f" % w) for w in fit_widths(svs_path)]
            for _ in range(nSA):
                self.sa_one_site(betas, seed + k)
                if on_step:
                    on_step(k, self)
                k += 1
        for _ in range(n1):
            self.one_site()
            if on_step:
                on_step(k, self)
            k += 1


def fit_widths(svs_path):
    """Per-bond widths fitted to an SVS log (bin/nls.py), host only."""
    lib = library()
    n = ctypes.c_int()
    rc = lib.kv_sa_fit_widths(os.fsencode(svs_path), None, 0, ctypes.byref(n))
    if rc:
        raise KVError(rc, "kv_sa_fit_widths failed on %s" % svs_path)
    w = np.zeros(max(n.value, 1))
    rc = lib.kv_sa_fit_widths(os.fsencode(svs_path), w.ctypes.data_as(_DP), n.value, ctypes.byref(n))
    if rc:
        raise KVError(rc, "kv_sa_fit_widths failed on %s" % svs_path)
    return [float(x) for x in w[:n.value]]


def parse_stage(s: str):
    """'[nK,n2,nSA,n1]' -> 4 ints (holstein/tdvp.h:19-32: drops the first character and
    splits on commas)."""
    parts = s[1:].split(",")
    return [int(p.strip().rstrip("]")) for p in parts[:4]]
