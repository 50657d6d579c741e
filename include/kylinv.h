/*
 * kylinv.h -- C ABI of the MI355X (gfx950) Holstein TDVP engine.
 *
 * Plain C: opaque handles, int status codes (0 = ok, < 0 = error, message in
 * kv_last_error), row-major interleaved complex<double> payloads (re, im pairs),
 * int32 charges and shapes.  No torch or HIP types cross this boundary.
 *
 * The reference (YiHeXu517/Kylin-V) has no FFI: its hot path is a set of header-only
 * C++ templates called by value.  Each entry point below names the template it
 * replaces.  Tensor conventions follow the reference exactly:
 *   dirs +1 = In, -1 = Out, every block satisfies sum(dir*qn) == 0
 *                                             (holstein/tensor.h:20-27, :201-211)
 *   MPS site   (a, s, b)        dirs {+1,+1,-1}   (holstein/state.h:35-39)
 *   MPO site   (s', wL, s, wR)  dirs {+1,+1,-1,-1} (holstein/operator.h:35-39)
 *   left env   (a', w, a)       dirs {+1,-1,-1}, boundary qn (0,0,0)
 *   right env  (b', w, b)       dirs {-1,+1,+1}, boundary qn (Q,0,Q)
 *                                             (holstein/tdvp.h:302-311)
 * Every call runs on the context's HIP stream and returns when that work has finished
 * (the reference's flow reads every result on the host before the next call).  One
 * context per host thread; contexts on one device may run concurrently.
 */
#ifndef KYLINV_H
#define KYLINV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kv_ctx kv_ctx;
typedef struct kv_tensor kv_tensor;
typedef struct kv_tdvp kv_tdvp;

/* Arithmetic of the effective-Hamiltonian and environment contractions.  Lanczos
 * scalars, vector algebra, the Krylov exponential and the bond SVD are FP64 always. */
enum kv_precision {
    KV_FP64 = 0,   /* complex FP64 on v_mfma_f64_16x16x4f64                        */
    KV_FP32 = 1,   /* complex FP32 on v_mfma_f32_16x16x4f32 (three real products per
                      complex one), FP32 accumulation                               */
    KV_BF16X3 = 2  /* FP32 storage; H_eff GEMMs as split-BF16 products (hi*hi + hi*lo +
                      lo*hi) on v_mfma_f32_16x16x32_bf16, FP32 accumulation; environment
                      updates requested in this mode run their GEMMs in FP32              */
};

/* Status codes. */
enum kv_status {
    KV_OK = 0,
    KV_ERR_ARG = -1,        /* bad argument (null handle, rank, size ...)            */
    KV_ERR_DIRS = -2,       /* contracted legs' directions mismatch
                               (the reference exits: holstein/product.h:102-110)    */
    KV_ERR_HIP = -3,        /* HIP runtime error                                     */
    KV_ERR_SOLVER = -4,     /* rocSOLVER failure (the reference exits on a LAPACK
                               error: util/allocation.h:219, :229)                  */
    KV_ERR_FORMAT = -5,     /* bad MPS/MPO file magic or truncated file
                               (holstein/state.h:134, holstein/operator.h:121)      */
    KV_ERR_SYM = -6,        /* block breaks charge conservation
                               (holstein/tensor.h:201-211)                          */
    KV_ERR_INTERNAL = -7
};

/* ---- context ------------------------------------------------------------------ */
int kv_ctx_create(int device, kv_ctx** out);
int kv_ctx_destroy(kv_ctx* ctx);
/* Last error message of this context (empty string if none). */
const char* kv_last_error(const kv_ctx* ctx);
/* Block until all work queued on the context's stream has finished. */
int kv_sync(kv_ctx* ctx);
/* Return the context's cached device buffers to the device (live tensors are kept).
 * The reference has no counterpart (it mallocs per block); for long-lived contexts that
 * run many differently shaped problems. */
int kv_ctx_trim(kv_ctx* ctx);

/* ---- block-sparse tensors (holstein/block.h:23-162, holstein/tensor.h:16-212) ---- */
/* Create a tensor with the given block structure, zero-filled.  qn and shape are
 * nblocks x rank int arrays.  Blocks with equal qn are rejected (the reference merges
 * them with push_back_ch, holstein/tensor.h:173-186: pass unique blocks). */
int kv_tensor_create(kv_ctx* ctx, int rank, const int* dirs, int nblocks,
                     const int* qn, const int* shape, kv_tensor** out);
int kv_tensor_destroy(kv_tensor* t);
/* Structure queries.  Blocks are reported in the tensor's own order (ascending qn). */
int kv_tensor_rank(const kv_tensor* t, int* rank, int* nblocks, int64_t* nelem);
int kv_tensor_dirs(const kv_tensor* t, int* dirs);
int kv_tensor_blocks(const kv_tensor* t, int* qn, int* shape, int64_t* offset);
/* Copy the payload (nelem complex values, blocks in the tensor's order, each block
 * row-major) in or out.  data = 2*nelem doubles. */
int kv_tensor_upload(kv_tensor* t, const double* data);
int kv_tensor_download(const kv_tensor* t, double* data);
/* Euclidean norm (holstein/tensor.h:214-223) and <b|a> over blocks of equal qn
 * (overlap(a, b), holstein/tensor.h:225-241). */
int kv_tensor_norm(const kv_tensor* t, double* out);
int kv_tensor_overlap(const kv_tensor* a, const kv_tensor* b, double* re_im);
/* <b|a> exactly as kv_tensor_overlap computes it (the same kernels, the same order, so the
 * same bits), allocating no device memory.  It has no counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The block-pair list, the partial sums and the result
 * live in a static per-device scratch of 64 KiB (resolved by kv_ctx_create); at most 2048
 * matching block pairs (KV_ERR_ARG beyond).  Every call on a device uses that device's one
 * scratch, so calls that overlap in time on one device race. */
int kv_tensor_overlap_noalloc(const kv_tensor* a, const kv_tensor* b, double* re_im);

/* ---- contractions ---------------------------------------------------------------- */
/* product<N1,N2,nc>(m1, m2, T1, T2) (holstein/product.h:61-201).  op = 'N', 'T', 'C'.
 * 'N' contracts the last nc legs of m1 / first nc legs of m2; 'T'/'C' the first nc
 * legs of m1 / last nc legs of m2; 'C' conjugates. */
int kv_product(kv_ctx* ctx, const kv_tensor* m1, const kv_tensor* m2, int nc,
               char op1, char op2, int precision, kv_tensor** out);

/* Profiling form of kv_heff_apply: builds H_eff once, applies it once untimed (tile
 * autotuning), then `reps` times timed by HIP events on the context stream.
 * out[8] = {seconds per apply, algorithmic flop per apply, GEMM kernel seconds per apply,
 * GEMM flop per apply, GEMM launches per apply, MPO-fold seconds per apply, MPO-fold flop
 * per apply, MPO-fold launches per apply}.  No reference counterpart (the reference's
 * Timer brackets whole modules, holstein/tdvp.h:384). */
int kv_heff_profile(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                    const kv_tensor* W2, const kv_tensor* R, const kv_tensor* x, int precision,
                    int reps, double* out);

/* Tensor algebra used by the TDVP adjoint.
 * kv_tensor_transpose: transpose(t, axes) (holstein/tensor.h:243-260), out leg a = t leg axes[a].
 * kv_tensor_axpy: out = y + alpha x over the union of both block sets (operator+=,
 *   holstein/tensor.h:88-105, which axpys into equal blocks and appends new ones).
 * kv_tensor_scale: t *= alpha in place (operator* / operator*=, holstein/tensor.h:107-123). */
int kv_tensor_transpose(kv_ctx* ctx, const kv_tensor* t, const int* axes, kv_tensor** out);
int kv_tensor_axpy(kv_ctx* ctx, const kv_tensor* x, const kv_tensor* y, double alpha_re,
                   double alpha_im, kv_tensor** out);
int kv_tensor_scale(kv_ctx* ctx, kv_tensor* t, double alpha_re, double alpha_im);

/* Effective Hamiltonian y = L . W1 . W2 . x . R.  nsite = 2 replaces apply_double
 * (holstein/product.h:478-486, W1 and W2 used); nsite = 1 replaces apply_single
 * (:547-553, W1 used, W2 ignored); nsite = 0 replaces apply_zero (:540-545, no MPO).
 * The output has the union block structure of x and H_eff(x). */
int kv_heff_apply(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                  const kv_tensor* W2, const kv_tensor* R, const kv_tensor* x,
                  int precision, kv_tensor** y);

/* H_eff with its sector-pairing lists built once per bond and kept on the device: the
 * build half of apply_double / apply_single / apply_zero (holstein/product.h:478-553),
 * which the reference repeats on every Lanczos iteration (holstein/exponential.h:113-124).
 * x gives the block structure; the plan's vectors have the union structure of x and
 * H_eff(x) (the output structure of kv_heff_apply).  L and R must outlive the plan (the
 * FP64 contractions read them at every apply); W1 and W2 are folded into the plan. */
typedef struct kv_heff_plan kv_heff_plan;
int kv_heff_plan_create(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                        const kv_tensor* W2, const kv_tensor* R, const kv_tensor* x,
                        int precision, kv_heff_plan** out);
/* y = scale * H_eff x, a new tensor with the plan's structure.  x with that structure (a
 * previous output) is read in place; x with a sub-structure is first spread into it. */
int kv_heff_plan_apply(kv_heff_plan* plan, const kv_tensor* x, double scale, kv_tensor** y);
int kv_heff_plan_destroy(kv_heff_plan* plan);

/* env_sweep(env, A, W, Left2Right) (holstein/product.h:488-503); with bra != NULL the
 * bra/ket form env_sweep(env, bra, W, ket) (holstein/product.h:505-521). */
int kv_env_update(kv_ctx* ctx, const kv_tensor* env, const kv_tensor* bra,
                  const kv_tensor* W, const kv_tensor* ket, int left_to_right,
                  int precision, kv_tensor** out);

/* local_krylov<Rank> (holstein/exponential.h:92-190): x <- exp(-/+ i (dt/2) H_eff) x/|x|
 * in place (the output tensor may gain blocks; the handle is updated).  reverse != 0
 * uses exp(+i ...).  Writes iterations used, final error estimate and 1 if the
 * Lanczos error reached tol (0 if the iteration cap was hit). */
typedef struct kv_krylov_exit {
    int converged;
    int iterations;
    double error;
    double loss_norm;
} kv_krylov_exit;
int kv_krylov_step(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                   const kv_tensor* W2, const kv_tensor* R, kv_tensor* x, double dt,
                   int reverse, double tol, int max_iter, int precision,
                   kv_krylov_exit* exit_info);

/* svd_impl<R1,R2>(A, cfg, LeftDir=-1) (holstein/svd.h:296-440): sector SVD in FP64,
 * global descending sort, drop entry i (never 0) while i >= maxdim or s_i <= cutoff;
 * S absorbed into V^dag if left_to_right else into U.  nkept receives the kept bond
 * dimension.  Sectors whose smaller side is at least three times maxdim + 32 take a block
 * subspace route: the isometric side's leading singular subspace from a seeded random
 * block of maxdim + 32 columns with power steps, Cholesky QR between the products and
 * Rayleigh-Ritz, each Ritz pair's residual checked against 1e-13 of the largest value;
 * a cut the certified values cannot settle sends the sector to the full route.  Otherwise
 * sectors whose smaller side is at least 64 are solved through their Gram matrix on the
 * isometric side (Hermitian eigen-decomposition: from side 512, or for sectors of
 * different sides, the engine's batched Householder tridiagonalisation, rocSOLVER's dstedc
 * and the kept vectors carried back through the reflectors; below that rocSOLVER's zheevd;
 * a second pass over the tail below 1e-5 of the largest value); smaller ones by rocSOLVER's
 * Jacobi SVD.  KV_SVD=full in the environment turns the block route off. */
int kv_svd_truncate(kv_ctx* ctx, const kv_tensor* A, int r1, double cutoff, int maxdim,
                    int left_to_right, kv_tensor** left, kv_tensor** right, int* nkept);

/* ---- TDVP driver (holstein/tdvp.h:34-386) ------------------------------------------ */
typedef struct kv_tdvp_config {
    double dt;          /* -v (src/kln-tdvp.cpp:41-44)                                 */
    double cutoff;      /* -t SVD cutoff, default 1e-8 (src/kln-tdvp.cpp:27)           */
    int maxdim;         /* -D bond cap, default 1000 (src/kln-tdvp.cpp:26)             */
    double krylov_tol;  /* fixed 1e-8 in the reference (src/kln-tdvp.cpp:27)           */
    int krylov_maxiter; /* fixed 15 in the reference (src/kln-tdvp.cpp:27)             */
    int precision;      /* kv_precision of H_eff / environment contractions            */
} kv_tdvp_config;

/* Takes copies of the MPS (nsites tensors) and MPO; builds the boundary environments
 * and the right-environment stack (holstein/tdvp.h:300-316). */
int kv_tdvp_create(kv_ctx* ctx, int nsites, kv_tensor* const* mps, kv_tensor* const* mpo,
                   const kv_tdvp_config* cfg, kv_tdvp** out);
int kv_tdvp_destroy(kv_tdvp* s);
/* One time step dt: TDVP::two_site (holstein/tdvp.h:144-227) or TDVP::one_site
 * (holstein/tdvp.h:78-141). */
int kv_tdvp_two_site(kv_tdvp* s);
int kv_tdvp_one_site(kv_tdvp* s);
/* One SA-1TDVP step (TDVP::sa_site, holstein/tdvp.h:230-294, with sa_svd_impl,
 * holstein/svd.h:443-614): the one-site step with every SVD left untruncated and enriched
 * by new zero-weight states, their number drawn from the quasi-Gaussian tail model with
 * betas[bond] (the reference's betaSet, 1 / the fitted width: holstein/tdvp.h:343-358; at
 * least nsites-1 values, each finite and > 0).  Seeded redesign of the reference's
 * std::random_device draws: the same (state, betas, seed) gives the same result.  The new
 * states are random vectors made orthonormal to the sector's singular vectors, in the
 * sectors with room in ascending row charge. */
int kv_tdvp_sa_one_site(kv_tdvp* s, const double* betas, int nbetas, uint64_t seed);
/* Starts the SVS singular-value log at path ("Total Site = N", holstein/tdvp.h:71-74);
 * every later two-site SVD appends its kept count and log10(s_i s_{i-2} / s_{i-1}^2),
 * i >= 2 (holstein/svd.h:358, :390-400).  KV_ERR_ARG when the file cannot be written. */
int kv_tdvp_svs_open(kv_tdvp* s, const char* path);
/* Host only (no device): the per-bond widths bin/nls.py fits to an SVS file (50-bin
 * histogram per bond, least-squares fit of its quasi-Gaussian density over width > 0, 1.0
 * for bonds with at most two values).  Writes min(nbonds, max_bonds) widths and the bond
 * count; KV_ERR_ARG if the file cannot be read, KV_ERR_FORMAT if it is not an SVS log. */
int kv_sa_fit_widths(const char* svs_path, double* widths, int max_bonds, int* nbonds);
/* One global Krylov step of the whole MPS: exp(-i dt H) by a Lanczos recurrence over MPS
 * (global_main, holstein/krylov.h:206-284, with H scaled by dt as holstein/tdvp.h:318-327),
 * truncating with the config's cutoff and maxdim, then rebuilding the environments.
 * iterations (may be NULL) receives the number of Krylov vectors. */
int kv_tdvp_global_krylov(kv_tdvp* s, int* iterations);
/* A copy of MPS site i (caller destroys it). */
int kv_tdvp_get_site(kv_tdvp* s, int i, kv_tensor** out);
/* Bond dimensions (nsites-1 ints), cf. State::max_bond (holstein/state.h:187-200). */
int kv_tdvp_bond_dims(kv_tdvp* s, int* dims);
/* Per-module timing in seconds, names of the reference's Timer modules
 * (holstein/tdvp.h:85-214): "Evl-2","Evl-1","Evl-0","SVD","Update","Sweep", plus
 * GPU time (HIP events on the context stream) and algorithmic flops of the H_eff matvec
 * kernels and of the environment-update kernels.  Writes up to n values in this order:
 *   0 Evl-2  1 Evl-1  2 Evl-0  3 SVD  4 Update  5 Sweep  (wall seconds)
 *   6 heff_gpu_s  7 heff_flop  8 env_gpu_s  9 env_flop  10 matvecs
 *   11 heff_gemm_gpu_s  12 heff_gemm_flop  13 heff_gemm_launches
 *   14 heff_wfold_gpu_s 15 heff_wfold_flop 16 heff_wfold_launches
 *   17 heff_gemm_bytes (algorithmic: each GEMM's operands and result counted once)
 *   18 heff_gemm32_gpu_s 19 heff_gemm32_flop 20 heff_gemm32_launches 21 heff_gemm32_bytes
 *      (the FP32-operand GEMM launches alone; the mixed path's first matvec per local
 *      step and every FP64-path launch are FP64)
 *   22 trd_gpu_s 23 trd_bytes 24 trd_columns: the bond SVD's Householder
 *      tridiagonalisations on the engine's own kernels (HIP-event time; algorithmic bytes =
 *      the Hermitian trailing triangle read once per column, 16 B per element; columns
 *      summed over the matrices of each batched call)
 *   25 blk_gemm_gpu_s 26 blk_gemm_flop 27 blk_gemm_launches 28 blk_gemm_bytes: the FP64
 *      GEMM launches of the bond SVD's block subspace route (HIP-event time; 8 m n k flop,
 *      operands and result once each)
 *   29 blk_orth_gpu_s 30 blk_orth_calls: its Cholesky QR kernels (factor + inverse)
 *   31 blk_eig_gpu_s 32 blk_eig_calls: its Rayleigh-Ritz eigensolves
 *   33 blk_sectors 34 blk_fallbacks: sectors solved by the route, and sectors it handed to
 *      the full eigensolver
 *   35 sa_added: states added by SA-1TDVP enrichment SVDs since the driver was created */
int kv_tdvp_stats(kv_tdvp* s, double* out, int n);
int kv_tdvp_reset_stats(kv_tdvp* s);

/* ---- observables (holstein/operator.h:178-198, holstein/state.h:206-227) ------------ */
/* <bra|O|ket> (Operator::join) and <s1|s2> (overlap(State, State)). */
int kv_join(kv_ctx* ctx, int nsites, kv_tensor* const* mpo, kv_tensor* const* bra,
            kv_tensor* const* ket, double* re_im);
int kv_overlap_states(kv_ctx* ctx, int nsites, kv_tensor* const* s1, kv_tensor* const* s2,
                      double* re_im);

/* ---- reference binary formats (holstein/state.h:107-157, operator.h:94-143) --------- */
/* MPS file: INT 111, INT ns, per site INT nb, per block INT shape[3], INT qn[3],
 * complex<double>[prod(shape)].  MPO: magic 112, rank-4 blocks.  load returns
 * KV_ERR_FORMAT on a bad magic (the reference exits). */
int kv_mps_load(kv_ctx* ctx, const char* path, int max_sites, int* nsites, kv_tensor** sites);
int kv_mpo_load(kv_ctx* ctx, const char* path, int max_sites, int* nsites, kv_tensor** sites);
int kv_mps_save(kv_ctx* ctx, const char* path, int nsites, kv_tensor* const* sites);
int kv_mpo_save(kv_ctx* ctx, const char* path, int nsites, kv_tensor* const* sites);

#ifdef __cplusplus
}
#endif
#endif
