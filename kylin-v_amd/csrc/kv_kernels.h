// Launcher entry points of the gfx950 kernels.  Work descriptors live in kv_types.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kv_types.h"

namespace kv {

// cfg selects the tile shape (kGemmConfigs in kv_types.h); tiles must have been cut for it
template <class T2>
void launch_gemm(hipStream_t st, int opA, int opB, const T2* A, const T2* B, T2* C,
                 const GemmTile* tiles, int ntiles, const GemmGroup* groups,
                 const GemmSeg* segs, double alpha_re, double alpha_im, int cfg = 0);
// Tile configs built per op pair and precision (every config is free of register spills).
bool gemm_config_supports(int cfg, int opA, int opB, bool fp64);
// Split-BF16 variant on the tile list of config 0 (64x64) or 4 (128x128): FP32 complex
// operands and output.
void launch_gemm_bf16x3(hipStream_t st, int opA, int opB, const float2* A, const float2* B,
                        float2* C, const GemmTile* tiles, int ntiles, const GemmGroup* groups,
                        const GemmSeg* segs, double alpha_re, double alpha_im, int cfg);

// MPO fold; maxPQ = the largest task's P x Q (sets the workgroups per task).
template <class T2>
void launch_wfold(hipStream_t st, const T2* in, T2* out, const WTask* tasks, int ntasks,
                  const WTerm* terms, int64_t maxPQ);

template <class T2>
void launch_perm(hipStream_t st, const T2* in, T2* out, const PermTask* tasks, int ntasks,
                 int64_t maxsize);

template <class T2>
void launch_copy2d(hipStream_t st, const T2* src, T2* dst, const Copy2D* cps, int ncopies,
                   int maxrows, const double* scale);

// dst[c*ld_dst + r] = conj(src[r*ld_src + c]), src is rows x cols
void launch_ctranspose(hipStream_t st, const double2* src, double2* dst, int rows, int cols,
                       int ld_src, int ld_dst);

// FP64 vector algebra over flat slabs of n complex values.
// bytes must be a multiple of 4
void launch_zero(hipStream_t st, void* p, size_t bytes);
// out[0] = sum conj(a) b  (double2); scratch >= 2*1024 doubles
void launch_dotc(hipStream_t st, const double2* a, const double2* b, int64_t n,
                 double2* out, double* scratch);
// out[0] = sum |a|^2
void launch_nrm2sq(hipStream_t st, const double2* a, int64_t n, double* out, double* scratch);
// y -= c[0] * x   (c on the device)
void launch_axpy_neg_dev(hipStream_t st, double2* y, const double2* x, const double2* c,
                         int64_t n);
// y -= c[0] x (c on the device), then out = sum conj(y) z (complex) or sum |y|^2: the
// subtraction and the reduction that follows it in one pass over y, with the arithmetic of
// launch_axpy_neg_dev followed by launch_dotc / launch_nrm2sq.
void launch_axpy_dotc(hipStream_t st, double2* y, const double2* x, const double2* c, const double2* z, int64_t n,
                      double2* out, double* scratch);
void launch_axpy_nrm2sq(hipStream_t st, double2* y, const double2* x, const double2* c, int64_t n, double* out,
                        double* scratch);
// dst = src / sqrt(nsq[0])  (nsq on the device)
void launch_scale_invsqrt_dev(hipStream_t st, double2* dst, const double2* src,
                              const double* nsq, int64_t n);
// dst = src * s
void launch_scale(hipStream_t st, double2* dst, const double2* src, double s, int64_t n);
void launch_zaxpby(hipStream_t st, double2* out, const double2* x, const double2* y, double2 a,
                   double2 b, int64_t n);
// out = sum_i c[i] * V[i*n + e],  c on the device
void launch_lincomb(hipStream_t st, double2* out, const double2* V, const double2* c, int nv,
                    int64_t n);
void launch_f64_to_f32(hipStream_t st, const double2* src, float2* dst, int64_t n);
// dst = s * src
void launch_f32_to_f64(hipStream_t st, const float2* src, double2* dst, double s, int64_t n);

// A[i ld + i] = (v, 0) for i in [from, to) (the diagonal of a square matrix).
void launch_set_diag(hipStream_t st, double2* A, int64_t ld, int from, int to, double v);
// dst (n x n, ld n) = the lower (j <= i) or upper (j >= i) triangle of src (ld lds), the rest zero.
void launch_tri_copy(hipStream_t st, const double2* src, int64_t lds, double2* dst, int n, bool lower);

// Back-transformation of a lower-triangle Householder tridiagonalisation (LAPACK layout:
// reflector i stored below the subdiagonal of column i, tau_i separately).
// V (column-major n x ncols): the reflector vectors written out with their unit entries.
void launch_hh_explicit(hipStream_t st, const double2* A, double2* V, int n, int ncols);
// T (bw x bw per block, column-major, upper; bw <= 128): triangular factor of reflectors
// [b bw, min((b+1) bw, nref)) from S_b = V_b^H V_b, so the block is I - V_b T_b V_b^H.
void launch_hh_tfactor(hipStream_t st, const double2* S, const double2* tau, double2* T, int nblocks,
                       int bw, int nref);
// out[e] = sum_{p < np} part[p n + e]: partial products of a split-K GEMM summed.
void launch_sum_parts(hipStream_t st, const double2* part, int np, int64_t n, double2* out);
// Z[j n + i] = C[(c0 + j) ldc + i]  (i < n, j < k): real columns to complex.
void launch_real_cols(hipStream_t st, const double* C, int ldc, int c0, int k, int n, double2* Z);

// Batched Householder tridiagonalisation (lower, LAPACK zhetrd storage) of column-major
// n x n Hermitian matrices at stride n * n, driven column by column from the host
// (hetrd_lower): panel width kTrdPanel; workspace per matrix VW (n x 2 nb), y (n), q (2 nb),
// part (trd_column_blocks(n, 0) complex) and count (one unsigned, zeroed before the first
// step; every step leaves it zero).
constexpr int kTrdPanel = 32;
void launch_herm_fill(hipStream_t st, double2* A, int n, int batch);
// The same reduction (D, E, tau, stored reflectors) for sides up to kTrdSmallMax in one
// workgroup per matrix; A must hold both triangles (launch_herm_fill).
constexpr int kTrdSmallMax = 256;
void launch_trd_small(hipStream_t st, double2* A, int n, double* D, double* E, double2* tau, int batch);
int trd_column_blocks(int n, int j);
void launch_trd_column(hipStream_t st, double2* A, double2* VW, const double2* y, const double2* q, double* D,
                       double* E, double2* tau, double2* part, unsigned* count, int n, int j, int i, int nb,
                       bool prev, bool col, int batch);
void launch_trd_matvec(hipStream_t st, const double2* A, const double2* VW, double2* y, double2* q, int n, int s,
                       int i, int nb, int batch);

// Block subspace route of the bond SVD.
// p[e] = uniform [-1, 1)^2 complex value determined by (seed, e).
void launch_rand_fill(hipStream_t st, double2* p, int64_t n, uint64_t seed);
// Largest block of the route's Cholesky QR, and the panel width of its Cholesky.
constexpr int kCholMax = 1024;
constexpr int kCholLeftNB = 32;
// Batched left-looking blocked Cholesky S = R^H R (upper, row-major, ld, batch stride) with the
// rank-tolerant pivot drop (tau * the largest diagonal entry, into dmax), forming T = R^{-1}
// (same layout; dropped columns give zero rows and columns) and the live flags.  S's tiles
// right of each diagonal block receive R; its diagonal tiles keep their input.  Two launches
// per 32-column panel; Rd (batch x ceil(ld / 32) x 32 x 32 complex) receives the inverted
// diagonal blocks.
inline int64_t chol_left_rd_elems(int ld, int batch) {
    return (int64_t)batch * ((ld + kCholLeftNB - 1) / kCholLeftNB) * kCholLeftNB * kCholLeftNB;
}
void launch_chol_left(hipStream_t st, double2* S, const int* sizes, int nmax, int ld, int64_t stride, double tau,
                      int* live, double* dmax, double2* T, double2* Rd, int batch);
// Columns c < ld of Y (n x ld, row-major) with live[c] == 0 refilled with seeded random values
// whose squared column norm is about dmax[0].
void launch_refill_cols(hipStream_t st, double2* Y, int n, int ld, const int* live, const double* dmax,
                        uint64_t seed);
// part[y kc + j] = sum over row chunk y of |G[r][j] - theta_j Z[r][j]|^2.
void launch_resid_cols(hipStream_t st, const double2* G, int ldg, const double2* Z, int ldz, const double* theta,
                       int n, int kc, int rows_per, double* part);

// out[0] = sum over the pairs of conj(x) y (double2); part holds 2 x kRedBlocks doubles of
// partial sums.  Deterministic: a fixed element-to-block assignment and a fixed final order.
constexpr int kRedBlocks = 1024;
void launch_dotc_pairs(hipStream_t st, const double2* x, const double2* y, const DotPair* pairs, int npairs,
                       int64_t total, double* part, double2* out);

// Per-device static scratch of the allocation-free overlap entry point: the pair list
// (kNoAllocMaxPairs DotPair), the partial sums and the result.  noalloc_scratch_init resolves
// its address on the current device (once per device); noalloc_scratch returns it.
constexpr int kNoAllocMaxPairs = 2048;
constexpr size_t kNoAllocPairBytes = (size_t)kNoAllocMaxPairs * sizeof(DotPair);
constexpr size_t kNoAllocPartBytes = (size_t)2 * kRedBlocks * sizeof(double);
constexpr size_t kNoAllocScratchBytes = kNoAllocPairBytes + kNoAllocPartBytes + 256;
void noalloc_scratch_init(int device);
void* noalloc_scratch(int device);

}  // namespace kv
