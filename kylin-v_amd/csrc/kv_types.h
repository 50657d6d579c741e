// Work descriptors shared by the host planners and the gfx950 kernels.  Plain data,
// no HIP types, so the planners build and run on any C++17 host compiler.
#pragma once

#include <cstdint>

namespace kv {

// ---- grouped complex GEMM ---------------------------------------------------------
// C_g[m,n] = alpha * sum_{s in segments of g} opA(A)_s[m,k] opB(B)_s[k,n]
// opA: 0 'N' A[m*lda+k], 1 'T' A[k*lda+m], 2 'C' conj(A[k*lda+m])
// opB: 0 'N' B[k*ldb+n], 1 'T' B[n*ldb+k], 2 'C' conj(B[n*ldb+k])
// C is row-major with leading dimension ldc.  Offsets count complex elements.
struct GemmSeg {
    int64_t a_off, b_off;
    int k, lda, ldb, pad;
};
struct GemmGroup {
    int64_t c_off;
    int m, n, ldc, seg_begin, seg_count, pad;
};
struct GemmTile {
    int group, tm, tn, pad;
};

constexpr int GEMM_BM = 64, GEMM_BN = 64, GEMM_BK = 8;

// Tile shapes of the grouped GEMM (complex outputs per workgroup, K depth per LDS stage).
// Config 0 exists for every op pair; the others for the pairs on the hot path
// (NN, NT, CN, NC).
struct GemmConfig {
    int bm, bn, bk;
};
constexpr int kNumGemmConfigs = 7;
constexpr GemmConfig kGemmConfigs[kNumGemmConfigs] = {
    {64, 64, 8}, {64, 64, 16}, {128, 64, 8}, {64, 128, 8}, {128, 128, 8}, {128, 64, 16}, {64, 128, 16}};

// ---- MPO fold -----------------------------------------------------------------------
// out[out_base + p*out_pstride + q] = sum_t w_t * in[t.in_base + p*t.in_pstride + q]
// for p < P, q < Q.  A task with no terms writes zeros.
struct WTerm {
    int64_t in_base, in_pstride;
    double re, im;
};
struct WTask {
    int64_t out_base, out_pstride;
    int P, Q, term_begin, term_count;
};

// ---- block permutation (tensor transpose) ----------------------------------------------
constexpr int PERM_MAXR = 6;
struct PermTask {
    int64_t in_off, out_off, size;
    int rank, pad;
    int oshape[PERM_MAXR];
    int64_t istride[PERM_MAXR];  // input stride of each OUTPUT axis
};

// ---- 2-D strided copies (sector gather / scatter for the SVD) -------------------------
// dst[dst_off + r*dst_ld + c] = f * src[src_off + r*src_ld + c], f = 1, s[scale_off+r]
// (scale_mode 1) or s[scale_off+c] (scale_mode 2).
struct Copy2D {
    int64_t src_off, dst_off;
    int rows, cols, src_ld, dst_ld, scale_mode, scale_off;
};

// ---- overlap over matching block pairs ------------------------------------------------
// sum over pairs of conj(x[x_off + e]) y[y_off + e], e < size
struct DotPair {
    int64_t x_off, y_off, size;
};

}  // namespace kv
