// gfx950 (CDNA4) kernels of the Holstein TDVP engine.
//
//  * zgemm_grouped: complex GEMM over a device-resident list of U(1) block pairs.  One
//    launch covers every output block of a contraction stage; a group sums its K
//    segments (the block pairs that land on the same output block) inside one tile,
//    so there are no atomics and results are deterministic.  Complex products run as
//    four real MFMAs (re.re - im.im, re.im + im.re) on v_mfma_f64_16x16x4f64, and in
//    FP32 as three (re.re, im.im, (re+im)(re+im): Gauss's product) on
//    v_mfma_f32_16x16x4f32; 64x64 complex tile per 256-thread workgroup, 4 waves in
//    a 2x2 arrangement of 32x32 complex wave tiles, operands staged through LDS.
//  * wfold: applies the sparse MPO tensor as a gather-sum over contiguous slices.
//  * perm / copy2d: block transposes and sector gather/scatter.
//  * FP64 vector algebra for the Lanczos recurrence (deterministic two-pass reductions).
#include <mutex>

#include "kv_kernels.h"
#include "kv_runtime.h"

namespace kv {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class T> struct Cx;
template <> struct Cx<double> {
    using T2 = double2;
    using acc = f64x4;
    __device__ static inline acc mfma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // v_mfma_f64_16x16x4f64: lane l, accumulator item v holds D[l/16 + 4v][l%16]
    // (measured on gfx950 with one-hot operands)
    __device__ static inline int out_row(int lk, int v) { return lk + 4 * v; }
};
template <> struct Cx<float> {
    using T2 = float2;
    using acc = f32x4;
    __device__ static inline acc mfma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // v_mfma_f32_16x16x4f32: lane l, accumulator item v holds D[4(l/16) + v][l%16]
    __device__ static inline int out_row(int lk, int v) { return 4 * lk + v; }
};

template <class T2> struct ScalarOf;
template <> struct ScalarOf<double2> { using T = double; };
template <> struct ScalarOf<float2> { using T = float; };

// ------------------------------------------------------------------------------------
// grouped complex GEMM
// ------------------------------------------------------------------------------------
// Tile BM x BN complex outputs per 256-thread workgroup; 4 waves in a 2x2 arrangement, each
// owning (BM/2) x (BN/2) outputs as 16x16 MFMA sub-tiles; BK complex K per LDS stage.
// Workgroup i runs tile i.  Workgroups are dispatched round-robin over the 8 XCDs and the
// tile list is row-major per group, so tiles that share a B panel (same tile column, rows
// a multiple of 8 tiles apart when the tile-column count is) land on one XCD's L2; a
// remap giving each XCD a contiguous run of tiles measured 20% slower on the H_eff stages.
template <class T, int OPA, int OPB, int BM, int BN, int BK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BM * BN <= 4096 ? 3 : 1))) void zgemm_grouped_kernel(
    const typename Cx<T>::T2* __restrict__ A, const typename Cx<T>::T2* __restrict__ B,
    typename Cx<T>::T2* __restrict__ C, const GemmTile* __restrict__ tiles,
    const GemmGroup* __restrict__ groups, const GemmSeg* __restrict__ segs, T alpha_re,
    T alpha_im) {
    using T2 = typename Cx<T>::T2;
    using acc_t = typename Cx<T>::acc;
    constexpr int SI = BM / 32, SJ = BN / 32;  // 16x16 sub-tiles per wave
    __shared__ T2 As[2][BK][BM + 1];  // double-buffered: one barrier per K stage
    __shared__ T2 Bs[2][BK][BN + 1];

    const GemmTile tile = tiles[blockIdx.x];
    const GemmGroup g = groups[tile.group];
    const int m0 = tile.tm * BM, n0 = tile.tn * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
    const int lr = lane & 15, lk = lane >> 4;

    // FP64: cre = sum re.re - im.im, cim = sum re.im + im.re (four MFMAs per step).
    // FP32: three MFMAs per step (Gauss): cre = sum re.re, cp2 = sum im.im,
    // cim = sum (re + im)(re + im); the epilogue forms re = cre - cp2, im = cim - cre - cp2.
    constexpr bool kGauss = sizeof(T) == sizeof(float);
    acc_t cre[SI][SJ], cim[SI][SJ], cp2[SI][SJ];
#pragma unroll
    for (int i = 0; i < SI; ++i)
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            cre[i][j] = acc_t{0, 0, 0, 0};
            cim[i][j] = acc_t{0, 0, 0, 0};
            cp2[i][j] = acc_t{0, 0, 0, 0};
        }

    // K stages run over (segment, k0) pairs.  The next stage's operands are loaded into
    // registers while the current one is multiplied out of LDS, so global latency hides
    // behind the MFMAs of the same workgroup.
    constexpr int NA = (BM * BK) / 256, NB = (BN * BK) / 256;
    T2 ra[NA], rb[NB];
    // FP32 complex operands load two neighbours along each operand's contiguous index per
    // thread (one 16-byte access when the segment's offset and leading dimension are even).
    constexpr bool kPair = sizeof(T2) == 8;
    auto load = [&](int s, int k0) {
        const GemmSeg sg = segs[g.seg_begin + s];
        const T2* __restrict__ a = A + sg.a_off;
        const T2* __restrict__ b = B + sg.b_off;
        const int K = sg.k;
        if constexpr (kPair) {
            const bool alA = ((sg.a_off | (int64_t)sg.lda) & 1) == 0;
            const bool alB = ((sg.b_off | (int64_t)sg.ldb) & 1) == 0;
#pragma unroll
            for (int r = 0; r < NA / 2; ++r) {
                const int e = tid + r * 256;
                const int mm = (OPA == 0) ? e / (BK / 2) : 2 * (e % (BM / 2));
                const int kk = (OPA == 0) ? 2 * (e % (BK / 2)) : e / (BM / 2);
                const int gm = m0 + mm, gk = k0 + kk;
                T2 v0, v1;
                v0.x = 0;
                v0.y = 0;
                v1 = v0;
                if (OPA == 0) {  // neighbours along k
                    if (gm < g.m) {
                        const T2* pa = a + (int64_t)gm * sg.lda + gk;
                        if (alA && gk + 1 < K) {
                            const float4 q = *reinterpret_cast<const float4*>(pa);
                            v0.x = q.x;
                            v0.y = q.y;
                            v1.x = q.z;
                            v1.y = q.w;
                        } else {
                            if (gk < K) v0 = pa[0];
                            if (gk + 1 < K) v1 = pa[1];
                        }
                    }
                } else {  // neighbours along m
                    if (gk < K) {
                        const T2* pa = a + (int64_t)gk * sg.lda + gm;
                        if (alA && gm + 1 < g.m) {
                            const float4 q = *reinterpret_cast<const float4*>(pa);
                            v0.x = q.x;
                            v0.y = q.y;
                            v1.x = q.z;
                            v1.y = q.w;
                        } else {
                            if (gm < g.m) v0 = pa[0];
                            if (gm + 1 < g.m) v1 = pa[1];
                        }
                    }
                }
                if (OPA == 2) {
                    v0.y = -v0.y;
                    v1.y = -v1.y;
                }
                ra[2 * r] = v0;
                ra[2 * r + 1] = v1;
            }
#pragma unroll
            for (int r = 0; r < NB / 2; ++r) {
                const int e = tid + r * 256;
                const int nn = (OPB == 0) ? 2 * (e % (BN / 2)) : e / (BK / 2);
                const int kk = (OPB == 0) ? e / (BN / 2) : 2 * (e % (BK / 2));
                const int gn = n0 + nn, gk = k0 + kk;
                T2 v0, v1;
                v0.x = 0;
                v0.y = 0;
                v1 = v0;
                if (OPB == 0) {  // neighbours along n
                    if (gk < K) {
                        const T2* pb = b + (int64_t)gk * sg.ldb + gn;
                        if (alB && gn + 1 < g.n) {
                            const float4 q = *reinterpret_cast<const float4*>(pb);
                            v0.x = q.x;
                            v0.y = q.y;
                            v1.x = q.z;
                            v1.y = q.w;
                        } else {
                            if (gn < g.n) v0 = pb[0];
                            if (gn + 1 < g.n) v1 = pb[1];
                        }
                    }
                } else {  // neighbours along k
                    if (gn < g.n) {
                        const T2* pb = b + (int64_t)gn * sg.ldb + gk;
                        if (alB && gk + 1 < K) {
                            const float4 q = *reinterpret_cast<const float4*>(pb);
                            v0.x = q.x;
                            v0.y = q.y;
                            v1.x = q.z;
                            v1.y = q.w;
                        } else {
                            if (gk < K) v0 = pb[0];
                            if (gk + 1 < K) v1 = pb[1];
                        }
                    }
                }
                if (OPB == 2) {
                    v0.y = -v0.y;
                    v1.y = -v1.y;
                }
                rb[2 * r] = v0;
                rb[2 * r + 1] = v1;
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < NA; ++r) {
            const int e = tid + r * 256;
            int mm, kk;
            if (OPA == 0) {
                mm = e / BK;
                kk = e % BK;
            } else {
                kk = e / BM;
                mm = e % BM;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            T2 v;
            v.x = 0;
            v.y = 0;
            if (gm < g.m && gk < K) {
                v = (OPA == 0) ? a[(int64_t)gm * sg.lda + gk] : a[(int64_t)gk * sg.lda + gm];
                if (OPA == 2) v.y = -v.y;
            }
            ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = tid + r * 256;
            int nn, kk;
            if (OPB == 0) {
                kk = e / BN;
                nn = e % BN;
            } else {
                nn = e / BK;
                kk = e % BK;
            }
            const int gn = n0 + nn, gk = k0 + kk;
            T2 v;
            v.x = 0;
            v.y = 0;
            if (gn < g.n && gk < K) {
                v = (OPB == 0) ? b[(int64_t)gk * sg.ldb + gn] : b[(int64_t)gn * sg.ldb + gk];
                if (OPB == 2) v.y = -v.y;
            }
            rb[r] = v;
        }
    };
    auto store = [&](int buf) {
        if constexpr (kPair) {
#pragma unroll
            for (int r = 0; r < NA / 2; ++r) {
                const int e = tid + r * 256;
                const int mm = (OPA == 0) ? e / (BK / 2) : 2 * (e % (BM / 2));
                const int kk = (OPA == 0) ? 2 * (e % (BK / 2)) : e / (BM / 2);
                As[buf][kk][mm] = ra[2 * r];
                if (OPA == 0)
                    As[buf][kk + 1][mm] = ra[2 * r + 1];
                else
                    As[buf][kk][mm + 1] = ra[2 * r + 1];
            }
#pragma unroll
            for (int r = 0; r < NB / 2; ++r) {
                const int e = tid + r * 256;
                const int nn = (OPB == 0) ? 2 * (e % (BN / 2)) : e / (BK / 2);
                const int kk = (OPB == 0) ? e / (BN / 2) : 2 * (e % (BK / 2));
                Bs[buf][kk][nn] = rb[2 * r];
                if (OPB == 0)
                    Bs[buf][kk][nn + 1] = rb[2 * r + 1];
                else
                    Bs[buf][kk + 1][nn] = rb[2 * r + 1];
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < NA; ++r) {
            const int e = tid + r * 256;
            const int mm = (OPA == 0) ? e / BK : e % BM;
            const int kk = (OPA == 0) ? e % BK : e / BM;
            As[buf][kk][mm] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = tid + r * 256;
            const int nn = (OPB == 0) ? e % BN : e / BK;
            const int kk = (OPB == 0) ? e / BN : e % BK;
            Bs[buf][kk][nn] = rb[r];
        }
    };
    int s = 0, k0 = 0, buf = 0;
    while (s < g.seg_count && segs[g.seg_begin + s].k <= 0) ++s;
    if (s < g.seg_count) {
        load(s, 0);
        store(0);
    }
    __syncthreads();
    // iteration t multiplies out of buffer buf while the operands of stage t+1 go into the
    // other one; the single barrier ends both (the other buffer was last read in t-1)
    while (s < g.seg_count) {
        int s2 = s, k2 = k0 + BK;
        if (k2 >= segs[g.seg_begin + s].k) {
            k2 = 0;
            ++s2;
            while (s2 < g.seg_count && segs[g.seg_begin + s2].k <= 0) ++s2;
        }
        if (s2 < g.seg_count) load(s2, k2);
#pragma unroll
        for (int kq = 0; kq < BK; kq += 4) {
            T2 av[SI], bv[SJ];
#pragma unroll
            for (int i = 0; i < SI; ++i) av[i] = As[buf][kq + lk][wm + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < SJ; ++j) bv[j] = Bs[buf][kq + lk][wn + j * 16 + lr];
            if constexpr (kGauss) {
                T as[SI], bs[SJ];
#pragma unroll
                for (int i = 0; i < SI; ++i) as[i] = av[i].x + av[i].y;
#pragma unroll
                for (int j = 0; j < SJ; ++j) bs[j] = bv[j].x + bv[j].y;
#pragma unroll
                for (int i = 0; i < SI; ++i)
#pragma unroll
                    for (int j = 0; j < SJ; ++j) {
                        cre[i][j] = Cx<T>::mfma(av[i].x, bv[j].x, cre[i][j]);
                        cp2[i][j] = Cx<T>::mfma(av[i].y, bv[j].y, cp2[i][j]);
                        cim[i][j] = Cx<T>::mfma(as[i], bs[j], cim[i][j]);
                    }
            } else {
#pragma unroll
                for (int i = 0; i < SI; ++i)
#pragma unroll
                    for (int j = 0; j < SJ; ++j) {
                        cre[i][j] = Cx<T>::mfma(av[i].x, bv[j].x, cre[i][j]);
                        cre[i][j] = Cx<T>::mfma(-av[i].y, bv[j].y, cre[i][j]);
                        cim[i][j] = Cx<T>::mfma(av[i].x, bv[j].y, cim[i][j]);
                        cim[i][j] = Cx<T>::mfma(av[i].y, bv[j].x, cim[i][j]);
                    }
            }
        }
        if (s2 < g.seg_count) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
        s = s2;
        k0 = k2;
    }
    // epilogue: MFMA output layout per type (Cx<T>::out_row)
#pragma unroll
    for (int i = 0; i < SI; ++i)
#pragma unroll
        for (int j = 0; j < SJ; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + Cx<T>::out_row(lk, v);
                const int col = n0 + wn + j * 16 + lr;
                if (row < g.m && col < g.n) {
                    T re = cre[i][j][v], im = cim[i][j][v];
                    if constexpr (kGauss) {
                        const T p1 = re, p2 = cp2[i][j][v];
                        re = p1 - p2;
                        im = im - p1 - p2;
                    }
                    T2 o;
                    o.x = alpha_re * re - alpha_im * im;
                    o.y = alpha_re * im + alpha_im * re;
                    C[g.c_off + (int64_t)row * g.ldc + col] = o;
                }
            }
}

template <class T2>
void launch_gemm(hipStream_t st, int opA, int opB, const T2* A, const T2* B, T2* C,
                 const GemmTile* tiles, int ntiles, const GemmGroup* groups,
                 const GemmSeg* segs, double alpha_re, double alpha_im, int cfg) {
    using T = typename ScalarOf<T2>::T;
    if (ntiles <= 0) return;
    const dim3 grid(ntiles), block(256);
    const T ar = (T)alpha_re, ai = (T)alpha_im;
#define KV_GEMM_CASE(C_, OA, OB, BM_, BN_, BK_)                                                  \
    if (cfg == C_ && opA == OA && opB == OB) {                                                 \
        hipLaunchKernelGGL((zgemm_grouped_kernel<T, OA, OB, BM_, BN_, BK_>), grid, block, 0, st, \
                           A, B, C, tiles, groups, segs, ar, ai);                              \
        KV_HIP(hipGetLastError());                                                             \
        return;                                                                                \
    }
#define KV_GEMM_ALL_OPS(C_, BM_, BN_, BK_)  \
    KV_GEMM_CASE(C_, 0, 0, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 0, 1, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 0, 2, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 1, 0, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 1, 1, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 1, 2, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 2, 0, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 2, 1, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 2, 2, BM_, BN_, BK_)
#define KV_GEMM_HOT_OPS(C_, BM_, BN_, BK_)  \
    KV_GEMM_CASE(C_, 0, 0, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 0, 1, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 2, 0, BM_, BN_, BK_)   \
    KV_GEMM_CASE(C_, 0, 2, BM_, BN_, BK_)
    KV_GEMM_ALL_OPS(0, 64, 64, 8)
    KV_GEMM_HOT_OPS(1, 64, 64, 16)
    KV_GEMM_HOT_OPS(2, 128, 64, 8)
    KV_GEMM_HOT_OPS(3, 64, 128, 8)
    if constexpr (sizeof(T) == sizeof(float)) {  // the FP64 128x128 tile spills registers
        KV_GEMM_HOT_OPS(4, 128, 128, 8)
    }
    KV_GEMM_HOT_OPS(5, 128, 64, 16)
    KV_GEMM_HOT_OPS(6, 64, 128, 16)
#undef KV_GEMM_HOT_OPS
#undef KV_GEMM_ALL_OPS
#undef KV_GEMM_CASE
    if (cfg != 0) {  // op pair without a specialised tile: fall back to the 64x64 tile
        throw Error(ERR_INTERNAL, "launch_gemm: tile config not built for this op pair");
    }
    throw Error(ERR_ARG, "launch_gemm: bad op codes");
}

bool gemm_config_supports(int cfg, int opA, int opB, bool fp64) {
    if (cfg == 0) return opA >= 0 && opA <= 2 && opB >= 0 && opB <= 2;
    if (cfg == 4 && fp64) return false;
    return (opA == 0 && opB == 0) || (opA == 0 && opB == 1) || (opA == 2 && opB == 0) ||
           (opA == 0 && opB == 2);
}

template void launch_gemm<double2>(hipStream_t, int, int, const double2*, const double2*,
                                   double2*, const GemmTile*, int, const GemmGroup*,
                                   const GemmSeg*, double, double, int);
template void launch_gemm<float2>(hipStream_t, int, int, const float2*, const float2*,
                                  float2*, const GemmTile*, int, const GemmGroup*,
                                  const GemmSeg*, double, double, int);

// ------------------------------------------------------------------------------------
// grouped complex GEMM, split-BF16 (FP32 operands and accumulation)
// ------------------------------------------------------------------------------------
// Every FP32 operand x is split into x_hi = bf16(x) and x_lo = bf16(x - x_hi) while it is
// staged into LDS; a real product is x_hi*y_hi + x_hi*y_lo + x_lo*y_hi on
// v_mfma_f32_16x16x32_bf16 (the lo*lo term, ~2^-16 relative, is dropped).  A complex
// product is four real products, i.e. 12 MFMAs per 16x16x32 complex step.  Same tiles,
// groups and K segments as zgemm_grouped_kernel.
// LDS holds each operand as four planes (re_hi, re_lo, im_hi, im_lo), K contiguous per row
// of A and per column of B, so a lane's eight K values are one 16-byte LDS read:
// v_mfma_f32_16x16x32_bf16 lane l takes A[l%16][8(l/16)..8(l/16)+7] and
// B[8(l/16)..8(l/16)+7][l%16]; accumulator item v holds D[4(l/16)+v][l%16].
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline void split2(float x0, float x1, bf16x2& hi, bf16x2& lo) {
    const __bf16 h0 = (__bf16)x0, h1 = (__bf16)x1;
    hi = bf16x2{h0, h1};
    lo = bf16x2{(__bf16)(x0 - (float)h0), (__bf16)(x1 - (float)h1)};
}

// NT threads = 64 * (BM / (16 SI)) * (BN / (16 SJ)); each wave owns SI x SJ sub-tiles of
// 16x16 outputs.  The im.im product is accumulated with the B imaginary planes sign-
// flipped (exact in bf16), so each sub-tile keeps two accumulators (re, im).
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ inline bf16x8 bf16_neg(bf16x8 v) {
    return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u16x8, v) ^ (u16x8)(unsigned short)0x8000);
}

template <int OPA, int OPB, int BM, int BN, int SI, int SJ>
__global__ __launch_bounds__(64 * (BM / (16 * SI)) * (BN / (16 * SJ))) void zgemm_bf16x3_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C,
    const GemmTile* __restrict__ tiles, const GemmGroup* __restrict__ groups,
    const GemmSeg* __restrict__ segs, float alpha_re, float alpha_im) {
    constexpr int BK = 32, LDK = BK + 8;  // row pitch stays a multiple of 16 bytes
    constexpr int WN = BN / (16 * SJ), NT = 64 * (BM / (16 * SI)) * WN;
    __shared__ __attribute__((aligned(16))) __bf16 As[4][BM][LDK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[4][BN][LDK];

    const GemmTile tile = tiles[blockIdx.x];
    const GemmGroup g = groups[tile.group];
    const int m0 = tile.tm * BM, n0 = tile.tn * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WN) * (16 * SI), wn = (wave % WN) * (16 * SJ);
    const int lr = lane & 15, lk = lane >> 4;

    f32x4 cre[SI][SJ], cim[SI][SJ];
#pragma unroll
    for (int i = 0; i < SI; ++i)
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            cre[i][j] = f32x4{0, 0, 0, 0};
            cim[i][j] = f32x4{0, 0, 0, 0};
        }

    // K stages run over (segment, k0) pairs.  The raw FP32 operands of the next stage are
    // loaded into registers while the current one is multiplied out of LDS; they are split
    // into the bf16 planes and stored once every wave has finished reading.
    constexpr int RA = (BM * BK / 2) / NT, RB = (BN * BK / 2) / NT;
    float2 ra[RA][2], rb[RB][2];
    auto load = [&](int s, int k0) {
        const GemmSeg sg = segs[g.seg_begin + s];
        const float2* __restrict__ a = A + sg.a_off;
        const float2* __restrict__ b = B + sg.b_off;
        const int K = sg.k;
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int e = tid + r * NT;
            int mm, kk;
            if (OPA == 0) {
                mm = e / (BK / 2);
                kk = 2 * (e % (BK / 2));
            } else {
                mm = e % BM;
                kk = 2 * (e / BM);
            }
            const int gm = m0 + mm, gk = k0 + kk;
            float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
            if (gm < g.m) {
                if (OPA == 0) {
                    const float2* row = a + (int64_t)gm * sg.lda;
                    if (gk < K) v0 = row[gk];
                    if (gk + 1 < K) v1 = row[gk + 1];
                } else {
                    if (gk < K) v0 = a[(int64_t)gk * sg.lda + gm];
                    if (gk + 1 < K) v1 = a[(int64_t)(gk + 1) * sg.lda + gm];
                }
            }
            if (OPA == 2) {
                v0.y = -v0.y;
                v1.y = -v1.y;
            }
            ra[r][0] = v0;
            ra[r][1] = v1;
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int e = tid + r * NT;
            int nn, kk;
            if (OPB == 0) {
                nn = e % BN;
                kk = 2 * (e / BN);
            } else {
                nn = e / (BK / 2);
                kk = 2 * (e % (BK / 2));
            }
            const int gn = n0 + nn, gk = k0 + kk;
            float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
            if (gn < g.n) {
                if (OPB == 0) {
                    if (gk < K) v0 = b[(int64_t)gk * sg.ldb + gn];
                    if (gk + 1 < K) v1 = b[(int64_t)(gk + 1) * sg.ldb + gn];
                } else {
                    const float2* col = b + (int64_t)gn * sg.ldb;
                    if (gk < K) v0 = col[gk];
                    if (gk + 1 < K) v1 = col[gk + 1];
                }
            }
            if (OPB == 2) {
                v0.y = -v0.y;
                v1.y = -v1.y;
            }
            rb[r][0] = v0;
            rb[r][1] = v1;
        }
    };
    auto store = [&] {
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int e = tid + r * NT;
            const int mm = (OPA == 0) ? e / (BK / 2) : e % BM;
            const int kk = (OPA == 0) ? 2 * (e % (BK / 2)) : 2 * (e / BM);
            bf16x2 h, l;
            split2(ra[r][0].x, ra[r][1].x, h, l);
            *(bf16x2*)&As[0][mm][kk] = h;
            *(bf16x2*)&As[1][mm][kk] = l;
            split2(ra[r][0].y, ra[r][1].y, h, l);
            *(bf16x2*)&As[2][mm][kk] = h;
            *(bf16x2*)&As[3][mm][kk] = l;
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int e = tid + r * NT;
            const int nn = (OPB == 0) ? e % BN : e / (BK / 2);
            const int kk = (OPB == 0) ? 2 * (e / BN) : 2 * (e % (BK / 2));
            bf16x2 h, l;
            split2(rb[r][0].x, rb[r][1].x, h, l);
            *(bf16x2*)&Bs[0][nn][kk] = h;
            *(bf16x2*)&Bs[1][nn][kk] = l;
            split2(rb[r][0].y, rb[r][1].y, h, l);
            *(bf16x2*)&Bs[2][nn][kk] = h;
            *(bf16x2*)&Bs[3][nn][kk] = l;
        }
    };
    int s = 0, k0 = 0;
    while (s < g.seg_count && segs[g.seg_begin + s].k <= 0) ++s;
    if (s < g.seg_count) {
        load(s, 0);
        store();
    }
    __syncthreads();
    while (s < g.seg_count) {
        int s2 = s, k2 = k0 + BK;
        if (k2 >= segs[g.seg_begin + s].k) {
            k2 = 0;
            ++s2;
            while (s2 < g.seg_count && segs[g.seg_begin + s2].k <= 0) ++s2;
        }
        if (s2 < g.seg_count) load(s2, k2);
        {
            bf16x8 af[SI][4], bf[SJ][4], bn[SJ][2];
#pragma unroll
            for (int i = 0; i < SI; ++i)
#pragma unroll
                for (int p = 0; p < 4; ++p)
                    af[i][p] = *(const bf16x8*)&As[p][wm + i * 16 + lr][8 * lk];
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
#pragma unroll
                for (int p = 0; p < 4; ++p)
                    bf[j][p] = *(const bf16x8*)&Bs[p][wn + j * 16 + lr][8 * lk];
                bn[j][0] = bf16_neg(bf[j][2]);
                bn[j][1] = bf16_neg(bf[j][3]);
            }
#pragma unroll
            for (int i = 0; i < SI; ++i)
#pragma unroll
                for (int j = 0; j < SJ; ++j) {
                    f32x4 re = cre[i][j], im = cim[i][j];
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bf[j][0], re, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bf[j][1], re, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][1], bf[j][0], re, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][2], bn[j][0], re, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][2], bn[j][1], re, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][3], bn[j][0], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bf[j][2], im, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][0], bf[j][3], im, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][1], bf[j][2], im, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][2], bf[j][0], im, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][2], bf[j][1], im, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][3], bf[j][0], im, 0, 0, 0);
                    cre[i][j] = re;
                    cim[i][j] = im;
                }
        }
        __syncthreads();  // every wave has read this stage's planes
        if (s2 < g.seg_count) {
            store();
            __syncthreads();
        }
        s = s2;
        k0 = k2;
    }
#pragma unroll
    for (int i = 0; i < SI; ++i)
#pragma unroll
        for (int j = 0; j < SJ; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + 4 * lk + v;
                const int col = n0 + wn + j * 16 + lr;
                if (row < g.m && col < g.n) {
                    const float re = cre[i][j][v], im = cim[i][j][v];
                    float2 o;
                    o.x = alpha_re * re - alpha_im * im;
                    o.y = alpha_re * im + alpha_im * re;
                    C[g.c_off + (int64_t)row * g.ldc + col] = o;
                }
            }
}

// cfg 0: 64x64 tile (256 threads); cfg 4: 128x128 tile (512 threads), same tile lists as
// the FP32 kernel's configs 0 and 4.
void launch_gemm_bf16x3(hipStream_t st, int opA, int opB, const float2* A, const float2* B,
                        float2* C, const GemmTile* tiles, int ntiles, const GemmGroup* groups,
                        const GemmSeg* segs, double alpha_re, double alpha_im, int cfg) {
    if (ntiles <= 0) return;
    const float ar = (float)alpha_re, ai = (float)alpha_im;
#define KV_BF_CASE(OA, OB)                                                                     \
    if (opA == OA && opB == OB) {                                                              \
        if (cfg == 4)                                                                          \
            hipLaunchKernelGGL((zgemm_bf16x3_kernel<OA, OB, 128, 128, 4, 2>), dim3(ntiles),     \
                               dim3(512), 0, st, A, B, C, tiles, groups, segs, ar, ai);        \
        else                                                                                   \
            hipLaunchKernelGGL((zgemm_bf16x3_kernel<OA, OB, 64, 64, 2, 2>), dim3(ntiles),       \
                               dim3(256), 0, st, A, B, C, tiles, groups, segs, ar, ai);        \
        KV_HIP(hipGetLastError());                                                             \
        return;                                                                                \
    }
    if (cfg != 0 && cfg != 4) throw Error(ERR_ARG, "launch_gemm_bf16x3: tile config 0 or 4");
    KV_BF_CASE(0, 0)
    KV_BF_CASE(0, 1)
    KV_BF_CASE(0, 2)
    KV_BF_CASE(1, 0)
    KV_BF_CASE(1, 1)
    KV_BF_CASE(1, 2)
    KV_BF_CASE(2, 0)
    KV_BF_CASE(2, 1)
    KV_BF_CASE(2, 2)
#undef KV_BF_CASE
    throw Error(ERR_ARG, "launch_gemm_bf16x3: bad op codes");
}

// ------------------------------------------------------------------------------------
// MPO fold
// ------------------------------------------------------------------------------------
// One workgroup column per task (an output (s', w') row of a block, P x Q elements), the
// task's P x Q elements spread over all threads of gridDim.y workgroups in row-major order
// (consecutive lanes on consecutive q, so every lane works even when Q < 256); the terms
// are summed in order, four loads issued together.
template <class T2>
__global__ __launch_bounds__(256) void wfold_kernel(const T2* __restrict__ in,
                                                    T2* __restrict__ out,
                                                    const WTask* __restrict__ tasks,
                                                    const WTerm* __restrict__ terms) {
    using T = typename ScalarOf<T2>::T;
    const WTask tk = tasks[blockIdx.x];
    const WTerm* __restrict__ tt = terms + tk.term_begin;
    const int nt = tk.term_count, Q = tk.Q;
    const int PQ = tk.P * Q;
    if constexpr (sizeof(T2) == 8) {
        // FP32 complex: two neighbouring q per thread as one 16-byte load per term, when every
        // offset and stride of the task is even (the addresses are then 16-byte aligned)
        int64_t odd = tk.out_base | tk.out_pstride | (int64_t)Q;
        for (int t = 0; t < nt; ++t) odd |= tt[t].in_base | tt[t].in_pstride;
        if ((odd & 1) == 0) {
            const int Q2 = Q >> 1, PQ2 = PQ >> 1;
            for (int e = blockIdx.y * 256 + threadIdx.x; e < PQ2; e += gridDim.y * 256) {
                const int p = e / Q2, q = 2 * (e - p * Q2);
                float re0 = 0.f, im0 = 0.f, re1 = 0.f, im1 = 0.f;
#pragma unroll 4
                for (int t = 0; t < nt; ++t) {
                    const WTerm w = tt[t];
                    const float4 v = *reinterpret_cast<const float4*>(in + w.in_base + (int64_t)p * w.in_pstride + q);
                    const float wr = (float)w.re, wi = (float)w.im;
                    re0 += wr * v.x - wi * v.y;
                    im0 += wr * v.y + wi * v.x;
                    re1 += wr * v.z - wi * v.w;
                    im1 += wr * v.w + wi * v.z;
                }
                *reinterpret_cast<float4*>(out + tk.out_base + (int64_t)p * tk.out_pstride + q) =
                    make_float4(re0, im0, re1, im1);
            }
            return;
        }
    }
    for (int e = blockIdx.y * 256 + threadIdx.x; e < PQ; e += gridDim.y * 256) {
        const int p = e / Q, q = e - p * Q;
        T re = 0, im = 0;
        int t = 0;
        for (; t + 4 <= nt; t += 4) {
            T2 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = in[tt[t + u].in_base + (int64_t)p * tt[t + u].in_pstride + q];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T wr = (T)tt[t + u].re, wi = (T)tt[t + u].im;
                re += wr * v[u].x - wi * v[u].y;
                im += wr * v[u].y + wi * v[u].x;
            }
        }
        for (; t < nt; ++t) {
            const WTerm w = tt[t];
            const T2 v = in[w.in_base + (int64_t)p * w.in_pstride + q];
            const T wr = (T)w.re, wi = (T)w.im;
            re += wr * v.x - wi * v.y;
            im += wr * v.y + wi * v.x;
        }
        T2 o;
        o.x = re;
        o.y = im;
        out[tk.out_base + (int64_t)p * tk.out_pstride + q] = o;
    }
}

template <class T2>
void launch_wfold(hipStream_t st, const T2* in, T2* out, const WTask* tasks, int ntasks,
                  const WTerm* terms, int64_t maxPQ) {
    if (ntasks <= 0) return;
    int64_t gy = (maxPQ + 255) / 256;
    if (gy > 64) gy = 64;
    if (gy < 1) gy = 1;
    hipLaunchKernelGGL((wfold_kernel<T2>), dim3(ntasks, (unsigned)gy), dim3(256), 0, st, in, out, tasks,
                       terms);
    KV_HIP(hipGetLastError());
}
template void launch_wfold<double2>(hipStream_t, const double2*, double2*, const WTask*, int,
                                    const WTerm*, int64_t);
template void launch_wfold<float2>(hipStream_t, const float2*, float2*, const WTask*, int,
                                   const WTerm*, int64_t);

// ------------------------------------------------------------------------------------
// block permutation
// ------------------------------------------------------------------------------------
template <class T2>
__global__ __launch_bounds__(256) void perm_kernel(const T2* __restrict__ in,
                                                   T2* __restrict__ out,
                                                   const PermTask* __restrict__ tasks) {
    // fields are read through the pointer: a by-value copy indexed by r would live in
    // scratch memory
    const PermTask* tk = tasks + blockIdx.x;
    const int64_t size = tk->size, in_off = tk->in_off, out_off = tk->out_off;
    const int rank = tk->rank;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < size;
         e += (int64_t)gridDim.y * blockDim.x) {
        int64_t rem = e, src = 0;
        for (int r = rank - 1; r >= 0; --r) {
            const int64_t sh = tk->oshape[r];
            const int64_t idx = rem % sh;
            rem /= sh;
            src += idx * tk->istride[r];
        }
        out[out_off + e] = in[in_off + src];
    }
}

template <class T2>
void launch_perm(hipStream_t st, const T2* in, T2* out, const PermTask* tasks, int ntasks,
                 int64_t maxsize) {
    if (ntasks <= 0 || maxsize <= 0) return;
    int64_t gy = (maxsize + 255) / 256;
    if (gy > 1024) gy = 1024;
    hipLaunchKernelGGL((perm_kernel<T2>), dim3(ntasks, (unsigned)gy), dim3(256), 0, st, in,
                       out, tasks);
    KV_HIP(hipGetLastError());
}
template void launch_perm<double2>(hipStream_t, const double2*, double2*, const PermTask*, int,
                                   int64_t);
template void launch_perm<float2>(hipStream_t, const float2*, float2*, const PermTask*, int,
                                  int64_t);

// ------------------------------------------------------------------------------------
// 2-D strided copies with optional diagonal scaling
// ------------------------------------------------------------------------------------
template <class T2>
__global__ __launch_bounds__(256) void copy2d_kernel(const T2* __restrict__ src,
                                                     T2* __restrict__ dst,
                                                     const Copy2D* __restrict__ cps,
                                                     const double* __restrict__ scale) {
    using T = typename ScalarOf<T2>::T;
    const Copy2D c = cps[blockIdx.x];
    for (int r = blockIdx.y; r < c.rows; r += gridDim.y) {
        for (int q = threadIdx.x; q < c.cols; q += blockDim.x) {
            T2 v = src[c.src_off + (int64_t)r * c.src_ld + q];
            if (c.scale_mode == 1) {
                const T f = (T)scale[c.scale_off + r];
                v.x *= f;
                v.y *= f;
            } else if (c.scale_mode == 2) {
                const T f = (T)scale[c.scale_off + q];
                v.x *= f;
                v.y *= f;
            }
            dst[c.dst_off + (int64_t)r * c.dst_ld + q] = v;
        }
    }
}

template <class T2>
void launch_copy2d(hipStream_t st, const T2* src, T2* dst, const Copy2D* cps, int ncopies,
                   int maxrows, const double* scale) {
    if (ncopies <= 0 || maxrows <= 0) return;
    const int gy = maxrows < 256 ? maxrows : 256;
    hipLaunchKernelGGL((copy2d_kernel<T2>), dim3(ncopies, gy), dim3(256), 0, st, src, dst, cps,
                       scale);
    KV_HIP(hipGetLastError());
}
template void launch_copy2d<double2>(hipStream_t, const double2*, double2*, const Copy2D*, int,
                                     int, const double*);

// dst[c*ld_dst + r] = conj(src[r*ld_src + c]) for r < rows, c < cols (32x32 LDS tiles)
__global__ __launch_bounds__(256) void ctranspose_kernel(const double2* __restrict__ src,
                                                         double2* __restrict__ dst, int rows,
                                                         int cols, int ld_src, int ld_dst) {
    __shared__ double2 t[32][33];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + tx;
        if (r < rows && c < cols) t[i][tx] = src[(int64_t)r * ld_src + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols) {
            const double2 v = t[tx][i];
            dst[(int64_t)c * ld_dst + r] = make_double2(v.x, -v.y);
        }
    }
}

void launch_ctranspose(hipStream_t st, const double2* src, double2* dst, int rows, int cols,
                       int ld_src, int ld_dst) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(ctranspose_kernel, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0,
                       st, src, dst, rows, cols, ld_src, ld_dst);
    KV_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------
// FP64 vector algebra
// ------------------------------------------------------------------------------------
constexpr int RED_BLOCKS = kRedBlocks;

__global__ __launch_bounds__(256) void zero_kernel(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
        p[e] = 0u;
}

// Zero fill by a kernel (not hipMemsetAsync), so the call stays safe inside captured graphs.
void launch_zero(hipStream_t st, void* p, size_t bytes) {
    if (!bytes) return;
    const int64_t n = (int64_t)(bytes / 4);
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)b), dim3(256), 0, st, (uint32_t*)p, n);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void dotc_partial_kernel(const double2* __restrict__ a,
                                                           const double2* __restrict__ b,
                                                           int64_t n, double* __restrict__ part) {
    __shared__ double sr[256], si[256];
    double re = 0, im = 0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const double2 x = a[e], y = b[e];
        re += x.x * y.x + x.y * y.y;  // conj(x) * y
        im += x.x * y.y - x.y * y.x;
    }
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sr[0];
        part[2 * blockIdx.x + 1] = si[0];
    }
}

__global__ __launch_bounds__(256) void reduce2_final_kernel(const double* __restrict__ part,
                                                            int nb, double* __restrict__ out,
                                                            int ncomp) {
    __shared__ double s0[256], s1[256];
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        a += part[ncomp * i];
        if (ncomp == 2) b += part[ncomp * i + 1];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s0[threadIdx.x] += s0[threadIdx.x + s];
            s1[threadIdx.x] += s1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        if (ncomp == 2) out[1] = s1[0];
    }
}

static int red_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > RED_BLOCKS) b = RED_BLOCKS;
    if (b < 1) b = 1;
    return (int)b;
}

void launch_dotc(hipStream_t st, const double2* a, const double2* b, int64_t n, double2* out,
                 double* scratch) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(dotc_partial_kernel, dim3(nb), dim3(256), 0, st, a, b, n, scratch);
    KV_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce2_final_kernel, dim3(1), dim3(256), 0, st, scratch, nb,
                       reinterpret_cast<double*>(out), 2);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void nrm2sq_partial_kernel(const double2* __restrict__ a,
                                                             int64_t n, double* __restrict__ part) {
    __shared__ double sr[256];
    double re = 0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const double2 x = a[e];
        re += x.x * x.x + x.y * x.y;
    }
    sr[threadIdx.x] = re;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sr[threadIdx.x] += sr[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sr[0];
}

void launch_nrm2sq(hipStream_t st, const double2* a, int64_t n, double* out, double* scratch) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(nrm2sq_partial_kernel, dim3(nb), dim3(256), 0, st, a, n, scratch);
    KV_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce2_final_kernel, dim3(1), dim3(256), 0, st, scratch, nb, out, 1);
    KV_HIP(hipGetLastError());
}

static unsigned ew_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (unsigned)b;
}

__global__ __launch_bounds__(256) void axpy_neg_dev_kernel(double2* __restrict__ y,
                                                           const double2* __restrict__ x,
                                                           const double2* __restrict__ c,
                                                           int64_t n) {
    const double2 cc = c[0];
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const double2 v = x[e];
        double2 o = y[e];
        o.x -= cc.x * v.x - cc.y * v.y;
        o.y -= cc.x * v.y + cc.y * v.x;
        y[e] = o;
    }
}

void launch_axpy_neg_dev(hipStream_t st, double2* y, const double2* x, const double2* c,
                         int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(axpy_neg_dev_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, y, x, c, n);
    KV_HIP(hipGetLastError());
}

// Lanczos passes fused with the subtraction that precedes them: y -= c[0] x, written back,
// then per-workgroup partial sums of conj(y) z (DOT) or of |y|^2 for reduce2_final_kernel.
// Element formulas, the element-to-block assignment and the reduction tree are those of
// axpy_neg_dev_kernel followed by dotc_partial_kernel / nrm2sq_partial_kernel.
template <bool DOT>
__global__ __launch_bounds__(256) void axpy_reduce_partial_kernel(double2* __restrict__ y,
                                                                  const double2* __restrict__ x,
                                                                  const double2* __restrict__ c,
                                                                  const double2* __restrict__ z, int64_t n,
                                                                  double* __restrict__ part) {
    __shared__ double sr[256], si[256];
    const double2 cc = c[0];
    double re = 0, im = 0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const double2 v = x[e];
        double2 o = y[e];
        o.x -= cc.x * v.x - cc.y * v.y;
        o.y -= cc.x * v.y + cc.y * v.x;
        y[e] = o;
        if constexpr (DOT) {
            const double2 w = z[e];
            re += o.x * w.x + o.y * w.y;  // conj(o) * w
            im += o.x * w.y - o.y * w.x;
        } else {
            re += o.x * o.x + o.y * o.y;
        }
    }
    sr[threadIdx.x] = re;
    if constexpr (DOT) si[threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            if constexpr (DOT) si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if constexpr (DOT) {
            part[2 * blockIdx.x] = sr[0];
            part[2 * blockIdx.x + 1] = si[0];
        } else {
            part[blockIdx.x] = sr[0];
        }
    }
}

void launch_axpy_dotc(hipStream_t st, double2* y, const double2* x, const double2* c, const double2* z, int64_t n,
                      double2* out, double* scratch) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(axpy_reduce_partial_kernel<true>, dim3(nb), dim3(256), 0, st, y, x, c, z, n, scratch);
    KV_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce2_final_kernel, dim3(1), dim3(256), 0, st, scratch, nb, reinterpret_cast<double*>(out),
                       2);
    KV_HIP(hipGetLastError());
}

void launch_axpy_nrm2sq(hipStream_t st, double2* y, const double2* x, const double2* c, int64_t n, double* out,
                        double* scratch) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(axpy_reduce_partial_kernel<false>, dim3(nb), dim3(256), 0, st, y, x, c, nullptr, n, scratch);
    KV_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce2_final_kernel, dim3(1), dim3(256), 0, st, scratch, nb, out, 1);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void scale_invsqrt_dev_kernel(double2* __restrict__ dst,
                                                                const double2* __restrict__ src,
                                                                const double* __restrict__ nsq,
                                                                int64_t n) {
    const double f = 1.0 / sqrt(nsq[0]);
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        double2 v = src[e];
        v.x *= f;
        v.y *= f;
        dst[e] = v;
    }
}

void launch_scale_invsqrt_dev(hipStream_t st, double2* dst, const double2* src,
                              const double* nsq, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(scale_invsqrt_dev_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, dst, src,
                       nsq, n);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void scale_kernel(double2* __restrict__ dst,
                                                    const double2* __restrict__ src, double s,
                                                    int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        double2 v = src[e];
        v.x *= s;
        v.y *= s;
        dst[e] = v;
    }
}

void launch_scale(hipStream_t st, double2* dst, const double2* src, double s, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(scale_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, dst, src, s, n);
    KV_HIP(hipGetLastError());
}

// out = a*x + b*y elementwise (complex a, b); x or y may be null (term dropped).
__global__ __launch_bounds__(256) void zaxpby_kernel(double2* __restrict__ out,
                                                     const double2* __restrict__ x,
                                                     const double2* __restrict__ y, double2 a,
                                                     double2 b, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        double re = 0, im = 0;
        if (x) {
            const double2 v = x[e];
            re += a.x * v.x - a.y * v.y;
            im += a.x * v.y + a.y * v.x;
        }
        if (y) {
            const double2 v = y[e];
            re += b.x * v.x - b.y * v.y;
            im += b.x * v.y + b.y * v.x;
        }
        out[e] = make_double2(re, im);
    }
}

void launch_zaxpby(hipStream_t st, double2* out, const double2* x, const double2* y, double2 a,
                   double2 b, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zaxpby_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, out, x, y, a, b, n);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void lincomb_kernel(double2* __restrict__ out,
                                                      const double2* __restrict__ V,
                                                      const double2* __restrict__ c, int nv,
                                                      int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        double re = 0, im = 0;
        for (int i = 0; i < nv; ++i) {
            const double2 v = V[(int64_t)i * n + e];
            const double2 ci = c[i];
            re += ci.x * v.x - ci.y * v.y;
            im += ci.x * v.y + ci.y * v.x;
        }
        out[e] = make_double2(re, im);
    }
}

void launch_lincomb(hipStream_t st, double2* out, const double2* V, const double2* c, int nv,
                    int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(lincomb_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, out, V, c, nv, n);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void f64_to_f32_kernel(const double2* __restrict__ s,
                                                         float2* __restrict__ d, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const double2 v = s[e];
        d[e] = make_float2((float)v.x, (float)v.y);
    }
}

void launch_f64_to_f32(hipStream_t st, const double2* src, float2* dst, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, src, dst, n);
    KV_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void f32_to_f64_kernel(const float2* __restrict__ s,
                                                         double2* __restrict__ d, double f,
                                                         int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const float2 v = s[e];
        d[e] = make_double2(f * (double)v.x, f * (double)v.y);
    }
}

void launch_f32_to_f64(hipStream_t st, const float2* src, double2* dst, double s, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, src, dst, s, n);
    KV_HIP(hipGetLastError());
}

// A[i ld + i] = (v, 0) for i in [from, to): the diagonal of a square matrix (either order).
__global__ __launch_bounds__(256) void set_diag_kernel(double2* __restrict__ A, int64_t ld, int from,
                                                       int to, double v) {
    for (int i = from + blockIdx.x * 256 + threadIdx.x; i < to; i += gridDim.x * 256)
        A[(int64_t)i * ld + i] = make_double2(v, 0.0);
}

void launch_set_diag(hipStream_t st, double2* A, int64_t ld, int from, int to, double v) {
    if (to <= from) return;
    hipLaunchKernelGGL(set_diag_kernel, dim3(ew_blocks(to - from)), dim3(256), 0, st, A, ld, from, to, v);
    KV_HIP(hipGetLastError());
}

// dst (n x n, row-major, ld n) = one triangle of src (ld lds), the other triangle zero:
// dst[i][j] = src[i * lds + j] where j <= i (lower) or j >= i (upper).
__global__ __launch_bounds__(256) void tri_copy_kernel(const double2* __restrict__ src, int64_t lds,
                                                       double2* __restrict__ dst, int n, int lower) {
    const int64_t total = (int64_t)n * n;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int i = (int)(e / n), j = (int)(e % n);
        const bool keep = lower ? j <= i : j >= i;
        dst[e] = keep ? src[(int64_t)i * lds + j] : make_double2(0.0, 0.0);
    }
}

void launch_tri_copy(hipStream_t st, const double2* src, int64_t lds, double2* dst, int n, bool lower) {
    if (n <= 0) return;
    hipLaunchKernelGGL(tri_copy_kernel, dim3(ew_blocks((int64_t)n * n)), dim3(256), 0, st, src, lds, dst, n,
                       lower ? 1 : 0);
    KV_HIP(hipGetLastError());
}

// ---- back-transformation of a Householder tridiagonalisation ------------------------
// Lower-triangle reduction Q^H A Q = T of a column-major n x n matrix: reflector i < n-1 is
// H_i = I - tau_i v_i v_i^H with v_i[r] = 0 for r <= i, 1 at r = i + 1 and A[i n + r] below.
// V (column-major, n rows, ncols >= n - 1 columns) receives every v_i explicitly; columns
// from n - 1 on are zero.
__global__ __launch_bounds__(256) void hh_explicit_kernel(const double2* __restrict__ A,
                                                          double2* __restrict__ V, int n,
                                                          int ncols) {
    const int64_t total = (int64_t)n * ncols;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int i = (int)(e / n), r = (int)(e % n);
        double2 v = make_double2(0.0, 0.0);
        if (i < n - 1) {
            if (r == i + 1)
                v.x = 1.0;
            else if (r > i + 1)
                v = A[e];
        }
        V[e] = v;
    }
}

void launch_hh_explicit(hipStream_t st, const double2* A, double2* V, int n, int ncols) {
    const int64_t total = (int64_t)n * ncols;
    if (total <= 0) return;
    hipLaunchKernelGGL(hh_explicit_kernel, dim3(ew_blocks(total)), dim3(256), 0, st, A, V, n, ncols);
    KV_HIP(hipGetLastError());
}

// Triangular factors of blocks of bw consecutive reflectors (forward, columnwise):
// H_{i0} ... H_{i0+w-1} = I - V_b T_b V_b^H with T_b(c,c) = tau_c and
// T_b(0:c, c) = -tau_c T_b(0:c, 0:c) S_b(0:c, c), S_b = V_b^H V_b.  S and T hold one bw x bw
// column-major matrix per block; entries of columns c >= w (past the last reflector) are 0.
// One workgroup per block; the upper triangle of T_b lives packed in LDS (column j at
// j (j + 1) / 2).  Each new column is a triangular matrix-vector product: row r is summed
// by kTfSplit threads over interleaved j, partial sums combined through LDS.  The next
// column of S is loaded into registers while the current column is formed, so the global
// load is off the column-to-column chain; two barriers per column.
constexpr int kTfactorMax = 128, kTfSplit = 4;
__global__ __launch_bounds__(kTfactorMax * kTfSplit) void hh_tfactor_kernel(
    const double2* __restrict__ S, const double2* __restrict__ tau, double2* __restrict__ T, int bw,
    int nref) {
    __shared__ double2 tp[kTfactorMax * (kTfactorMax + 1) / 2];
    __shared__ double2 sc[kTfactorMax], tl[kTfactorMax];
    __shared__ double2 part[kTfSplit][kTfactorMax];
    const int b = blockIdx.x;
    const int i0 = b * bw;
    const int w = min(bw, nref - i0);
    const int r = threadIdx.x % kTfactorMax, q = threadIdx.x / kTfactorMax;
    const double2* Sb = S + (int64_t)b * bw * bw;
    double2* Tb = T + (int64_t)b * bw * bw;
    const double2 zero2 = make_double2(0.0, 0.0);
    // prologue: taus, column 0 (tau_0), S(:, 1) into sc and S(:, 2) into registers
    double2 pre = zero2;
    if (q == 0) {
        double2 t = zero2;
        if (r < w) t = tau[i0 + r];
        tl[r] = t;
        if (r == 0 && w > 0) tp[0] = t;
        double2 s1 = zero2;
        if (r < 1 && 1 < w) s1 = Sb[(int64_t)1 * bw + r];
        sc[r] = s1;
        if (r < 2 && 2 < w) pre = Sb[(int64_t)2 * bw + r];
    }
    __syncthreads();
    for (int c = 1; c < w; ++c) {
        // partial of sum_{j=r}^{c-1} T(r, j) S(j, c) over j = r + q, r + q + kTfSplit, ...
        double sr = 0.0, si = 0.0;
        if (r < c) {
#pragma unroll 4
            for (int j = r + q; j < c; j += kTfSplit) {
                const double2 a = tp[j * (j + 1) / 2 + r], x = sc[j];
                sr += a.x * x.x - a.y * x.y;
                si += a.x * x.y + a.y * x.x;
            }
        }
        part[q][r] = make_double2(sr, si);
        __syncthreads();
        if (q == 0) {
            const double2 t = tl[c];
            if (r < c) {
                double ar = 0.0, ai = 0.0;
#pragma unroll
                for (int u = 0; u < kTfSplit; ++u) {
                    ar += part[u][r].x;
                    ai += part[u][r].y;
                }
                tp[c * (c + 1) / 2 + r] = make_double2(-(t.x * ar - t.y * ai), -(t.x * ai + t.y * ar));
            } else if (r == c) {
                tp[c * (c + 1) / 2 + c] = t;
            }
            // column c + 1 of S for the next step, and the load of column c + 2
            sc[r] = pre;
            pre = zero2;
            if (r < c + 2 && c + 2 < w) pre = Sb[(int64_t)(c + 2) * bw + r];
        }
        __syncthreads();  // column c and sc complete before they are read
    }
    for (int e = threadIdx.x; e < bw * bw; e += blockDim.x) {
        const int c = e / bw, rr = e % bw;
        double2 t = zero2;
        if (c < w && rr <= c) t = tp[c * (c + 1) / 2 + rr];
        Tb[e] = t;
    }
}

void launch_hh_tfactor(hipStream_t st, const double2* S, const double2* tau, double2* T, int nblocks,
                       int bw, int nref) {
    if (nblocks <= 0) return;
    KV_REQUIRE(bw >= 1 && bw <= kTfactorMax, ERR_ARG, "reflector block width must be in [1, 128]");
    KV_REQUIRE((int64_t)nblocks * bw >= nref && (int64_t)(nblocks - 1) * bw < nref, ERR_ARG,
               "reflector blocks do not cover the reflectors");
    hipLaunchKernelGGL(hh_tfactor_kernel, dim3(nblocks), dim3(kTfactorMax * kTfSplit), 0, st, S, tau, T, bw,
                       nref);
    KV_HIP(hipGetLastError());
}

// out[e] = sum_{p < np} part[p * n + e]  (split-K partial products summed).
__global__ __launch_bounds__(256) void sum_parts_kernel(const double2* __restrict__ part, int np,
                                                        int64_t n, double2* __restrict__ out) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        double re = 0, im = 0;
        for (int p = 0; p < np; ++p) {
            const double2 v = part[(int64_t)p * n + e];
            re += v.x;
            im += v.y;
        }
        out[e] = make_double2(re, im);
    }
}

void launch_sum_parts(hipStream_t st, const double2* part, int np, int64_t n, double2* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(sum_parts_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, part, np, n, out);
    KV_HIP(hipGetLastError());
}

// Z[j n + i] = C[(c0 + j) ldc + i] for i < n, j < k (real columns to complex).
__global__ __launch_bounds__(256) void real_cols_kernel(const double* __restrict__ C, int ldc, int c0,
                                                        int k, int n, double2* __restrict__ Z) {
    const int64_t total = (int64_t)n * k;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int j = (int)(e / n), i = (int)(e % n);
        Z[e] = make_double2(C[(int64_t)(c0 + j) * ldc + i], 0.0);
    }
}

void launch_real_cols(hipStream_t st, const double* C, int ldc, int c0, int k, int n, double2* Z) {
    const int64_t total = (int64_t)n * k;
    if (total <= 0) return;
    hipLaunchKernelGGL(real_cols_kernel, dim3(ew_blocks(total)), dim3(256), 0, st, C, ldc, c0, k, n, Z);
    KV_HIP(hipGetLastError());
}


// ---- Householder tridiagonalisation (lower triangle, LAPACK storage) ----------------
// Batched reduction Q^H A Q = T of column-major n x n Hermitian matrices (one per grid.y),
// blocked by panels of kTrdPanel columns: within a panel column j is brought up to date by
// the panel's earlier reflectors (V) and their companions (W), reflected (larfg), and its
// companion W(:, i) = tau (A v - V (W^H v) - W (V^H v)) + alpha v formed from a matvec over
// the trailing matrix as it stood at the panel start; after the panel the trailing matrix
// takes the rank-2 update A -= V W^H + W V^H on both triangles.  A is kept fully Hermitian
// (both triangles) so the matvec is a set of independent column dot products.  On exit
// D = diag(T), E = subdiag(T), tau and the reflector tails below the subdiagonal of A follow
// LAPACK zhetrd('L').  VW holds V in columns [0, nb) and W in [nb, 2 nb) (column-major,
// n rows); q holds W^H v in [0, nb) and V^H v in [nb, 2 nb).

// Mirror the lower triangle of each column-major n x n matrix into its upper triangle
// (conjugated) and drop the imaginary part of the diagonal.  32 x 32 tiles, grid (T, T, batch).
__global__ __launch_bounds__(256) void herm_fill_kernel(double2* __restrict__ A, int n, int64_t stride) {
    const int ti = blockIdx.x, tj = blockIdx.y;  // source tile rows ti, columns tj (lower: ti >= tj)
    if (ti < tj) return;
    A += (int64_t)blockIdx.z * stride;
    __shared__ double2 t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int c = ty; c < 32; c += 8) {
        const int gr = ti * 32 + tx, gc = tj * 32 + c;
        if (gr < n && gc < n) t[c][tx] = A[(int64_t)gc * n + gr];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int dc = ti * 32 + r, dr = tj * 32 + tx;  // destination (dr, dc) = conj(source (dc, dr))
        if (dc < n && dr < n) {
            const double2 v = t[tx][r];
            if (dc > dr)
                A[(int64_t)dc * n + dr] = make_double2(v.x, -v.y);
            else if (dc == dr)
                A[(int64_t)dc * n + dr] = make_double2(v.x, 0.0);
        }
    }
}

constexpr int kTrdRows = 64;  // rows per block of trd_column_kernel (x 4 slices of the panel index)
constexpr int kTrdReg = 8;    // rows per thread whose loads are issued together (256 threads: 2048 rows)

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ inline double2 wave_sum2(double2 v) {  // fixed butterfly over the 64 lanes
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    return v;
}

// One column step of the panel, rows [j, n), grid (ceil((n - j) / 64), batch).
// prev: finish the companion of column j - 1 (local index i - 1): w' = tau_{j-1} (y - V q1 - W q2)
//   over the panel's columns before i - 1, alpha = -tau_{j-1} (w'^H v) / 2, W(:, i-1) = w' + alpha v.
// col: bring column j up to date with the panel's columns before i and reflect it (j < n - 1);
//   with v(j) = 1 the contribution of column i - 1 is -v conj(w'(j)) - w' - 2 Re(alpha) v.
// Each block writes its share of w'^H v; the last block to finish (ticket on count[b], which
// it resets) sums the shares in block order and does the O(n) tail: W, the final column, its
// norm, larfg and the reflector, with each thread's rows loaded together.
__global__ __launch_bounds__(256) void trd_column_kernel(double2* __restrict__ A, double2* __restrict__ VW,
                                                         const double2* __restrict__ y,
                                                         const double2* __restrict__ q, double* __restrict__ Dg,
                                                         double* __restrict__ Eg, double2* __restrict__ tau,
                                                         double2* __restrict__ part, unsigned* __restrict__ count,
                                                         int n, int j, int i, int nb, int prev, int col) {
    const int b = blockIdx.y;
    A += (int64_t)b * n * n;
    VW += (int64_t)b * n * 2 * nb;
    y += (int64_t)b * n;
    q += (int64_t)b * 2 * nb;
    Dg += (int64_t)b * n;
    Eg += (int64_t)b * n;
    tau += (int64_t)b * n;
    part += (int64_t)b * gridDim.x;
    const double2* V = VW;
    double2* W = VW + (int64_t)nb * n;
    const double2 zero2 = make_double2(0.0, 0.0);
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int K = prev ? i - 1 : 0;  // panel columns with final companions
    __shared__ double2 vj[kTrdPanel], wj[kTrdPanel], q1[kTrdPanel], q2[kTrdPanel];
    __shared__ double2 rw[4][kTrdRows], rc[4][kTrdRows], red[kTrdRows];
    __shared__ double2 wpj_s, alpha_s, alf_s, scale_s;
    __shared__ double ssq[4];
    __shared__ int last;
    const double2 taup = prev ? tau[j - 1] : zero2;
    if (threadIdx.x < K) {
        vj[threadIdx.x] = V[(int64_t)threadIdx.x * n + j];
        wj[threadIdx.x] = W[(int64_t)threadIdx.x * n + j];
        q1[threadIdx.x] = q[threadIdx.x];
        q2[threadIdx.x] = q[nb + threadIdx.x];
    }
    const int r = j + blockIdx.x * kTrdRows + tx;
    double2 yr = zero2, ar = zero2, vr = zero2;
    if (ty == 0 && r < n) {
        if (prev) {
            yr = y[r];
            vr = V[(int64_t)(i - 1) * n + r];
        }
        if (col) ar = A[(int64_t)j * n + r];
    }
    __syncthreads();
    if (prev && threadIdx.x < 64) {  // w'(j), needed by every row of column j: one panel column per lane
        double2 t = zero2;
        if (threadIdx.x < K) {
            const double2 a = cmul(vj[threadIdx.x], q1[threadIdx.x]), c = cmul(wj[threadIdx.x], q2[threadIdx.x]);
            t = make_double2(a.x + c.x, a.y + c.y);
        }
        t = wave_sum2(t);
        if (threadIdx.x == 0) {
            const double2 yj = y[j];
            wpj_s = cmul(taup, make_double2(yj.x - t.x, yj.y - t.y));
        }
    }
    double2 sw = zero2, sc = zero2;
    if (r < n) {
        for (int k = ty; k < K; k += 4) {
            const double2 vk = V[(int64_t)k * n + r], wk = W[(int64_t)k * n + r];
            const double2 a = cmul(vk, q1[k]), c = cmul(wk, q2[k]);
            sw.x += a.x + c.x;
            sw.y += a.y + c.y;
            const double2 e = cmulc(vk, wj[k]), f = cmulc(wk, vj[k]);
            sc.x += e.x + f.x;
            sc.y += e.y + f.y;
        }
    }
    rw[ty][tx] = sw;
    rc[ty][tx] = sc;
    __syncthreads();
    if (ty == 0) {
        double2 d = zero2;
        if (r < n) {
            for (int t = 1; t < 4; ++t) {
                sw.x += rw[t][tx].x;
                sw.y += rw[t][tx].y;
                sc.x += rc[t][tx].x;
                sc.y += rc[t][tx].y;
            }
            double2 wp = zero2;
            if (prev) {
                wp = cmul(taup, make_double2(yr.x - sw.x, yr.y - sw.y));
                W[(int64_t)(i - 1) * n + r] = wp;
                d = cmulc(vr, wp);  // conj(wp) v
            }
            if (col) {
                double2 c0 = make_double2(ar.x - sc.x, ar.y - sc.y);
                if (prev) {
                    const double2 t2 = cmulc(vr, wpj_s);
                    c0.x -= t2.x + wp.x;
                    c0.y -= t2.y + wp.y;
                }
                A[(int64_t)j * n + r] = c0;
            }
        }
        d = wave_sum2(d);
        if (tx == 0) {
            part[blockIdx.x] = d;
            __threadfence();
            last = atomicAdd(&count[b], 1u) == gridDim.x - 1;
        }
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // ---- tail, one block: rows [j, n) ----
    if (threadIdx.x < 64) {
        double2 d = zero2;
        if (prev)
            for (unsigned t = threadIdx.x; t < gridDim.x; t += 64) {
                d.x += part[t].x;
                d.y += part[t].y;
            }
        d = wave_sum2(d);
        if (threadIdx.x == 0) {
            alpha_s = prev ? cmul(make_double2(-0.5 * taup.x, -0.5 * taup.y), d) : zero2;
            alf_s = zero2;
            count[b] = 0u;
        }
    }
    __syncthreads();
    const double2 alpha = alpha_s;
    const double g2 = 2.0 * alpha.x;
    const int m = n - j;
    const bool spill = m > 256 * kTrdReg;  // rows beyond the registers go back through A
    double2 cv[kTrdReg];
    double s2 = 0.0;
    for (int base = 0; base < m; base += 256 * kTrdReg) {
        double2 vv[kTrdReg], ww[kTrdReg], cc[kTrdReg];
#pragma unroll
        for (int u = 0; u < kTrdReg; ++u) {
            const int rr = j + base + u * 256 + threadIdx.x;
            vv[u] = zero2;
            ww[u] = zero2;
            cc[u] = zero2;
            if (rr < n) {
                if (prev) {
                    vv[u] = V[(int64_t)(i - 1) * n + rr];
                    ww[u] = W[(int64_t)(i - 1) * n + rr];
                }
                if (col) cc[u] = A[(int64_t)j * n + rr];
            }
        }
#pragma unroll
        for (int u = 0; u < kTrdReg; ++u) {
            const int rr = j + base + u * 256 + threadIdx.x;
            if (rr >= n) continue;
            if (prev) {
                const double2 av = cmul(alpha, vv[u]);
                W[(int64_t)(i - 1) * n + rr] = make_double2(ww[u].x + av.x, ww[u].y + av.y);
            }
            if (col) {
                const double2 c = make_double2(cc[u].x - g2 * vv[u].x, cc[u].y - g2 * vv[u].y);
                if (base == 0) cv[u] = c;
                if (spill) A[(int64_t)j * n + rr] = c;
                if (rr == j) Dg[j] = c.x;
                if (rr == j + 1) alf_s = c;
                if (rr >= j + 2) s2 += c.x * c.x + c.y * c.y;
            }
        }
    }
    if (!col || j >= n - 1) return;
    {
        double2 t = wave_sum2(make_double2(s2, 0.0));
        if (tx == 0) ssq[ty] = t.x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // larfg: H^H [alpha; x] = [beta; 0], H = I - tau v v^H, v = [1; x / (alpha - beta)]
        const double2 a = alf_s;
        const double xn2 = ((ssq[0] + ssq[1]) + ssq[2]) + ssq[3];
        double beta, tr = 0.0, ti = 0.0;
        double2 z = make_double2(1.0, 0.0);
        if (xn2 == 0.0 && a.y == 0.0) {
            beta = a.x;
        } else {
            beta = -copysign(sqrt(a.x * a.x + a.y * a.y + xn2), a.x);
            tr = (beta - a.x) / beta;
            ti = -a.y / beta;
            const double zr = a.x - beta, zi = a.y, z2 = zr * zr + zi * zi;
            z = make_double2(zr / z2, -zi / z2);
        }
        Eg[j] = beta;
        tau[j] = make_double2(tr, ti);
        scale_s = z;
        A[(int64_t)j * n + j + 1] = make_double2(beta, 0.0);
        VW[(int64_t)i * n + j + 1] = make_double2(1.0, 0.0);
    }
    __syncthreads();
    const double2 scl = scale_s;
    for (int base = 0; base < m; base += 256 * kTrdReg) {
#pragma unroll
        for (int u = 0; u < kTrdReg; ++u) {
            const int rr = j + base + u * 256 + threadIdx.x;
            if (rr < j + 2 || rr >= n) continue;
            const double2 v = cmul(base == 0 ? cv[u] : A[(int64_t)j * n + rr], scl);
            A[(int64_t)j * n + rr] = v;
            VW[(int64_t)i * n + rr] = v;
        }
    }
}

// y[c] = sum_{r >= s} conj(A(r, c)) v(r) for c in [s, n) (= (A v)(c), A Hermitian), then
// q[k] = W(:, k)^H v and q[nb + k] = V(:, k)^H v over rows [s, n) for k < i, with v = V(:, i).
// Two outputs per block of 256 threads, the rows spread over the block with each thread's
// loads issued together; grid (ceil((n - s + 2 i) / 2), batch).
__global__ __launch_bounds__(256) void trd_matvec_kernel(const double2* __restrict__ A,
                                                         const double2* __restrict__ VW, double2* __restrict__ y,
                                                         double2* __restrict__ q, int n, int s, int i, int nb) {
    const int b = blockIdx.y;
    A += (int64_t)b * n * n;
    VW += (int64_t)b * n * 2 * nb;
    y += (int64_t)b * n;
    q += (int64_t)b * 2 * nb;
    const int m = n - s, ntask = m + 2 * i;
    const int t0 = blockIdx.x * 2;
    const bool two = t0 + 1 < ntask;
    auto colp = [&](int t) -> const double2* {
        if (t < m) return A + (int64_t)(s + t) * n;
        t -= m;
        return t < i ? VW + (int64_t)(nb + t) * n : VW + (int64_t)(t - i) * n;  // W(:, t) or V(:, t - i)
    };
    const double2* c0 = colp(t0);
    const double2* c1 = two ? colp(t0 + 1) : c0;
    const double2* v = VW + (int64_t)i * n;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    for (int base = s; base < n; base += 256 * kTrdReg) {
        double2 vv[kTrdReg], x0[kTrdReg], x1[kTrdReg];
#pragma unroll
        for (int u = 0; u < kTrdReg; ++u) {
            const int r = base + u * 256 + threadIdx.x;
            vv[u] = make_double2(0.0, 0.0);
            x0[u] = vv[u];
            x1[u] = vv[u];
            if (r < n) {
                vv[u] = v[r];
                x0[u] = c0[r];
                x1[u] = c1[r];
            }
        }
#pragma unroll
        for (int u = 0; u < kTrdReg; ++u) {
            const double2 e = cmulc(vv[u], x0[u]), f = cmulc(vv[u], x1[u]);  // conj(x) v
            a0.x += e.x;
            a0.y += e.y;
            a1.x += f.x;
            a1.y += f.y;
        }
    }
    a0 = wave_sum2(a0);
    a1 = wave_sum2(a1);
    __shared__ double2 red[2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = a0;
        red[1][wave] = a1;
    }
    __syncthreads();
    if (threadIdx.x < 2 && (threadIdx.x == 0 || two)) {
        const int t = t0 + threadIdx.x;
        double2 acc = red[threadIdx.x][0];
        for (int w = 1; w < 4; ++w) {
            acc.x += red[threadIdx.x][w].x;
            acc.y += red[threadIdx.x][w].y;
        }
        if (t < m)
            y[s + t] = acc;
        else if (t - m < i)
            q[t - m] = acc;
        else
            q[nb + t - m - i] = acc;
    }
}

// Householder tridiagonalisation in zhetrd('L')'s form of Hermitian matrices up to
// kTrdSmallMax wide, one 768-thread workgroup per matrix (column-major, both triangles
// valid, as herm_fill leaves them).  Unblocked, with each step's rank-2 update fused into the
// next step's matrix-vector product: after column j's reflector v_j and y_j = A v_j,
//   w_j = tau_j y_j - (tau_j / 2) (tau_j y_j)^H v_j v_j,
//   column j + 1 brought up to date (A -= v_j w_j^H + w_j v_j^H on it) and reflected by
//   larfg's formulas (v_{j+1}, tau_{j+1}, beta = E[j+1]; the reflector's tail stored below
//   the subdiagonal of column j + 1 and beta on it),
//   one pass over A(j+2:, j+2:) applying the same update and forming y_{j+1} = A v_{j+1}
//   (as a combination of columns, exact for the Hermitian A the updates keep).
// D, E, tau and the stored reflectors are hetrd_lower's (the back-transformation reads them).
constexpr int kTrdSmallThreads = 768;
__global__ __launch_bounds__(kTrdSmallThreads) void trd_small_kernel(double2* __restrict__ A, int n,
                                                                     double* __restrict__ Dg, double* __restrict__ Eg,
                                                                     double2* __restrict__ tau) {
    constexpr int NW = kTrdSmallThreads / 64, RU = kTrdSmallMax / 64;
    const int b = blockIdx.x;
    A += (int64_t)b * n * n;
    Dg += (int64_t)b * n;
    Eg += (int64_t)b * n;
    tau += (int64_t)b * n;
    __shared__ double2 vbuf[2][kTrdSmallMax], w0[kTrdSmallMax], y[kTrdSmallMax], col[kTrdSmallMax];
    __shared__ double2 part[NW][kTrdSmallMax];
    __shared__ double2 red2[NW];
    __shared__ double red1[NW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double2 zero2 = make_double2(0.0, 0.0);
    if (n <= 0) return;
    // larfg on col[j + 1 ..) (every thread the same arithmetic from the reduced norm): beta into
    // E and the subdiagonal, tau, the reflector into v (rows >= j + 1, v(j + 1) = 1) and below
    // the subdiagonal of column j.  Returns tau_j; ends synchronised.
    auto reflect = [&](int j, double2* v, double s2) -> double2 {
        for (int off = 32; off > 0; off >>= 1) s2 += __shfl_xor(s2, off, 64);
        if (lane == 0) red1[wv] = s2;
        __syncthreads();
        double xn2 = 0.0;
        for (int q = 0; q < NW; ++q) xn2 += red1[q];
        const double2 a = col[j + 1];
        double beta, tr = 0.0, ti = 0.0;
        double2 z = make_double2(1.0, 0.0);
        if (xn2 == 0.0 && a.y == 0.0) {
            beta = a.x;
        } else {
            beta = -copysign(sqrt(a.x * a.x + a.y * a.y + xn2), a.x);
            tr = (beta - a.x) / beta;
            ti = -a.y / beta;
            const double zr = a.x - beta, zi = a.y, z2 = zr * zr + zi * zi;
            z = make_double2(zr / z2, -zi / z2);
        }
        if (tid == 0) {
            Eg[j] = beta;
            tau[j] = make_double2(tr, ti);
            A[(int64_t)j * n + j + 1] = make_double2(beta, 0.0);
        }
        for (int r = tid; r < n; r += kTrdSmallThreads) {
            double2 vr = zero2;
            if (r == j + 1) {
                vr = make_double2(1.0, 0.0);
            } else if (r > j + 1) {
                vr = cmul(col[r], z);
                A[(int64_t)j * n + r] = vr;
            }
            v[r] = vr;
        }
        __syncthreads();
        return make_double2(tr, ti);
    };
    // y = A(s:, s:) v over rows and columns >= s, with the pending update (vp, w0) applied to
    // A(s:, s:) first and written back when `upd`; each lane's rows of a column are loaded
    // together before they are used
    // y = A(s:, s:) v over rows and columns >= s as a combination of columns (A is Hermitian,
    // so no reduction across lanes is needed): lane rows s + lane + 64 u, wave columns
    // s + wv CG + g + NW CG t; the pending update (vp, w0) is applied to A(s:, s:) first and
    // written back when `upd`.  The waves' partial vectors are summed in wave order.
    auto pass = [&](int s, const double2* v, const double2* vp, bool upd) {
        // columns per wave whose loads are issued before the first store (2 at 12 waves: 168 VGPRs,
        // no spills; 4 at 8 waves)
        constexpr int CG = kTrdSmallThreads > 512 ? 2 : 4;
        double2 acc[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) acc[u] = zero2;
        for (int c0 = s + wv * CG; c0 < n; c0 += NW * CG) {
            double2 a[CG][RU];
#pragma unroll
            for (int g = 0; g < CG; ++g)
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    const int r = s + lane + 64 * u, c = c0 + g;
                    a[g][u] = zero2;
                    if (r < n && c < n) a[g][u] = A[(int64_t)c * n + r];
                }
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                const int c = c0 + g;
                if (c >= n) break;  // uniform over the wave
                const double2 wc = upd ? w0[c] : zero2, vc = upd ? vp[c] : zero2, xc = v[c];
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    const int r = s + lane + 64 * u;
                    if (r < n) {
                        double2 x = a[g][u];
                        if (upd) {
                            const double2 p = cmulc(vp[r], wc), q = cmulc(w0[r], vc);  // v(r) conj(w(c)) + w(r) conj(v(c))
                            x = make_double2(x.x - p.x - q.x, x.y - p.y - q.y);
                            A[(int64_t)c * n + r] = x;
                        }
                        const double2 e = cmul(x, xc);
                        acc[u].x += e.x;
                        acc[u].y += e.y;
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int r = s + lane + 64 * u;
            if (r < n) part[wv][r] = acc[u];
        }
        __syncthreads();
        for (int r = s + tid; r < n; r += kTrdSmallThreads) {
            double2 t = part[0][r];
            for (int q = 1; q < NW; ++q) {
                t.x += part[q][r].x;
                t.y += part[q][r].y;
            }
            y[r] = t;
        }
        __syncthreads();
    };
    // column 0
    double s2 = 0.0;
    for (int r = tid; r < n; r += kTrdSmallThreads) {
        const double2 x = A[r];
        col[r] = x;
        if (r >= 2) s2 += x.x * x.x + x.y * x.y;
    }
    if (tid == 0) Dg[0] = A[0].x;
    if (n == 1) return;
    __syncthreads();
    double2 tj = reflect(0, vbuf[0], s2);
    pass(1, vbuf[0], vbuf[0], false);
    for (int j = 0; j + 1 < n; ++j) {
        const double2* v0 = vbuf[j & 1];
        double2* v1 = vbuf[(j + 1) & 1];
        // alpha = -(tau/2) (tau y)^H v over rows >= j + 1
        double2 d = zero2;
        for (int r = j + 1 + tid; r < n; r += kTrdSmallThreads) {
            const double2 e = cmulc(v0[r], cmul(tj, y[r]));  // conj(tau y) v
            d.x += e.x;
            d.y += e.y;
        }
        d = wave_sum2(d);
        if (lane == 0) red2[wv] = d;
        __syncthreads();
        double2 t = zero2;
        for (int q = 0; q < NW; ++q) {
            t.x += red2[q].x;
            t.y += red2[q].y;
        }
        const double2 al = cmul(make_double2(-0.5 * tj.x, -0.5 * tj.y), t);
        for (int r = j + 1 + tid; r < n; r += kTrdSmallThreads) {
            const double2 x = cmul(tj, y[r]), av = cmul(al, v0[r]);
            w0[r] = make_double2(x.x + av.x, x.y + av.y);
        }
        __syncthreads();
        // column j + 1, up to date, and the norm of its tail
        const int c1 = j + 1;
        const double2 wc = w0[c1], vc = v0[c1];
        s2 = 0.0;
        for (int r = c1 + tid; r < n; r += kTrdSmallThreads) {
            const double2 a = A[(int64_t)c1 * n + r];
            const double2 p = cmulc(v0[r], wc), q = cmulc(w0[r], vc);
            const double2 x = make_double2(a.x - p.x - q.x, a.y - p.y - q.y);
            col[r] = x;
            if (r == c1) Dg[c1] = x.x;
            if (r >= c1 + 2) s2 += x.x * x.x + x.y * x.y;
        }
        if (c1 + 1 >= n) break;
        __syncthreads();
        tj = reflect(c1, v1, s2);
        pass(c1 + 1, v1, v0, true);
    }
}

void launch_trd_small(hipStream_t st, double2* A, int n, double* D, double* E, double2* tau, int batch) {
    KV_REQUIRE(n >= 1 && n <= kTrdSmallMax && batch >= 1, ERR_INTERNAL, "trd_small: side above kTrdSmallMax");
    hipLaunchKernelGGL(trd_small_kernel, dim3(batch), dim3(kTrdSmallThreads), 0, st, A, n, D, E, tau);
    KV_HIP(hipGetLastError());
}

void launch_herm_fill(hipStream_t st, double2* A, int n, int batch) {
    if (n <= 0 || batch <= 0) return;
    const int T = (n + 31) / 32;
    hipLaunchKernelGGL(herm_fill_kernel, dim3(T, T, batch), dim3(256), 0, st, A, n, (int64_t)n * n);
    KV_HIP(hipGetLastError());
}

int trd_column_blocks(int n, int j) { return (n - j + kTrdRows - 1) / kTrdRows; }

void launch_trd_column(hipStream_t st, double2* A, double2* VW, const double2* y, const double2* q, double* D,
                       double* E, double2* tau, double2* part, unsigned* count, int n, int j, int i, int nb,
                       bool prev, bool col, int batch) {
    KV_REQUIRE(nb <= kTrdPanel && i <= nb && j >= 0 && j < n && (prev ? i >= 1 : i == 0) && j - i >= 0, ERR_INTERNAL,
               "trd_column: bad step");
    hipLaunchKernelGGL(trd_column_kernel, dim3(trd_column_blocks(n, j), batch), dim3(256), 0, st, A, VW, y, q, D,
                       E, tau, part, count, n, j, i, nb, prev ? 1 : 0, col ? 1 : 0);
    KV_HIP(hipGetLastError());
}

void launch_trd_matvec(hipStream_t st, const double2* A, const double2* VW, double2* y, double2* q, int n, int s,
                       int i, int nb, int batch) {
    KV_REQUIRE(s >= 1 && s < n && i >= 0 && i < nb && nb <= kTrdPanel, ERR_INTERNAL, "trd_matvec: bad step");
    const int ntask = n - s + 2 * i;
    hipLaunchKernelGGL(trd_matvec_kernel, dim3((ntask + 1) / 2, batch), dim3(256), 0, st, A, VW, y, q, n, s, i,
                       nb);
    KV_HIP(hipGetLastError());
}

// ---- block subspace route of the bond SVD --------------------------------------------
// Randomised range finding on a sector matrix (Halko, Martinsson and Tropp, SIAM Review 53
// (2011), algorithms 4.3/4.4 with Rayleigh-Ritz): a seeded random start block, Cholesky QR
// orthonormalisation with rank-deficient columns dropped, and the Ritz residual check.

// splitmix64 output function (Steele, Lea and Flood, OOPSLA 2014): a bijective 64-bit mix.
__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// p[e] = (u, v) uniform in [-1, 1)^2, a function of (seed, e) only.
__global__ __launch_bounds__(256) void rand_fill_kernel(double2* __restrict__ p, int64_t n, uint64_t seed) {
    const uint64_t s = mix64(seed);
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const uint64_t a = mix64(s ^ (uint64_t)(2 * e)), b = mix64(s ^ (uint64_t)(2 * e + 1));
        p[e] = make_double2((double)(a >> 11) * 0x1.0p-52 - 1.0, (double)(b >> 11) * 0x1.0p-52 - 1.0);
    }
}

void launch_rand_fill(hipStream_t st, double2* p, int64_t n, uint64_t seed) {
    if (n <= 0) return;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(rand_fill_kernel, dim3(blocks), dim3(256), 0, st, p, n, seed);
    KV_HIP(hipGetLastError());
}

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline double2 cfma_conj(double2 acc, double2 a, double2 b) {  // acc + conj(a) b
    return make_double2(acc.x + a.x * b.x + a.y * b.y, acc.y + a.x * b.y - a.y * b.x);
}
__device__ inline double2 cfma(double2 acc, double2 a, double2 b) {  // acc + a b
    return make_double2(acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x);
}

// Largest real diagonal entry of each batched matrix (the Cholesky drop threshold's scale),
// read before any panel step overwrites S.  One workgroup per matrix.
__global__ __launch_bounds__(256) void diag_max_kernel(const double2* __restrict__ S, const int* __restrict__ sizes,
                                                       int ld, int64_t stride, double* __restrict__ dmax) {
    const int b = blockIdx.x, n = sizes[b];
    S += (int64_t)b * stride;
    __shared__ double red[4];
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) mx = fmax(mx, S[(int64_t)i * ld + i].x);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) dmax[b] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Left-looking blocked Cholesky S = R^H R with T = R^{-1}, panel k in two launches.
// Shared by both: 32 x 32 tile sums over rows of R (or T) staged through LDS in 32-row
// chunks, the four waves taking every fourth row and each lane a 4 x 4 block, the waves'
// partial tiles added in wave order.
struct CholTiles {
    double2 At[kCholLeftNB][kCholLeftNB + 1], Bt[kCholLeftNB][kCholLeftNB + 1];
    double2 red[3][kCholLeftNB][kCholLeftNB + 1];
};

// acc (this lane's 4 x 4 block of a 32 x 32 tile) += sum over rows i in [i_begin, i_end) of
// op(A)(i, r) B(i, c), with A(i, r) = S[i ld + a_col + r] (conj, arows = false) or
// T[(a_row + r) ld + i] (arows = true), B(i, c) = S[i ld + b_col + c] for c < bw; then the
// waves' blocks are added into wave 0's.
__device__ __forceinline__ void chol_tile_sum(CholTiles& L, double2 (&acc)[4][4], const double2* __restrict__ S,
                                     const double2* __restrict__ T, int ld, int i_begin, int i_end, int a_col,
                                     int a_row, bool arows, int aw, int b_col, int bw) {
    constexpr int NB = kCholLeftNB;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r4 = (lane >> 3) * 4, c4 = (lane & 7) * 4;
    const double2 zero2 = make_double2(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = zero2;
    // the next chunk's operands are loaded into registers while the current one is summed
    constexpr int NQ = NB * NB / 256;
    double2 pa[NQ], pb[NQ];
    auto fetch = [&](int i0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q, ii = e / NB, cc = e % NB;
            pa[q] = zero2;
            pb[q] = zero2;
            if (arows)
                pa[q] = T[(int64_t)(a_row + ii) * ld + i0 + cc];  // A^T staged: At[r][i]
            else if (cc < aw)
                pa[q] = S[(int64_t)(i0 + ii) * ld + a_col + cc];
            if (cc < bw) pb[q] = S[(int64_t)(i0 + ii) * ld + b_col + cc];
        }
    };
    if (i_begin < i_end) fetch(i_begin);
    for (int i0 = i_begin; i0 < i_end; i0 += NB) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q;
            L.At[e / NB][e % NB] = pa[q];
            L.Bt[e / NB][e % NB] = pb[q];
        }
        __syncthreads();
        if (i0 + NB < i_end) fetch(i0 + NB);
        for (int ii = wv; ii < NB; ii += 4) {
            double2 av[4], bv[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                av[x] = arows ? L.At[r4 + x][ii] : L.At[ii][r4 + x];
                bv[x] = L.Bt[ii][c4 + x];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
                    acc[x][y] = arows ? cfma(acc[x][y], av[x], bv[y]) : cfma_conj(acc[x][y], av[x], bv[y]);
        }
        __syncthreads();
    }
    if (wv > 0)
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) L.red[wv - 1][r4 + x][c4 + y] = acc[x][y];
    __syncthreads();
    if (wv == 0)
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const double2 v = L.red[q][r4 + x][c4 + y];
                    acc[x][y].x += v.x;
                    acc[x][y].y += v.y;
                }
    __syncthreads();
}

// Launch A of panel k, one workgroup per matrix: C_kk = S_kk - R_{<k,k}^H R_{<k,k}, its
// factor with the rank-tolerant pivot rule (a pivot at or below tau * dmax drops its column:
// R's row and T's row and column are zero; live flags), and its inverse Di, into T's tile
// (k, k) and Rd[b][k].  The factor is right-looking in LDS with all threads; the inverse of
// the triangular factor by recursive doubling: T_AC = -T_A R_AC T_C for pairs of diagonal
// blocks of 1, 2, 4, 8 and 16.
__global__ __launch_bounds__(256) void chol_diag_kernel(const double2* __restrict__ S, const int* __restrict__ sizes,
                                                        int ld, int64_t stride, double tau,
                                                        int* __restrict__ live, const double* __restrict__ dmax,
                                                        double2* __restrict__ T, double2* __restrict__ Rd, int k) {
    constexpr int NB = kCholLeftNB;
    const int b = blockIdx.x, n = sizes[b];
    const int p0 = k * NB;
    if (p0 >= n) return;  // uniform over the workgroup
    const int w = min(NB, n - p0);
    S += (int64_t)b * stride;
    T += (int64_t)b * stride;
    live += (int64_t)b * ld;
    Rd += ((int64_t)b * ((ld + NB - 1) / NB) + k) * NB * NB;
    const double thr = tau * dmax[b];
    __shared__ CholTiles L;
    __shared__ double2 Dg[NB][NB + 1], Ti[NB][NB + 1], Rf[NB][NB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r4 = (lane >> 3) * 4, c4 = (lane & 7) * 4;
    const double2 zero2 = make_double2(0.0, 0.0);
    double2 ak[4][4];
    chol_tile_sum(L, ak, S, T, ld, 0, p0, p0, 0, false, w, p0, w);
    if (wv == 0)
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int r = r4 + x, c = c4 + y;
                double2 v = zero2;
                if (r < w && c < w && c >= r) {
                    const double2 sv = S[(int64_t)(p0 + r) * ld + p0 + c];
                    v = make_double2(sv.x - ak[x][y].x, sv.y - ak[x][y].y);
                }
                Dg[r][c] = v;
                Rf[r][c] = zero2;
            }
    __syncthreads();
    // factor in LDS, right-looking, all threads, one phase per pivot: every thread reads the
    // pivot; row kk of R (Dg's row over sqrt(piv)) goes to Rf and the trailing triangle takes
    // the update with Dg's unscaled row over piv.  Dg's row kk is only read in the phase, and
    // each thread loads all its operands before it stores.
    constexpr int NQ = NB * NB / 256;
    unsigned lv = 0u;
    for (int kk = 0; kk < w; ++kk) {
        const double piv = Dg[kk][kk].x;
        const bool lk = piv > thr && piv > 0.0;  // NaN is not live
        const double sp = lk ? sqrt(piv) : 0.0;
        const double ri = lk ? 1.0 / sp : 0.0, rp = ri * ri;
        lv |= lk ? 1u << kk : 0u;
        double2 xa[NQ], ya[NQ], za[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q, i = e / NB, c = e % NB;
            xa[q] = zero2;
            za[q] = zero2;
            ya[q] = Dg[kk][c];
            if (i > kk && c >= i && c < w) {
                xa[q] = Dg[kk][i];
                za[q] = Dg[i][c];
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q, i = e / NB, c = e % NB;
            if (i == kk && c >= kk && c < w)
                Rf[kk][c] = c == kk ? make_double2(sp, 0.0) : make_double2(ya[q].x * ri, ya[q].y * ri);
            if (lk && i > kk && c >= i && c < w) {
                const double2 x = xa[q], y = ya[q];
                Dg[i][c] = make_double2(za[q].x - (x.x * y.x + x.y * y.y) * rp, za[q].y - (x.x * y.y - x.y * y.x) * rp);
            }
        }
        __syncthreads();
    }
    // Ti = diag(1 / R_ii) over live indices
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e / NB, c = e % NB;
        Ti[i][c] = (i == c && i < w && ((lv >> i) & 1u)) ? make_double2(1.0 / Rf[i][i].x, 0.0) : zero2;
    }
    __syncthreads();
    // recursive doubling: for pairs (A = [p, p+s), C = [p+s, p+2s)): M = R_AC T_C, T_AC = -T_A M
    // (M in Dg, free after the factor)
    for (int s = 1; s < NB; s *= 2) {
        const int per = s * s, tot = (NB / (2 * s)) * per;  // outputs per pair, all pairs
        for (int e = tid; e < tot; e += 256) {
            const int pr = e / per, rr = (e % per) / s, cc = e % s, p = pr * 2 * s;
            double2 m = zero2;
            for (int t = 0; t <= cc; ++t) m = cfma(m, Rf[p + rr][p + s + t], Ti[p + s + t][p + s + cc]);
            Dg[p + rr][p + s + cc] = m;
        }
        __syncthreads();
        for (int e = tid; e < tot; e += 256) {
            const int pr = e / per, rr = (e % per) / s, cc = e % s, p = pr * 2 * s;
            double2 m = zero2;
            for (int t = rr; t < s; ++t) m = cfma(m, Ti[p + rr][p + t], Dg[p + t][p + s + cc]);
            Ti[p + rr][p + s + cc] = make_double2(-m.x, -m.y);
        }
        __syncthreads();
    }
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e / NB, c = e % NB;
        Rd[e] = Ti[i][c];
        if (i < w && c < w) T[(int64_t)(p0 + i) * ld + p0 + c] = Ti[i][c];
    }
    if (tid < w) live[p0 + tid] = (lv >> tid) & 1u;
}

// Launch B of panel k, grid (ceil(nmax / 32), batch): workgroup j with 0 < j < (panels from k
// on) writes R's tile (k, k + j) = Di^H (S_kc - R_{<k,k}^H R_{<k,c}) into S and a zero tile
// (k + j, k) into T; workgroup j < k writes T's tile (j, k) = -(T_{j, j..k-1} R_{j..k-1, k}) Di.
// Tiles of S and T written here are read by no other workgroup of the launch.
__global__ __launch_bounds__(256) void chol_panel_kernel(double2* __restrict__ S, const int* __restrict__ sizes,
                                                         int ld, int64_t stride, double2* __restrict__ T,
                                                         const double2* __restrict__ Rd, int k) {
    constexpr int NB = kCholLeftNB;
    const int b = blockIdx.y, n = sizes[b];
    const int p0 = k * NB;
    if (p0 >= n) return;  // uniform over the workgroup
    const int w = min(NB, n - p0);
    const int j = blockIdx.x;
    const int npk = (n - p0 + NB - 1) / NB;
    const bool off = j > 0 && j < npk;
    const bool inv = j < k;
    if (!off && !inv) return;  // uniform over the workgroup
    S += (int64_t)b * stride;
    T += (int64_t)b * stride;
    Rd += ((int64_t)b * ((ld + NB - 1) / NB) + k) * NB * NB;
    __shared__ CholTiles L;
    __shared__ double2 Di[NB][NB + 1], Ct[NB][NB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r4 = (lane >> 3) * 4, c4 = (lane & 7) * 4;
    const double2 zero2 = make_double2(0.0, 0.0);
    for (int e = tid; e < NB * NB; e += 256) Di[e / NB][e % NB] = Rd[e];
    const int r = tid >> 3, cq = (tid & 7) * 4;
    if (off) {
        const int c0 = p0 + j * NB, wc = min(NB, n - c0);
        double2 ac[4][4];
        chol_tile_sum(L, ac, S, T, ld, 0, p0, p0, 0, false, w, c0, wc);
        if (wv == 0)
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int rr = r4 + x, c = c4 + y;
                    double2 v = zero2;
                    if (rr < w && c < wc) {
                        const double2 sv = S[(int64_t)(p0 + rr) * ld + c0 + c];
                        v = make_double2(sv.x - ac[x][y].x, sv.y - ac[x][y].y);
                    }
                    Ct[rr][c] = v;
                }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cq + q;
            double2 yv = zero2;
            for (int m = 0; m <= r; ++m) yv = cfma_conj(yv, Di[m][r], Ct[m][c]);
            if (r < w && c < wc) S[(int64_t)(p0 + r) * ld + c0 + c] = yv;
            if (c < w && r < wc) T[(int64_t)(c0 + r) * ld + p0 + c] = zero2;
        }
        __syncthreads();  // Ct is reused
    }
    if (inv) {
        const int R0 = j * NB;
        double2 xa[4][4];
        chol_tile_sum(L, xa, S, T, ld, R0, p0, 0, R0, true, NB, p0, w);
        if (wv == 0)
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) Ct[r4 + x][c4 + y] = xa[x][y];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cq + q;
            double2 yv = zero2;
            for (int m = 0; m <= c; ++m) yv = cfma(yv, Ct[r][m], Di[m][c]);
            if (c < w) T[(int64_t)(R0 + r) * ld + p0 + c] = make_double2(-yv.x, -yv.y);
        }
    }
}

void launch_chol_left(hipStream_t st, double2* S, const int* sizes, int nmax, int ld, int64_t stride, double tau,
                      int* live, double* dmax, double2* T, double2* Rd, int batch) {
    KV_REQUIRE(nmax <= ld && batch >= 1, ERR_INTERNAL, "chol_left: side above the leading dimension");
    hipLaunchKernelGGL(diag_max_kernel, dim3(batch), dim3(256), 0, st, S, sizes, ld, stride, dmax);
    KV_HIP(hipGetLastError());
    const int np = (nmax + kCholLeftNB - 1) / kCholLeftNB;
    for (int k = 0; k < np; ++k) {
        hipLaunchKernelGGL(chol_diag_kernel, dim3(batch), dim3(256), 0, st, S, sizes, ld, stride, tau, live, dmax, T,
                           Rd, k);
        KV_HIP(hipGetLastError());
        hipLaunchKernelGGL(chol_panel_kernel, dim3(np, batch), dim3(256), 0, st, S, sizes, ld, stride, T, Rd, k);
        KV_HIP(hipGetLastError());
    }
}

// Column c of Y (n x ld, row-major) for every c < ld with live[c] == 0: seeded uniform
// complex values scaled so the column's squared norm is about dmax[0] (E|u|^2 = 2/3).
__global__ __launch_bounds__(256) void refill_cols_kernel(double2* __restrict__ Y, int n, int ld,
                                                          const int* __restrict__ live,
                                                          const double* __restrict__ dmax, uint64_t seed) {
    const uint64_t s = mix64(seed);
    const double f = sqrt(1.5 * fmax(dmax[0], 0.0) / (double)n);
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < (int64_t)n * ld; e += (int64_t)gridDim.x * 256) {
        const int c = (int)(e % ld);
        if (live[c]) continue;
        const uint64_t a = mix64(s ^ (uint64_t)(2 * e)), b = mix64(s ^ (uint64_t)(2 * e + 1));
        Y[e] = make_double2(f * ((double)(a >> 11) * 0x1.0p-52 - 1.0), f * ((double)(b >> 11) * 0x1.0p-52 - 1.0));
    }
}

void launch_refill_cols(hipStream_t st, double2* Y, int n, int ld, const int* live, const double* dmax,
                        uint64_t seed) {
    const int64_t total = (int64_t)n * ld;
    if (total <= 0) return;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(refill_cols_kernel, dim3(blocks), dim3(256), 0, st, Y, n, ld, live, dmax, seed);
    KV_HIP(hipGetLastError());
}

// part[y kc + j] = sum over rows [y rows_per, (y + 1) rows_per) of |G[r][j] - theta_j Z[r][j]|^2
// (row-major G and Z, ld ldg and ldz).  grid (ceil(kc / 64), chunks), 4 row slices per block.
__global__ __launch_bounds__(256) void resid_cols_kernel(const double2* __restrict__ G, int ldg,
                                                         const double2* __restrict__ Z, int ldz,
                                                         const double* __restrict__ theta, int n, int kc,
                                                         int rows_per, double* __restrict__ part) {
    const int l = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + l;
    const int r0 = blockIdx.y * rows_per, r1 = min(n, r0 + rows_per);
    __shared__ double red[4][64];
    double s = 0.0;
    if (j < kc) {
        const double t = theta[j];
        for (int r = r0 + sl; r < r1; r += 4) {
            const double2 g = G[(int64_t)r * ldg + j], z = Z[(int64_t)r * ldz + j];
            const double dx = g.x - t * z.x, dy = g.y - t * z.y;
            s += dx * dx + dy * dy;
        }
    }
    red[sl][l] = s;
    __syncthreads();
    if (sl == 0 && j < kc) part[(int64_t)blockIdx.y * kc + j] = (red[0][l] + red[1][l]) + (red[2][l] + red[3][l]);
}

void launch_resid_cols(hipStream_t st, const double2* G, int ldg, const double2* Z, int ldz, const double* theta,
                       int n, int kc, int rows_per, double* part) {
    if (n <= 0 || kc <= 0) return;
    const int chunks = (n + rows_per - 1) / rows_per;
    hipLaunchKernelGGL(resid_cols_kernel, dim3((kc + 63) / 64, chunks), dim3(256), 0, st, G, ldg, Z, ldz, theta, n,
                       kc, rows_per, part);
    KV_HIP(hipGetLastError());
}

// ---- overlap over matching block pairs, and the allocation-free entry point's scratch -----

__global__ __launch_bounds__(256) void dotc_pairs_partial_kernel(const double2* __restrict__ x,
                                                                 const double2* __restrict__ y,
                                                                 const DotPair* __restrict__ pairs, int npairs,
                                                                 double* __restrict__ part) {
    __shared__ double sr[256], si[256];
    double re = 0, im = 0;
    for (int p = 0; p < npairs; ++p) {
        const DotPair d = pairs[p];
        for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < d.size; e += (int64_t)gridDim.x * 256) {
            const double2 a = x[d.x_off + e], b = y[d.y_off + e];
            re += a.x * b.x + a.y * b.y;  // conj(a) * b
            im += a.x * b.y - a.y * b.x;
        }
    }
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sr[0];
        part[2 * blockIdx.x + 1] = si[0];
    }
}

void launch_dotc_pairs(hipStream_t st, const double2* x, const double2* y, const DotPair* pairs, int npairs,
                       int64_t total, double* part, double2* out) {
    const int nb = red_blocks(total);  // every block writes its partial, also with no pairs
    hipLaunchKernelGGL(dotc_pairs_partial_kernel, dim3(nb), dim3(256), 0, st, x, y, pairs, npairs, part);
    KV_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce2_final_kernel, dim3(1), dim3(256), 0, st, part, nb, reinterpret_cast<double*>(out), 2);
    KV_HIP(hipGetLastError());
}

// The allocation-free entry point's scratch.  In host code `scratch` names a host-side
// stand-in, not device memory: only the address hipGetSymbolAddress returns for the
// current device may be handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex& scratch_mutex() {
    static std::mutex m;
    return m;
}
void* scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

void noalloc_scratch_init(int device) {
    KV_REQUIRE(device >= 0 && device < kMaxDevices, ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> g(scratch_mutex());
    if (scratch_ptr[device]) return;
    int cur = -1;
    KV_HIP(hipGetDevice(&cur));
    KV_REQUIRE(cur == device, ERR_INTERNAL, "noalloc scratch: device is not current");
    void* p = nullptr;
    KV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    KV_REQUIRE(p != nullptr, ERR_HIP, "hipGetSymbolAddress returned no address for the scratch");
    scratch_ptr[device] = p;
}

void* noalloc_scratch(int device) {
    KV_REQUIRE(device >= 0 && device < kMaxDevices, ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> g(scratch_mutex());
    KV_REQUIRE(scratch_ptr[device] != nullptr, ERR_INTERNAL, "noalloc scratch not initialised on this device");
    return scratch_ptr[device];
}

}  // namespace kv
