// Contraction planners: turn the block structures of the operands into device work
// lists (grouped-GEMM tiles, MPO-fold tasks, permutation tasks).  Host-only C++ (no
// HIP), so the plans can be checked on a CPU against the oracle.
//
// A plan is built once per bond and reused for every Lanczos matvec on that bond; the
// executors in product.h upload its arrays to the device once.
#pragma once

#include <array>
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "../kv_types.h"
#include "layout.h"

namespace kv {

inline int op_code(char op) {
    if (op == 'N') return 0;
    if (op == 'T') return 1;
    if (op == 'C') return 2;
    throw Error(ERR_ARG, std::string("bad op code ") + op);
}

// Output tiles of a group list for tile config cfg (kGemmConfigs), row-major per group.
inline std::vector<GemmTile> cut_tiles(const std::vector<GemmGroup>& groups, int cfg) {
    const int bm = kGemmConfigs[cfg].bm, bn = kGemmConfigs[cfg].bn;
    std::vector<GemmTile> t;
    for (int gi = 0; gi < (int)groups.size(); ++gi) {
        const int tm = (groups[gi].m + bm - 1) / bm, tn = (groups[gi].n + bn - 1) / bn;
        for (int a = 0; a < tm; ++a)
            for (int b = 0; b < tn; ++b) t.push_back(GemmTile{gi, a, b, 0});
    }
    return t;
}

struct ProductPlan {
    int opA = 0, opB = 0;
    Layout out;
    std::vector<GemmTile> tiles;
    std::vector<GemmGroup> groups;
    std::vector<GemmSeg> segs;
    double flop = 0;     // 8*m*n*k summed over block pairs
    double bytes = 0;    // operand + result bytes per complex element size 1
    double unique = 0;   // elements of A, B and the result, each counted once
};

// product<N1,N2,nc>(m1, m2, T1, T2) (holstein/product.h:61-201).
// 'N' contracts the last nc legs of A and the first nc legs of B; 'T'/'C' contract the
// first nc legs of A / the last nc legs of B, 'C' conjugating and negating the kept
// legs' dirs.  Blocks pair when all contracted legs agree in qn and shape; pairs with the
// same output charge tuple become the K segments of one output group.  With a target
// layout, output offsets come from it and target blocks without contributions become
// zero groups (so the kernel writes every element of the target).
inline ProductPlan plan_product(const Layout& A, const Layout& B, int nc, char opA, char opB,
                                const Layout* target = nullptr) {
    ProductPlan P;
    P.opA = op_code(opA);
    P.opB = op_code(opB);
    const int n1 = A.rank, n2 = B.rank;
    KV_REQUIRE(nc >= 1 && nc <= n1 && nc <= n2, ERR_ARG, "bad contraction count");
    KV_REQUIRE(n1 + n2 - 2 * nc >= 1 && n1 + n2 - 2 * nc <= MAXR, ERR_ARG, "bad result rank");
    const int iax1 = (opA == 'N') ? n1 - nc : 0;
    const int iax2 = (opB == 'N') ? 0 : n2 - nc;
    const int k1b = (opA == 'N') ? 0 : nc, k1e = (opA == 'N') ? n1 - nc : n1;
    const int k2b = (opB == 'N') ? nc : 0, k2e = (opB == 'N') ? n2 : n2 - nc;
    const int fac1 = (opA == 'C') ? -1 : 1, fac2 = (opB == 'C') ? -1 : 1;
    for (int ax = 0; ax < nc; ++ax)
        if (A.dirs[iax1 + ax] * fac1 != -1 * B.dirs[iax2 + ax] * fac2)
            throw Error(ERR_DIRS, "Contracted legs' directions mismatch");
    Layout out;
    out.rank = n1 + n2 - 2 * nc;
    {
        int r = 0;
        for (int a = k1b; a < k1e; ++a) out.dirs[r++] = A.dirs[a] * fac1;
        for (int a = k2b; a < k2e; ++a) out.dirs[r++] = B.dirs[a] * fac2;
    }
    // index B blocks by contracted (qn, shape)
    std::map<std::vector<int>, std::vector<int>> bidx;
    for (int j = 0; j < (int)B.blocks.size(); ++j) {
        std::vector<int> key(2 * nc);
        for (int ax = 0; ax < nc; ++ax) {
            key[ax] = B.blocks[j].qn[iax2 + ax];
            key[nc + ax] = B.blocks[j].shape[iax2 + ax];
        }
        bidx[key].push_back(j);
    }
    struct Acc {
        Idx shape;
        std::vector<std::pair<int, int>> pairs;
    };
    std::map<std::vector<int>, Acc> groups;
    for (int i = 0; i < (int)A.blocks.size(); ++i) {
        std::vector<int> key(2 * nc);
        for (int ax = 0; ax < nc; ++ax) {
            key[ax] = A.blocks[i].qn[iax1 + ax];
            key[nc + ax] = A.blocks[i].shape[iax1 + ax];
        }
        auto it = bidx.find(key);
        if (it == bidx.end()) continue;
        for (int j : it->second) {
            std::vector<int> q(out.rank);
            Idx sh{};
            int r = 0;
            for (int a = k1b; a < k1e; ++a, ++r) {
                q[r] = A.blocks[i].qn[a];
                sh[r] = A.blocks[i].shape[a];
            }
            for (int a = k2b; a < k2e; ++a, ++r) {
                q[r] = B.blocks[j].qn[a];
                sh[r] = B.blocks[j].shape[a];
            }
            auto g = groups.find(q);
            if (g == groups.end()) {
                Acc acc;
                acc.shape = sh;
                acc.pairs.push_back({i, j});
                groups.emplace(q, std::move(acc));
            } else {
                KV_REQUIRE(qn_equal(g->second.shape, sh, out.rank), ERR_ARG,
                           "product: equal output charges with different shapes");
                g->second.pairs.push_back({i, j});
            }
        }
    }
    if (target) {
        KV_REQUIRE(target->rank == out.rank, ERR_INTERNAL, "target layout rank mismatch");
        for (int r = 0; r < out.rank; ++r)
            KV_REQUIRE(target->dirs[r] == out.dirs[r], ERR_INTERNAL, "target layout dirs mismatch");
        out.blocks = target->blocks;
        out.total = target->total;
    } else {
        for (auto& kvp : groups) {
            BlockDesc bd;
            for (int r = 0; r < out.rank; ++r) bd.qn[r] = kvp.first[r];
            bd.shape = kvp.second.shape;
            out.blocks.push_back(bd);
        }
        out.finalize();
    }
    const int nkept1 = k1e - k1b;
    std::vector<char> covered(out.blocks.size(), 0);
    for (auto& kvp : groups) {
        Idx q{};
        for (int r = 0; r < out.rank; ++r) q[r] = kvp.first[r];
        const int ob = out.find(q);
        KV_REQUIRE(ob >= 0, ERR_INTERNAL, "product output block missing from target layout");
        KV_REQUIRE(qn_equal(out.blocks[ob].shape, kvp.second.shape, out.rank), ERR_INTERNAL,
                   "product output shape differs from target layout");
        covered[ob] = 1;
        GemmGroup g{};
        g.c_off = out.blocks[ob].off;
        int64_t m = 1, n = 1;
        for (int r = 0; r < nkept1; ++r) m *= kvp.second.shape[r];
        for (int r = nkept1; r < out.rank; ++r) n *= kvp.second.shape[r];
        KV_REQUIRE(m < (1ll << 31) && n < (1ll << 31), ERR_ARG, "block too large");
        g.m = (int)m;
        g.n = (int)n;
        g.ldc = (int)n;
        g.seg_begin = (int)P.segs.size();
        g.seg_count = (int)kvp.second.pairs.size();
        for (auto& pr : kvp.second.pairs) {
            const BlockDesc& ba = A.blocks[pr.first];
            const BlockDesc& bb = B.blocks[pr.second];
            int64_t k = 1;
            for (int ax = 0; ax < nc; ++ax) k *= ba.shape[iax1 + ax];
            GemmSeg s{};
            s.a_off = ba.off;
            s.b_off = bb.off;
            s.k = (int)k;
            s.lda = (opA == 'N') ? (int)k : (int)m;
            s.ldb = (opB == 'N') ? (int)n : (int)k;
            P.segs.push_back(s);
            P.flop += 8.0 * (double)m * (double)n * (double)k;
            P.bytes += (double)(m * k + k * n);
        }
        P.bytes += (double)(m * n);
        P.groups.push_back(g);
    }
    P.unique = (double)A.total + (double)B.total;
    // zero groups for target blocks that receive nothing
    for (size_t ob = 0; ob < out.blocks.size(); ++ob) {
        if (covered[ob]) continue;
        GemmGroup g{};
        g.c_off = out.blocks[ob].off;
        const int last = out.blocks[ob].shape[out.rank - 1];
        g.n = last;
        g.m = (int)(out.blocks[ob].size / last);
        g.ldc = last;
        g.seg_begin = (int)P.segs.size();
        g.seg_count = 0;
        P.groups.push_back(g);
    }
    P.tiles = cut_tiles(P.groups, 0);
    P.out = std::move(out);
    P.unique += (double)P.out.total;
    return P;
}

struct WFoldPlan {
    Layout out;
    std::vector<WTask> tasks;
    std::vector<WTerm> terms;
    int maxP = 0;
    int64_t maxPQ = 0;  // largest task's P x Q
    double flop = 0;
    // host-side provenance for compose_wfold: per task (output block, s', wR), -1s for a
    // zero task; per term (input block, wL, s)
    std::vector<std::array<int, 3>> task_src, term_src;
};

// MPO fold (the env_mps_mpo family, holstein/product.h:203-477):
//   Y[pre, s', wR, post] = sum_{wL, s} W[s', wL, s, wR] X[pre, wL, s, post]
// where X legs p, p+1 pair with W legs 1 (wL) and 2 (s) on equal qn and shape.  W is
// tiny and sparse, so each nonzero W entry becomes a scaled slice term; every (s', wR)
// row of every output block is one task summing its terms.
inline WFoldPlan plan_wfold(const Layout& X, int p, const Layout& W, const double* Wd,
                            const Layout* target = nullptr) {
    WFoldPlan F;
    KV_REQUIRE(W.rank == 4, ERR_ARG, "MPO tensor must have rank 4");
    KV_REQUIRE(p >= 0 && p + 2 <= X.rank, ERR_ARG, "bad MPO fold position");
    Layout out;
    out.rank = X.rank;
    {
        int r = 0;
        for (int a = 0; a < p; ++a) out.dirs[r++] = X.dirs[a];
        out.dirs[r++] = W.dirs[0];
        out.dirs[r++] = W.dirs[3];
        for (int a = p + 2; a < X.rank; ++a) out.dirs[r++] = X.dirs[a];
    }
    std::map<std::vector<int>, std::vector<int>> widx;  // (wL qn, s qn, wL dim, s dim)
    for (int j = 0; j < (int)W.blocks.size(); ++j) {
        const auto& b = W.blocks[j];
        widx[{b.qn[1], b.qn[2], b.shape[1], b.shape[2]}].push_back(j);
    }
    struct Acc {
        Idx shape;
        std::vector<std::pair<int, int>> pairs;
    };
    std::map<std::vector<int>, Acc> groups;
    for (int i = 0; i < (int)X.blocks.size(); ++i) {
        const auto& xb = X.blocks[i];
        auto it = widx.find({xb.qn[p], xb.qn[p + 1], xb.shape[p], xb.shape[p + 1]});
        if (it == widx.end()) continue;
        for (int j : it->second) {
            const auto& wb = W.blocks[j];
            std::vector<int> q(out.rank);
            Idx sh{};
            int r = 0;
            for (int a = 0; a < p; ++a, ++r) {
                q[r] = xb.qn[a];
                sh[r] = xb.shape[a];
            }
            q[r] = wb.qn[0];
            sh[r++] = wb.shape[0];
            q[r] = wb.qn[3];
            sh[r++] = wb.shape[3];
            for (int a = p + 2; a < X.rank; ++a, ++r) {
                q[r] = xb.qn[a];
                sh[r] = xb.shape[a];
            }
            auto g = groups.find(q);
            if (g == groups.end()) {
                Acc acc;
                acc.shape = sh;
                acc.pairs.push_back({i, j});
                groups.emplace(q, std::move(acc));
            } else {
                KV_REQUIRE(qn_equal(g->second.shape, sh, out.rank), ERR_ARG,
                           "MPO fold: equal output charges with different shapes");
                g->second.pairs.push_back({i, j});
            }
        }
    }
    if (target) {
        out.blocks = target->blocks;
        out.total = target->total;
    } else {
        for (auto& kvp : groups) {
            BlockDesc bd;
            for (int r = 0; r < out.rank; ++r) bd.qn[r] = kvp.first[r];
            bd.shape = kvp.second.shape;
            out.blocks.push_back(bd);
        }
        out.finalize();
    }
    std::vector<char> covered(out.blocks.size(), 0);
    for (auto& kvp : groups) {
        Idx q{};
        for (int r = 0; r < out.rank; ++r) q[r] = kvp.first[r];
        const int ob = out.find(q);
        KV_REQUIRE(ob >= 0, ERR_INTERNAL, "MPO fold output block missing from target");
        covered[ob] = 1;
        const BlockDesc& O = out.blocks[ob];
        int64_t Pd = 1, Qd = 1;
        for (int a = 0; a < p; ++a) Pd *= O.shape[a];
        for (int a = p + 2; a < out.rank; ++a) Qd *= O.shape[a];
        const int Sp = O.shape[p], WR = O.shape[p + 1];
        KV_REQUIRE(Pd < (1ll << 31) && Qd < (1ll << 31), ERR_ARG, "block too large");
        F.maxP = std::max<int>(F.maxP, (int)Pd);
        KV_REQUIRE(Pd * Qd < (1ll << 31), ERR_ARG, "MPO fold: block too large");
        F.maxPQ = std::max<int64_t>(F.maxPQ, Pd * Qd);
        for (int sp = 0; sp < Sp; ++sp)
            for (int wr = 0; wr < WR; ++wr) {
                WTask t{};
                t.out_base = O.off + ((int64_t)sp * WR + wr) * Qd;
                t.out_pstride = (int64_t)Sp * WR * Qd;
                t.P = (int)Pd;
                t.Q = (int)Qd;
                t.term_begin = (int)F.terms.size();
                for (auto& pr : kvp.second.pairs) {
                    const BlockDesc& xb = X.blocks[pr.first];
                    const BlockDesc& wb = W.blocks[pr.second];
                    const int WL = xb.shape[p], S = xb.shape[p + 1];
                    for (int wl = 0; wl < WL; ++wl)
                        for (int s = 0; s < S; ++s) {
                            const int64_t e =
                                wb.off + (((int64_t)sp * WL + wl) * S + s) * WR + wr;
                            const double re = Wd[2 * e], im = Wd[2 * e + 1];
                            if (re == 0.0 && im == 0.0) continue;
                            WTerm w{};
                            w.in_base = xb.off + ((int64_t)wl * S + s) * Qd;
                            w.in_pstride = (int64_t)WL * S * Qd;
                            w.re = re;
                            w.im = im;
                            F.terms.push_back(w);
                            F.term_src.push_back({pr.first, wl, s});
                            F.flop += 8.0 * (double)Pd * (double)Qd;
                        }
                }
                t.term_count = (int)F.terms.size() - t.term_begin;
                F.tasks.push_back(t);
                F.task_src.push_back({ob, sp, wr});
            }
    }
    for (size_t ob = 0; ob < out.blocks.size(); ++ob) {
        if (covered[ob]) continue;
        WTask t{};
        t.out_base = out.blocks[ob].off;
        t.out_pstride = 0;
        t.P = 1;
        KV_REQUIRE(out.blocks[ob].size < (1ll << 31), ERR_ARG, "block too large");
        t.Q = (int)out.blocks[ob].size;
        t.term_begin = (int)F.terms.size();
        t.term_count = 0;
        F.maxP = std::max(F.maxP, 1);
        F.maxPQ = std::max<int64_t>(F.maxPQ, t.Q);
        F.tasks.push_back(t);
        F.task_src.push_back({-1, -1, -1});
    }
    F.out = std::move(out);
    return F;
}

// The two folds of the two-site H_eff as one pass: F1 = plan_wfold(X, 1, W1) takes
// X[pre, wL, s1, s2, post] to Y[pre, s1', w1, s2, post] and F2 = plan_wfold(Y, 2, W2) takes
// Y to Z[pre, s1', s2', wR, post].  The composite reads X directly:
//   Z[pre, s1', s2', wR, post] = sum_{w1, s2} W2[s2', w1, s2, wR]
//                                  sum_{wL, s1} W1[s1', wL, s1, w1] X[pre, wL, s1, s2, post],
// one task per (F2 task, s1') over P = pre, Q = post, with a term per pair of nonzero W
// entries (weight W2 W1, X's slice of F1's term shifted to s2).  Z's layout is F2's.
inline WFoldPlan compose_wfold(const WFoldPlan& F1, const WFoldPlan& F2, const Layout& Y) {
    KV_REQUIRE(F1.term_src.size() == F1.terms.size() && F2.term_src.size() == F2.terms.size() &&
                   F1.task_src.size() == F1.tasks.size() && F2.task_src.size() == F2.tasks.size(),
               ERR_INTERNAL, "compose_wfold: plans without provenance");
    WFoldPlan C;
    C.out = F2.out;
    // F1's task writing Y block yb at (s1', w1)
    std::map<std::array<int, 3>, int> f1_task;
    for (size_t t = 0; t < F1.tasks.size(); ++t)
        if (F1.task_src[t][0] >= 0) f1_task[F1.task_src[t]] = (int)t;
    for (size_t t2 = 0; t2 < F2.tasks.size(); ++t2) {
        const WTask& T2 = F2.tasks[t2];
        if (F2.task_src[t2][0] < 0) {  // an output block no term reaches: zeros
            C.tasks.push_back(T2);
            C.tasks.back().term_begin = (int)C.terms.size();
            C.task_src.push_back({-1, -1, -1});
            C.maxP = std::max(C.maxP, T2.P);
            C.maxPQ = std::max<int64_t>(C.maxPQ, (int64_t)T2.P * T2.Q);
            continue;
        }
        const BlockDesc& Ob = C.out.blocks[F2.task_src[t2][0]];
        const int S1 = Ob.shape[1];  // s1' (F2's pre is (pre, s1'))
        KV_REQUIRE(S1 > 0 && T2.P % S1 == 0, ERR_INTERNAL, "compose_wfold: pre does not end in s1'");
        for (int s1 = 0; s1 < S1; ++s1) {
            WTask t{};
            t.out_base = T2.out_base + (int64_t)s1 * T2.out_pstride;
            t.out_pstride = (int64_t)S1 * T2.out_pstride;
            t.P = T2.P / S1;
            t.Q = T2.Q;
            t.term_begin = (int)C.terms.size();
            for (int k2 = T2.term_begin; k2 < T2.term_begin + T2.term_count; ++k2) {
                const WTerm& w2 = F2.terms[k2];
                const auto& src = F2.term_src[k2];  // (Y block, w1, s2)
                auto it = f1_task.find({src[0], s1, src[1]});
                if (it == f1_task.end()) continue;  // that slice of Y is zero
                const WTask& T1 = F1.tasks[it->second];
                KV_REQUIRE(T1.P == t.P, ERR_INTERNAL, "compose_wfold: pre extents differ");
                for (int k1 = T1.term_begin; k1 < T1.term_begin + T1.term_count; ++k1) {
                    const WTerm& w1 = F1.terms[k1];
                    WTerm w{};
                    w.in_base = w1.in_base + (int64_t)src[2] * t.Q;
                    w.in_pstride = w1.in_pstride;
                    w.re = w2.re * w1.re - w2.im * w1.im;
                    w.im = w2.re * w1.im + w2.im * w1.re;
                    C.terms.push_back(w);
                    C.term_src.push_back({-1, -1, -1});
                    C.flop += 8.0 * (double)t.P * (double)t.Q;
                }
            }
            t.term_count = (int)C.terms.size() - t.term_begin;
            C.tasks.push_back(t);
            C.task_src.push_back({-1, -1, -1});
            C.maxP = std::max(C.maxP, t.P);
            C.maxPQ = std::max<int64_t>(C.maxPQ, (int64_t)t.P * t.Q);
        }
    }
    (void)Y;
    return C;
}

struct PermPlan {
    Layout out;
    std::vector<PermTask> tasks;
    int64_t maxsize = 0;
};

// Leg permutation out-leg r = in-leg axes[r] (holstein/tensor.h:244-260).
inline PermPlan plan_perm(const Layout& X, const Idx& axes) {
    PermPlan P;
    P.out.rank = X.rank;
    for (int r = 0; r < X.rank; ++r) P.out.dirs[r] = X.dirs[axes[r]];
    for (const auto& b : X.blocks) {
        BlockDesc o;
        for (int r = 0; r < X.rank; ++r) {
            o.qn[r] = b.qn[axes[r]];
            o.shape[r] = b.shape[axes[r]];
        }
        P.out.blocks.push_back(o);
    }
    P.out.finalize();
    for (const auto& b : X.blocks) {
        Idx q{};
        for (int r = 0; r < X.rank; ++r) q[r] = b.qn[axes[r]];
        const int ob = P.out.find(q);
        PermTask t{};
        t.in_off = b.off;
        t.out_off = P.out.blocks[ob].off;
        t.size = b.size;
        t.rank = X.rank;
        int64_t istr[MAXR];
        int64_t s = 1;
        for (int r = X.rank - 1; r >= 0; --r) {
            istr[r] = s;
            s *= b.shape[r];
        }
        for (int r = 0; r < X.rank; ++r) {
            t.oshape[r] = b.shape[axes[r]];
            t.istride[r] = istr[axes[r]];
        }
        P.maxsize = std::max(P.maxsize, b.size);
        P.tasks.push_back(t);
    }
    return P;
}

// Host-side permutation of a small tensor (used for MPO tensors, which live on the host).
inline void host_perm(const Layout& X, const double* d, const Idx& axes, Layout& out,
                      std::vector<double>& od) {
    PermPlan P = plan_perm(X, axes);
    out = P.out;
    od.assign(2 * (size_t)out.total, 0.0);
    for (const auto& t : P.tasks) {
        for (int64_t e = 0; e < t.size; ++e) {
            int64_t rem = e, src = 0;
            for (int r = t.rank - 1; r >= 0; --r) {
                const int64_t idx = rem % t.oshape[r];
                rem /= t.oshape[r];
                src += idx * t.istride[r];
            }
            od[2 * (t.out_off + e)] = d[2 * (t.in_off + src)];
            od[2 * (t.out_off + e) + 1] = d[2 * (t.in_off + src) + 1];
        }
    }
}

// Effective-Hamiltonian plan.  nsite = 2: apply_double (holstein/product.h:478-486) as
//   T1 = L.x (product<3,4,1>), T2 = fold W1 at legs (1,2), T3 = fold W2 at legs (2,3),
//   y = T3.R^T (product<5,3,2>, N/T);
// nsite = 1: apply_single (:547-553) T1 = L.x, T2 = fold W at (1,2), y = T2.R^T;
// nsite = 0: apply_zero (:540-545) T1 = L.C, y = T1.R^T.
// The vector layout xl is the closure of x's blocks under H_eff, so every Krylov vector
// shares one slab layout.
struct HeffStructure {
    int nsite = 2;
    Layout xl;
    ProductPlan s1, s4;
    WFoldPlan s2, s3;
    WFoldPlan s23;  // two sites: both folds as one pass (compose_wfold)
    double flop = 0;
};

inline HeffStructure plan_heff(int nsite, const Layout& L, const Layout* W1, const double* W1d,
                               const Layout* W2, const double* W2d, const Layout& R,
                               const Layout& x) {
    KV_REQUIRE(nsite >= 0 && nsite <= 2, ERR_ARG, "nsite must be 0, 1 or 2");
    KV_REQUIRE(x.rank == nsite + 2, ERR_ARG, "x rank does not match nsite");
    HeffStructure H;
    H.nsite = nsite;
    Layout cur = x;
    for (int it = 0; it < 16; ++it) {
        H.s1 = plan_product(L, cur, 1, 'N', 'N');
        const Layout* t = &H.s1.out;
        if (nsite >= 1) {
            H.s2 = plan_wfold(*t, 1, *W1, W1d);
            t = &H.s2.out;
        }
        if (nsite == 2) {
            H.s3 = plan_wfold(*t, 2, *W2, W2d);
            t = &H.s3.out;
        }
        ProductPlan probe = plan_product(*t, R, 2, 'N', 'T');
        Layout u = union_layout(cur, probe.out);
        if (same_layout(u, cur)) break;
        KV_REQUIRE(it < 15, ERR_INTERNAL, "H_eff block closure did not converge");
        cur = u;
    }
    const Layout* t = nsite == 2 ? &H.s3.out : (nsite == 1 ? &H.s2.out : &H.s1.out);
    H.s4 = plan_product(*t, R, 2, 'N', 'T', &cur);
    H.xl = cur;
    if (nsite == 2) H.s23 = compose_wfold(H.s2, H.s3, H.s2.out);
    H.flop = H.s1.flop + H.s2.flop + H.s3.flop + H.s4.flop;
    return H;
}

// Environment update env_sweep (holstein/product.h:488-521).  L2R:
//   lm = env.ket (product<3,3,1>), lmo = fold W at (1,2), out = bra^dag.lmo (product<3,4,2>, C/N).
// R2L applies the same after transposing bra/ket by (2,1,0) and W by (0,3,2,1); the
// caller does those permutations.
struct EnvStructure {
    ProductPlan s1, s3;
    WFoldPlan s2;
    double flop = 0;
};

inline EnvStructure plan_env_l2r(const Layout& env, const Layout& bra, const Layout& W,
                                 const double* Wd, const Layout& ket) {
    EnvStructure E;
    E.s1 = plan_product(env, ket, 1, 'N', 'N');
    E.s2 = plan_wfold(E.s1.out, 1, W, Wd);
    E.s3 = plan_product(bra, E.s2.out, 2, 'C', 'N');
    E.flop = E.s1.flop + E.s2.flop + E.s3.flop;
    return E;
}

}  // namespace kv
