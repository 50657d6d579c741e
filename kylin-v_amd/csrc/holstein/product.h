// Executors of the contraction plans: block-sparse product, MPO fold, transposes, the
// effective-Hamiltonian matvec and the environment update, all on the context stream.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <type_traits>
#include <vector>

#include "plan.h"
#include "tensor.h"

namespace kv {

// PREC_BF16X3: FP32 storage, GEMMs as split-BF16 products with FP32 accumulation
enum Prec { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2 };

// Event pairs recorded around contraction kernels; resolved after the host has
// synchronised the stream anyway (no extra syncs on the hot path).
struct EventLog {
    struct Rec {
        hipEvent_t a, b;
        double flop;
        int kind;      // 0 gemm (FP64 operands), 1 wfold, 2 gemm (FP32 operands)
        double bytes;  // algorithmic bytes: operands and result once each
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> spare;
    hipEvent_t get() {
        if (!spare.empty()) {
            hipEvent_t e = spare.back();
            spare.pop_back();
            return e;
        }
        hipEvent_t e;
        KV_HIP(hipEventCreate(&e));
        return e;
    }
    // call only after the stream has been synchronised
    void resolve(KernelStats& st) {
        for (auto& r : pending) {
            float ms = 0;
            KV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            if (r.kind == 0 || r.kind == 2) {
                st.gemm_ms += ms;
                st.gemm_flop += r.flop;
                st.gemm_bytes += r.bytes;
                if (r.kind == 2) {
                    st.gemm32_ms += ms;
                    st.gemm32_flop += r.flop;
                    st.gemm32_bytes += r.bytes;
                    st.gemm32_launches += 1;
                }
                st.gemm_launches += 1;
            } else {
                st.wfold_ms += ms;
                st.wfold_flop += r.flop;
                st.wfold_launches += 1;
            }
            spare.push_back(r.a);
            spare.push_back(r.b);
        }
        pending.clear();
    }
    ~EventLog() {
        for (auto& r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : spare) (void)hipEventDestroy(e);
    }
};

// Shape class of a product plan for the tile autotuner: log2 buckets of the m, n and total
// k of its largest group.
inline uint64_t gemm_shape_key(const ProductPlan& p, int prec) {
    int64_t best = -1;
    int bm = 0, bn = 0, bk = 0;
    for (const auto& g : p.groups) {
        int64_t k = 0;
        for (int s = 0; s < g.seg_count; ++s) k += p.segs[g.seg_begin + s].k;
        const int64_t w = (int64_t)g.m * g.n * std::max<int64_t>(k, 1);
        if (w > best) {
            best = w;
            bm = g.m;
            bn = g.n;
            bk = (int)std::min<int64_t>(k, 1 << 30);
        }
    }
    auto lg = [](int v) {
        int l = 0;
        while ((1 << (l + 1)) <= v && l < 20) ++l;
        return l;
    };
    return ((uint64_t)prec << 40) | ((uint64_t)p.opA << 36) | ((uint64_t)p.opB << 32) |
           ((uint64_t)lg(bm) << 20) | ((uint64_t)lg(bn) << 10) | (uint64_t)lg(bk);
}

struct DevProduct {
    ProductPlan p;
    PackedUpload lists;  // groups, segments and config 0's tiles, one upload
    DevBuf tiles[kNumGemmConfigs];  // other configs' tiles, cut on first use
    const GemmTile* tp[kNumGemmConfigs] = {nullptr};
    int ntiles[kNumGemmConfigs] = {0};
    bool cut[kNumGemmConfigs] = {false};
    void build(Device& dev, ProductPlan plan) {
        p = std::move(plan);
        lists = upload_packed(dev, p.groups, p.segs, p.tiles);
        tp[0] = lists.at<GemmTile>(2);
        ntiles[0] = (int)p.tiles.size();
        cut[0] = true;
    }
    void ensure_tiles(Device& dev, int cfg) {
        if (cut[cfg]) return;
        std::vector<GemmTile> t = cut_tiles(p.groups, cfg);
        tiles[cfg] = upload_vec(dev, t);
        tp[cfg] = tiles[cfg].as<GemmTile>();
        ntiles[cfg] = (int)t.size();
        cut[cfg] = true;
    }
    template <class T2>
    void launch(Device& dev, int cfg, const T2* A, const T2* B, T2* C, double ar, double ai) {
        ensure_tiles(dev, cfg);
        launch_gemm<T2>(dev.stream, p.opA, p.opB, A, B, C, tp[cfg], ntiles[cfg], lists.at<GemmGroup>(0),
                        lists.at<GemmSeg>(1), ar, ai, cfg);
    }
    // Tile config for this plan: cached per shape class; on a first encounter every config
    // built for the op pair runs twice on the real operands and the fastest is kept.
    template <class T2>
    int choose(Device& dev, const T2* A, const T2* B, T2* C, double ar, double ai) {
        if (!dev.autotune || p.flop < 1e8) return 0;
        const uint64_t key = gemm_shape_key(p, sizeof(T2) == sizeof(double2) ? 0 : 1);
        auto it = dev.gemm_tuned.find(key);
        if (it != dev.gemm_tuned.end()) return it->second;
        EventTimer ev;
        int best = 0;
        float best_ms = 1e30f;
        const char* vb = std::getenv("KV_VERBOSE");
        const bool verbose = vb && vb[0] == '1';
        char line[512];
        int len = std::snprintf(line, sizeof line, "[kv] autotune %s op%d%d %.3g GFLOP:",
                                sizeof(T2) == sizeof(double2) ? "f64" : "f32", p.opA, p.opB,
                                p.flop * 1e-9);
        for (int cfg = 0; cfg < kNumGemmConfigs; ++cfg) {
            if (!gemm_config_supports(cfg, p.opA, p.opB, sizeof(T2) == sizeof(double2))) continue;
            launch<T2>(dev, cfg, A, B, C, ar, ai);
            ev.start(dev.stream);
            launch<T2>(dev, cfg, A, B, C, ar, ai);
            ev.stop(dev.stream);
            const float ms = ev.ms();
            if (verbose && len < (int)sizeof line - 40)
                len += std::snprintf(line + len, sizeof line - len, " %dx%dx%d=%.3fms",
                                     kGemmConfigs[cfg].bm, kGemmConfigs[cfg].bn,
                                     kGemmConfigs[cfg].bk, ms);
            if (ms < best_ms) {
                best_ms = ms;
                best = cfg;
            }
        }
        if (verbose) std::fprintf(stderr, "%s -> %d\n", line, best);
        dev.gemm_tuned[key] = best;
        return best;
    }
    void launch_split(Device& dev, int cfg, const float2* A, const float2* B, float2* C, double ar,
                      double ai) {
        ensure_tiles(dev, cfg);
        launch_gemm_bf16x3(dev.stream, p.opA, p.opB, A, B, C, tp[cfg], ntiles[cfg], lists.at<GemmGroup>(0),
                           lists.at<GemmSeg>(1), ar, ai, cfg);
    }
    // Split-BF16 tile (64x64 or 128x128), timed once per shape class like choose().
    int choose_split(Device& dev, const float2* A, const float2* B, float2* C, double ar, double ai) {
        if (const char* t = std::getenv("KV_SPLIT_TILE")) return std::atoi(t) == 128 ? 4 : 0;
        if (!dev.autotune || p.flop < 1e8) return 0;
        const uint64_t key = gemm_shape_key(p, 2);
        auto it = dev.gemm_tuned.find(key);
        if (it != dev.gemm_tuned.end()) return it->second;
        EventTimer ev;
        int best = 0;
        float best_ms = 1e30f;
        for (int cfg : {0, 4}) {
            launch_split(dev, cfg, A, B, C, ar, ai);
            ev.start(dev.stream);
            launch_split(dev, cfg, A, B, C, ar, ai);
            ev.stop(dev.stream);
            const float ms = ev.ms();
            if (ms < best_ms) {
                best_ms = ms;
                best = cfg;
            }
        }
        dev.gemm_tuned[key] = best;
        return best;
    }
    template <class T2>
    void run(Device& dev, const T2* A, const T2* B, T2* C, double alpha_re = 1.0,
             double alpha_im = 0.0, EventLog* log = nullptr, bool split = false) {
        int cfg = 0;
        if constexpr (std::is_same<T2, float2>::value) {
            cfg = split ? choose_split(dev, A, B, C, alpha_re, alpha_im)
                        : choose<T2>(dev, A, B, C, alpha_re, alpha_im);
        } else {
            cfg = choose<T2>(dev, A, B, C, alpha_re, alpha_im);
        }
        hipEvent_t a = nullptr, b = nullptr;
        if (log) {
            a = log->get();
            b = log->get();
            KV_HIP(hipEventRecord(a, dev.stream));
        }
        if constexpr (std::is_same<T2, float2>::value) {
            if (split)
                launch_split(dev, cfg, A, B, C, alpha_re, alpha_im);
            else
                launch<T2>(dev, cfg, A, B, C, alpha_re, alpha_im);
        } else {
            KV_REQUIRE(!split, ERR_INTERNAL, "split-BF16 GEMM takes FP32 operands");
            launch<T2>(dev, cfg, A, B, C, alpha_re, alpha_im);
        }
        if (log) {
            KV_HIP(hipEventRecord(b, dev.stream));
            log->pending.push_back({a, b, p.flop, sizeof(T2) == sizeof(float2) ? 2 : 0,
                                    p.unique * (double)sizeof(T2)});
        }
    }
};

struct DevWFold {
    WFoldPlan p;
    PackedUpload lists;  // tasks and terms, one upload
    void build(Device& dev, WFoldPlan plan) {
        p = std::move(plan);
        lists = upload_packed(dev, p.tasks, p.terms);
    }
    template <class T2>
    void run(Device& dev, const T2* in, T2* out, EventLog* log = nullptr) const {
        hipEvent_t a = nullptr, b = nullptr;
        if (log) {
            a = log->get();
            b = log->get();
            KV_HIP(hipEventRecord(a, dev.stream));
        }
        launch_wfold<T2>(dev.stream, in, out, lists.at<WTask>(0), (int)p.tasks.size(),
                         lists.at<WTerm>(1), p.maxPQ);
        if (log) {
            KV_HIP(hipEventRecord(b, dev.stream));
            log->pending.push_back({a, b, p.flop, 1, 0.0});
        }
    }
};

// product<N1,N2,nc>(m1, m2, T1, T2) on the device, FP64 (holstein/product.h:61-201).
inline std::unique_ptr<Tensor> product(const std::shared_ptr<Device>& dev, const Tensor& A,
                                       const Tensor& B, int nc, char opA, char opB) {
    DevProduct dp;
    dp.build(*dev, plan_product(A.lay, B.lay, nc, opA, opB));
    auto C = Tensor::create(dev, dp.p.out, false);
    dp.run<double2>(*dev, A.d(), B.d(), C->d());
    return C;
}

// transpose(t, axes) on the device (holstein/tensor.h:244-260).
inline std::unique_ptr<Tensor> transpose(const std::shared_ptr<Device>& dev, const Tensor& X,
                                         const Idx& axes) {
    PermPlan P = plan_perm(X.lay, axes);
    auto Y = Tensor::create(dev, P.out, false);
    DevBuf tasks = upload_vec(*dev, P.tasks);
    launch_perm<double2>(dev->stream, X.d(), Y->d(), tasks.as<PermTask>(), (int)P.tasks.size(),
                         P.maxsize);
    return Y;
}

// Copy X into the (larger) layout `to`, zero elsewhere.
inline std::unique_ptr<Tensor> relayout(const std::shared_ptr<Device>& dev, const Tensor& X,
                                        const Layout& to) {
    if (same_layout(X.lay, to)) return X.clone();
    auto Y = Tensor::create(dev, to, true);
    std::vector<Copy2D> cps;
    int maxrows = 1;
    for (const auto& b : X.lay.blocks) {
        const int ob = to.find(b.qn);
        KV_REQUIRE(ob >= 0 && qn_equal(to.blocks[ob].shape, b.shape, to.rank), ERR_INTERNAL,
                   "relayout: block missing from target layout");
        const int last = b.shape[X.lay.rank - 1];
        Copy2D c{};
        c.src_off = b.off;
        c.dst_off = to.blocks[ob].off;
        c.rows = (int)(b.size / last);
        c.cols = last;
        c.src_ld = last;
        c.dst_ld = last;
        maxrows = std::max(maxrows, c.rows);
        cps.push_back(c);
    }
    if (!cps.empty()) {
        DevBuf d = upload_vec(*dev, cps);
        launch_copy2d<double2>(dev->stream, X.d(), Y->d(), d.as<Copy2D>(), (int)cps.size(),
                               maxrows, nullptr);
    }
    return Y;
}

// y + alpha x over the union of both tensors' blocks (holstein/tensor.h:173-186 appends a
// block with a new (shape, qn) and axpys into an equal one).
inline std::unique_ptr<Tensor> axpy(const std::shared_ptr<Device>& dev, const Tensor& X,
                                    const Tensor& Y, double2 alpha) {
    KV_REQUIRE(X.lay.rank == Y.lay.rank, ERR_ARG, "axpy: rank mismatch");
    for (int a = 0; a < X.lay.rank; ++a)
        KV_REQUIRE(X.lay.dirs[a] == Y.lay.dirs[a], ERR_DIRS, "axpy: leg directions differ");
    const Layout U = union_layout(X.lay, Y.lay);
    auto Xu = relayout(dev, X, U);
    auto out = relayout(dev, Y, U);
    launch_zaxpby(dev->stream, out->d(), Xu->d(), out->d(), alpha, make_double2(1.0, 0.0),
                  U.total);
    dev->sync();
    return out;
}

// Effective Hamiltonian on one bond / site, built once per Lanczos call
// (apply_double / apply_single / apply_zero, holstein/product.h:478-553).
struct Heff {
    std::shared_ptr<Device> dev;
    HeffStructure S;
    Prec prec = PREC_F64;
    DevProduct s1, s4;
    // the MPO folds; two sites in FP32 / split-BF16 also both folds as one pass (s23)
    DevWFold s2, s3, s23;
    const Tensor* L = nullptr;
    const Tensor* R = nullptr;
    DevBuf L32, R32, x32, y32, t1, t2, t3;

    void build(const std::shared_ptr<Device>& d, int nsite, const Tensor& Lt, Tensor* W1,
               Tensor* W2, const Tensor& Rt, const Layout& x, Prec p) {
        dev = d;
        prec = p;
        L = &Lt;
        R = &Rt;
        const double* w1d = W1 ? W1->host_data().data() : nullptr;
        const double* w2d = W2 ? W2->host_data().data() : nullptr;
        S = plan_heff(nsite, Lt.lay, W1 ? &W1->lay : nullptr, w1d, W2 ? &W2->lay : nullptr, w2d,
                      Rt.lay, x);
        s1.build(*dev, S.s1);
        if (nsite >= 1) s2.build(*dev, S.s2);
        if (nsite == 2) s3.build(*dev, S.s3);
        if (nsite == 2 && p != PREC_F64) s23.build(*dev, S.s23);
        s4.build(*dev, S.s4);
        // intermediates sized for FP64 so either precision can run through them
        const size_t es = sizeof(double2);
        t1 = DevBuf(dev->pool.get(), (size_t)S.s1.out.total * es);
        if (nsite >= 1) t2 = DevBuf(dev->pool.get(), (size_t)S.s2.out.total * es);
        if (nsite == 2) t3 = DevBuf(dev->pool.get(), (size_t)S.s3.out.total * es);
        if (p != PREC_F64) {
            L32 = DevBuf(dev->pool.get(), (size_t)Lt.lay.total * sizeof(float2));
            R32 = DevBuf(dev->pool.get(), (size_t)Rt.lay.total * sizeof(float2));
            x32 = DevBuf(dev->pool.get(), (size_t)S.xl.total * sizeof(float2));
            y32 = DevBuf(dev->pool.get(), (size_t)S.xl.total * sizeof(float2));
            launch_f64_to_f32(dev->stream, Lt.d(), L32.as<float2>(), Lt.lay.total);
            launch_f64_to_f32(dev->stream, Rt.d(), R32.as<float2>(), Rt.lay.total);
        }
    }
    const Layout& layout() const { return S.xl; }
    double flop() const { return S.flop; }

    template <class T2>
    void chain(const T2* Ld, const T2* Rd, const T2* x, T2* y, double alpha, EventLog* log,
               bool split = false) {
        s1.run<T2>(*dev, Ld, x, t1.as<T2>(), 1.0, 0.0, log, split);
        const T2* last = t1.as<T2>();
        // FP64 keeps the two folds one after the other (the reference's order of sums);
        // the FP32 operands take the one-pass composite, which saves a pass over the
        // intermediate at the cost of a different rounding order
        const bool one_pass = S.nsite == 2 && !std::is_same<T2, double2>::value;
        if (one_pass) {
            s23.run<T2>(*dev, last, t3.as<T2>(), log);
            last = t3.as<T2>();
        } else {
            if (S.nsite >= 1) {
                s2.run<T2>(*dev, last, t2.as<T2>(), log);
                last = t2.as<T2>();
            }
            if (S.nsite == 2) {
                s3.run<T2>(*dev, last, t3.as<T2>(), log);
                last = t3.as<T2>();
            }
        }
        s4.run<T2>(*dev, last, Rd, y, alpha, 0.0, log, split);
    }
    // y = alpha * H_eff x, x and y FP64 slabs in layout() order; force64 runs the
    // contractions in FP64 whatever the plan's precision
    void apply(const double2* x, double2* y, double alpha, EventLog* log = nullptr,
               bool force64 = false) {
        if (prec == PREC_F64 || force64) {
            chain<double2>(L->d(), R->d(), x, y, alpha, log);
        } else {
            launch_f64_to_f32(dev->stream, x, x32.as<float2>(), S.xl.total);
            chain<float2>(L32.as<float2>(), R32.as<float2>(), x32.as<float2>(), y32.as<float2>(),
                          1.0, log, prec == PREC_BF16X3);
            launch_f32_to_f64(dev->stream, y32.as<float2>(), y, alpha, S.xl.total);
        }
    }
};

// env_sweep (holstein/product.h:488-521): L2R  L' = bra^dag . (L . ket . W);
// R2L: the same after transposing bra/ket (2,1,0) and W (0,3,2,1).
inline std::unique_ptr<Tensor> env_update(const std::shared_ptr<Device>& dev, const Tensor& env,
                                          const Tensor* bra, Tensor& W, const Tensor& ket,
                                          bool l2r, Prec prec, EventLog* log = nullptr,
                                          double* flop = nullptr) {
    std::unique_ptr<Tensor> kt, bt;
    const Tensor* k = &ket;
    const Tensor* b = bra ? bra : &ket;
    Layout wl;
    std::vector<double> wt;
    const Layout* wlay = &W.lay;
    const double* wd = W.host_data().data();
    if (!l2r) {
        const Idx ax{2, 1, 0, 0, 0, 0};
        kt = transpose(dev, ket, ax);
        k = kt.get();
        if (bra) {
            bt = transpose(dev, *bra, ax);
            b = bt.get();
        } else {
            b = k;
        }
        host_perm(W.lay, wd, Idx{0, 3, 2, 1, 0, 0}, wl, wt);
        wlay = &wl;
        wd = wt.data();
    }
    EnvStructure E = plan_env_l2r(env.lay, b->lay, *wlay, wd, k->lay);
    if (flop) *flop += E.flop;
    DevProduct s1, s3;
    DevWFold s2;
    s1.build(*dev, E.s1);
    s2.build(*dev, E.s2);
    s3.build(*dev, E.s3);
    auto out = Tensor::create(dev, E.s3.out, false);
    if (prec == PREC_F64) {
        DevBuf t1(dev->pool.get(), (size_t)E.s1.out.total * sizeof(double2));
        DevBuf t2(dev->pool.get(), (size_t)E.s2.out.total * sizeof(double2));
        s1.run<double2>(*dev, env.d(), k->d(), t1.as<double2>(), 1.0, 0.0, log);
        s2.run<double2>(*dev, t1.as<double2>(), t2.as<double2>(), log);
        s3.run<double2>(*dev, b->d(), t2.as<double2>(), out->d(), 1.0, 0.0, log);
        dev->sync();
    } else {
        DevBuf e32(dev->pool.get(), (size_t)env.lay.total * sizeof(float2));
        DevBuf k32(dev->pool.get(), (size_t)k->lay.total * sizeof(float2));
        DevBuf b32(dev->pool.get(), (size_t)b->lay.total * sizeof(float2));
        DevBuf o32(dev->pool.get(), (size_t)E.s3.out.total * sizeof(float2));
        DevBuf t1(dev->pool.get(), (size_t)E.s1.out.total * sizeof(float2));
        DevBuf t2(dev->pool.get(), (size_t)E.s2.out.total * sizeof(float2));
        launch_f64_to_f32(dev->stream, env.d(), e32.as<float2>(), env.lay.total);
        launch_f64_to_f32(dev->stream, k->d(), k32.as<float2>(), k->lay.total);
        launch_f64_to_f32(dev->stream, b->d(), b32.as<float2>(), b->lay.total);
        s1.run<float2>(*dev, e32.as<float2>(), k32.as<float2>(), t1.as<float2>(), 1.0, 0.0, log);
        s2.run<float2>(*dev, t1.as<float2>(), t2.as<float2>(), log);
        s3.run<float2>(*dev, b32.as<float2>(), t2.as<float2>(), o32.as<float2>(), 1.0, 0.0, log);
        launch_f32_to_f64(dev->stream, o32.as<float2>(), out->d(), 1.0, E.s3.out.total);
        dev->sync();
    }
    return out;
}

}  // namespace kv
