// extern "C" boundary of the engine (declared in include/kylinv.h).  Every entry point
// catches C++ exceptions and turns them into status codes plus a per-context message;
// nothing here calls exit().
#include "../../include/kylinv.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "holstein/io.h"
#include "holstein/krylov.h"
#include "holstein/tdvp.h"

struct kv_ctx {
    std::shared_ptr<kv::Device> dev;
    std::unique_ptr<kv::Solver> sol;
    std::string err;
};

struct kv_tensor {
    kv_ctx* ctx;
    std::unique_ptr<kv::Tensor> t;
};

struct kv_tdvp {
    kv_ctx* ctx;
    std::unique_ptr<kv::TDVP> s;
};

struct kv_heff_plan {
    kv_ctx* ctx;
    kv::Heff H;
};

namespace {

template <class F>
int guard(kv_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->err.clear();
        return KV_OK;
    } catch (const kv::Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return KV_ERR_INTERNAL;
    }
}

kv_tensor* wrap(kv_ctx* ctx, std::unique_ptr<kv::Tensor> t) {
    kv_tensor* h = new kv_tensor;
    h->ctx = ctx;
    h->t = std::move(t);
    return h;
}

kv::Prec prec_of(int p) {
    KV_REQUIRE(p == KV_FP64 || p == KV_FP32 || p == KV_BF16X3, kv::ERR_ARG, "unknown precision");
    return p == KV_FP64 ? kv::PREC_F64 : p == KV_FP32 ? kv::PREC_F32 : kv::PREC_BF16X3;
}

// Calls f(DotPair{b block offset, a block offset, size}) for every block of a whose qn and
// shape b also has, in a's block order (overlap(a, b) = <b|a>, holstein/tensor.h:225-241).
template <class F>
void for_overlap_pairs(const kv::Tensor& A, const kv::Tensor& B, F&& f) {
    for (const auto& ba : A.lay.blocks) {
        const int j = B.lay.find(ba.qn);
        if (j < 0 || !kv::qn_equal(B.lay.blocks[j].shape, ba.shape, A.lay.rank)) continue;
        f(kv::DotPair{B.lay.blocks[j].off, ba.off, ba.size});
    }
}

kv::Tensor* need(const kv_tensor* t, const char* what) {
    KV_REQUIRE(t && t->t, kv::ERR_ARG, std::string("null tensor: ") + what);
    return t->t.get();
}

}  // namespace

extern "C" {

int kv_ctx_create(int device, kv_ctx** out) {
    if (!out) return KV_ERR_ARG;
    try {
        kv_ctx* c = new kv_ctx;
        c->dev = std::make_shared<kv::Device>(device);
        c->sol.reset(new kv::Solver(*c->dev));
        kv::noalloc_scratch_init(device);  // the Device constructor made `device` current
        *out = c;
        return KV_OK;
    } catch (const kv::Error& e) {
        *out = nullptr;
        return e.code;
    } catch (...) {
        *out = nullptr;
        return KV_ERR_INTERNAL;
    }
}

int kv_ctx_destroy(kv_ctx* ctx) {
    if (!ctx) return KV_ERR_ARG;
    delete ctx;
    return KV_OK;
}

const char* kv_last_error(const kv_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int kv_ctx_trim(kv_ctx* ctx) {
    if (!ctx) return KV_ERR_ARG;
    return guard(ctx, [&] {
        ctx->dev->sync();
        ctx->dev->pool->trim(0);
    });
}

int kv_sync(kv_ctx* ctx) {
    if (!ctx) return KV_ERR_ARG;
    return guard(ctx, [&] { ctx->dev->sync(); });
}

int kv_tensor_create(kv_ctx* ctx, int rank, const int* dirs, int nblocks, const int* qn,
                     const int* shape, kv_tensor** out) {
    if (!ctx || !out) return KV_ERR_ARG;
    return guard(ctx, [&] {
        KV_REQUIRE(rank >= 1 && rank <= kv::MAXR, kv::ERR_ARG, "rank out of range");
        KV_REQUIRE(nblocks >= 0 && (nblocks == 0 || (qn && shape)) && dirs, kv::ERR_ARG,
                   "null structure arrays");
        kv::Layout l;
        l.rank = rank;
        for (int r = 0; r < rank; ++r) {
            KV_REQUIRE(dirs[r] == 1 || dirs[r] == -1, kv::ERR_ARG, "dirs must be +1 or -1");
            l.dirs[r] = dirs[r];
        }
        for (int b = 0; b < nblocks; ++b) {
            kv::BlockDesc bd;
            for (int r = 0; r < rank; ++r) {
                bd.qn[r] = qn[b * rank + r];
                bd.shape[r] = shape[b * rank + r];
            }
            l.blocks.push_back(bd);
        }
        l.finalize();
        KV_REQUIRE(l.check_sym(), kv::ERR_SYM, "block breaks charge conservation");
        *out = wrap(ctx, kv::Tensor::create(ctx->dev, l, true));
        ctx->dev->sync();
    });
}

int kv_tensor_destroy(kv_tensor* t) {
    if (!t) return KV_ERR_ARG;
    delete t;
    return KV_OK;
}

int kv_tensor_rank(const kv_tensor* t, int* rank, int* nblocks, int64_t* nelem) {
    if (!t || !t->t) return KV_ERR_ARG;
    if (rank) *rank = t->t->lay.rank;
    if (nblocks) *nblocks = (int)t->t->lay.blocks.size();
    if (nelem) *nelem = t->t->lay.total;
    return KV_OK;
}

int kv_tensor_dirs(const kv_tensor* t, int* dirs) {
    if (!t || !t->t || !dirs) return KV_ERR_ARG;
    for (int r = 0; r < t->t->lay.rank; ++r) dirs[r] = t->t->lay.dirs[r];
    return KV_OK;
}

int kv_tensor_blocks(const kv_tensor* t, int* qn, int* shape, int64_t* offset) {
    if (!t || !t->t) return KV_ERR_ARG;
    const kv::Layout& l = t->t->lay;
    for (size_t b = 0; b < l.blocks.size(); ++b) {
        for (int r = 0; r < l.rank; ++r) {
            if (qn) qn[b * l.rank + r] = l.blocks[b].qn[r];
            if (shape) shape[b * l.rank + r] = l.blocks[b].shape[r];
        }
        if (offset) offset[b] = l.blocks[b].off;
    }
    return KV_OK;
}

int kv_tensor_upload(kv_tensor* t, const double* data) {
    if (!t || !t->t || (!data && t->t->lay.total)) return KV_ERR_ARG;
    return guard(t->ctx, [&] { t->t->upload(data); });
}

int kv_tensor_download(const kv_tensor* t, double* data) {
    if (!t || !t->t || (!data && t->t->lay.total)) return KV_ERR_ARG;
    return guard(t->ctx, [&] { t->t->download(data); });
}

int kv_tensor_norm(const kv_tensor* t, double* out) {
    if (!t || !t->t || !out) return KV_ERR_ARG;
    return guard(t->ctx, [&] {
        kv::Device& D = *t->ctx->dev;
        kv::DevBuf s(D.pool.get(), 8), scr(D.pool.get(), 2 * 1024 * sizeof(double));
        kv::launch_nrm2sq(D.stream, t->t->d(), t->t->lay.total, s.as<double>(), scr.as<double>());
        double h = 0;
        KV_HIP(hipMemcpyAsync(&h, s.get(), 8, hipMemcpyDeviceToHost, D.stream));
        D.sync();
        *out = std::sqrt(h);
    });
}

int kv_tensor_overlap(const kv_tensor* a, const kv_tensor* b, double* re_im) {
    if (!a || !b || !re_im) return KV_ERR_ARG;
    return guard(a->ctx, [&] {
        kv::Tensor* A = need(a, "a");
        kv::Tensor* B = need(b, "b");
        kv::Device& D = *a->ctx->dev;
        // <b|a>: one reduction over every pair of blocks of equal qn and shape
        std::vector<kv::DotPair> pairs;
        int64_t total = 0;
        for_overlap_pairs(*A, *B, [&](const kv::DotPair& p) {
            pairs.push_back(p);
            total += p.size;
        });
        kv::DevBuf dp = kv::upload_vec(D, pairs);
        kv::DevBuf part(D.pool.get(), 2 * kv::kRedBlocks * sizeof(double)), out(D.pool.get(), 16);
        kv::launch_dotc_pairs(D.stream, B->d(), A->d(), dp.as<kv::DotPair>(), (int)pairs.size(), total,
                              part.as<double>(), out.as<double2>());
        KV_HIP(hipMemcpyAsync(re_im, out.get(), 16, hipMemcpyDeviceToHost, D.stream));
        D.sync();
    });
}

int kv_tensor_overlap_noalloc(const kv_tensor* a, const kv_tensor* b, double* re_im) {
    if (!a || !b || !re_im) return KV_ERR_ARG;
    return guard(a->ctx, [&] {
        kv::Tensor* A = need(a, "a");
        kv::Tensor* B = need(b, "b");
        kv::Device& D = *a->ctx->dev;
        // the pair list is staged on the stack and copied into the device's static scratch
        // (pairs, then the partial sums, then the result); nothing is allocated
        kv::DotPair pairs[kv::kNoAllocMaxPairs];
        int np = 0;
        int64_t total = 0;
        bool over = false;
        for_overlap_pairs(*A, *B, [&](const kv::DotPair& p) {
            if (np < kv::kNoAllocMaxPairs)
                pairs[np] = p;
            else
                over = true;
            ++np;
            total += p.size;
        });
        KV_REQUIRE(!over, KV_ERR_ARG,
                   "kv_tensor_overlap_noalloc: " + std::to_string(np) + " matching blocks exceed the " +
                       std::to_string(kv::kNoAllocMaxPairs) + " its scratch holds");
        char* base = static_cast<char*>(kv::noalloc_scratch(D.id));
        auto* dpairs = reinterpret_cast<kv::DotPair*>(base);
        auto* part = reinterpret_cast<double*>(base + kv::kNoAllocPairBytes);
        auto* out = reinterpret_cast<double2*>(base + kv::kNoAllocPairBytes + kv::kNoAllocPartBytes);
        if (np > 0)
            KV_HIP(hipMemcpyAsync(dpairs, pairs, (size_t)np * sizeof(kv::DotPair), hipMemcpyHostToDevice,
                                  D.stream));
        kv::launch_dotc_pairs(D.stream, B->d(), A->d(), dpairs, np, total, part, out);
        KV_HIP(hipMemcpyAsync(re_im, out, 16, hipMemcpyDeviceToHost, D.stream));
        D.sync();
    });
}

int kv_product(kv_ctx* ctx, const kv_tensor* m1, const kv_tensor* m2, int nc, char op1,
               char op2, int precision, kv_tensor** out) {
    if (!ctx || !out) return KV_ERR_ARG;
    return guard(ctx, [&] {
        KV_REQUIRE(precision == KV_FP64, kv::ERR_ARG, "kv_product runs in FP64");
        auto C = kv::product(ctx->dev, *need(m1, "m1"), *need(m2, "m2"), nc, op1, op2);
        ctx->dev->sync();
        *out = wrap(ctx, std::move(C));
    });
}

int kv_tensor_transpose(kv_ctx* ctx, const kv_tensor* t, const int* axes, kv_tensor** out) {
    if (!ctx || !axes || !out) return KV_ERR_ARG;
    return guard(ctx, [&] {
        const kv::Tensor* T = need(t, "t");
        kv::Idx ax{};
        for (int a = 0; a < T->lay.rank; ++a) ax[a] = axes[a];
        *out = wrap(ctx, kv::transpose(ctx->dev, *T, ax));
    });
}

int kv_tensor_axpy(kv_ctx* ctx, const kv_tensor* x, const kv_tensor* y, double alpha_re,
                   double alpha_im, kv_tensor** out) {
    if (!ctx || !out) return KV_ERR_ARG;
    return guard(ctx, [&] {
        *out = wrap(ctx, kv::axpy(ctx->dev, *need(x, "x"), *need(y, "y"),
                                  make_double2(alpha_re, alpha_im)));
    });
}

int kv_tensor_scale(kv_ctx* ctx, kv_tensor* t, double alpha_re, double alpha_im) {
    if (!ctx || !t) return KV_ERR_ARG;
    return guard(ctx, [&] {
        kv::Tensor* T = t->t.get();
        KV_REQUIRE(T, kv::ERR_ARG, "t is null");
        kv::launch_zaxpby(ctx->dev->stream, T->d(), T->d(), nullptr,
                          make_double2(alpha_re, alpha_im), make_double2(0.0, 0.0), T->lay.total);
        ctx->dev->sync();
    });
}

int kv_heff_apply(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                  const kv_tensor* W2, const kv_tensor* R, const kv_tensor* x, int precision,
                  kv_tensor** y) {
    if (!ctx || !y) return KV_ERR_ARG;
    return guard(ctx, [&] {
        kv::Tensor* w1 = nsite >= 1 ? need(W1, "W1") : nullptr;
        kv::Tensor* w2 = nsite == 2 ? need(W2, "W2") : nullptr;
        kv::Heff H;
        H.build(ctx->dev, nsite, *need(L, "L"), w1, w2, *need(R, "R"), need(x, "x")->lay,
                prec_of(precision));
        auto xin = kv::relayout(ctx->dev, *x->t, H.layout());
        auto out = kv::Tensor::create(ctx->dev, H.layout(), false);
        H.apply(xin->d(), out->d(), 1.0);
        ctx->dev->sync();
        *y = wrap(ctx, std::move(out));
    });
}

int kv_heff_plan_create(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                        const kv_tensor* W2, const kv_tensor* R, const kv_tensor* x, int precision,
                        kv_heff_plan** out) {
    if (!ctx || !out) return KV_ERR_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        KV_REQUIRE(nsite >= 0 && nsite <= 2, kv::ERR_ARG, "nsite must be 0, 1 or 2");
        kv::Tensor* w1 = nsite >= 1 ? need(W1, "W1") : nullptr;
        kv::Tensor* w2 = nsite == 2 ? need(W2, "W2") : nullptr;
        std::unique_ptr<kv_heff_plan> p(new kv_heff_plan);
        p->ctx = ctx;
        p->H.build(ctx->dev, nsite, *need(L, "L"), w1, w2, *need(R, "R"), need(x, "x")->lay,
                   prec_of(precision));
        ctx->dev->sync();
        *out = p.release();
    });
}

int kv_heff_plan_apply(kv_heff_plan* plan, const kv_tensor* x, double scale, kv_tensor** y) {
    if (!plan || !y) return KV_ERR_ARG;
    kv_ctx* ctx = plan->ctx;
    return guard(ctx, [&] {
        const kv::Tensor* xt = need(x, "x");
        KV_REQUIRE(x->ctx == ctx, kv::ERR_ARG, "x belongs to another context");
        std::unique_ptr<kv::Tensor> spread;
        if (!kv::same_layout(xt->lay, plan->H.layout())) {
            spread = kv::relayout(ctx->dev, *xt, plan->H.layout());
            xt = spread.get();
        }
        auto out = kv::Tensor::create(ctx->dev, plan->H.layout(), false);
        plan->H.apply(xt->d(), out->d(), scale);
        ctx->dev->sync();
        *y = wrap(ctx, std::move(out));
    });
}

int kv_heff_plan_destroy(kv_heff_plan* plan) {
    if (!plan) return KV_ERR_ARG;
    try {
        plan->ctx->dev->sync();
        delete plan;
        return KV_OK;
    } catch (...) {
        return KV_ERR_INTERNAL;
    }
}

int kv_heff_profile(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                    const kv_tensor* W2, const kv_tensor* R, const kv_tensor* x, int precision,
                    int reps, double* out) {
    if (!ctx || !out || reps < 1) return KV_ERR_ARG;
    return guard(ctx, [&] {
        kv::Tensor* w1 = nsite >= 1 ? need(W1, "W1") : nullptr;
        kv::Tensor* w2 = nsite == 2 ? need(W2, "W2") : nullptr;
        kv::Heff H;
        H.build(ctx->dev, nsite, *need(L, "L"), w1, w2, *need(R, "R"), need(x, "x")->lay,
                prec_of(precision));
        auto xin = kv::relayout(ctx->dev, *x->t, H.layout());
        auto y = kv::Tensor::create(ctx->dev, H.layout(), false);
        H.apply(xin->d(), y->d(), 1.0);  // first call: GEMM tile autotuning, not timed
        ctx->dev->sync();
        kv::EventLog log;
        kv::KernelStats st;
        kv::EventTimer ev;
        ev.start(ctx->dev->stream);
        for (int r = 0; r < reps; ++r) H.apply(xin->d(), y->d(), 1.0, &log);
        ev.stop(ctx->dev->stream);
        const double wall_ms = ev.ms();
        ctx->dev->sync();
        log.resolve(st);
        out[0] = 1e-3 * wall_ms / reps;
        out[1] = H.flop();
        out[2] = 1e-3 * st.gemm_ms / reps;
        out[3] = st.gemm_flop / reps;
        out[4] = (double)st.gemm_launches / reps;
        out[5] = 1e-3 * st.wfold_ms / reps;
        out[6] = st.wfold_flop / reps;
        out[7] = (double)st.wfold_launches / reps;
    });
}

int kv_env_update(kv_ctx* ctx, const kv_tensor* env, const kv_tensor* bra, const kv_tensor* W,
                  const kv_tensor* ket, int left_to_right, int precision, kv_tensor** out) {
    if (!ctx || !out) return KV_ERR_ARG;
    return guard(ctx, [&] {
        auto o = kv::env_update(ctx->dev, *need(env, "env"), bra ? need(bra, "bra") : nullptr,
                                *need(W, "W"), *need(ket, "ket"), left_to_right != 0,
                                prec_of(precision));
        *out = wrap(ctx, std::move(o));
    });
}

int kv_krylov_step(kv_ctx* ctx, int nsite, const kv_tensor* L, const kv_tensor* W1,
                   const kv_tensor* W2, const kv_tensor* R, kv_tensor* x, double dt, int reverse,
                   double tol, int max_iter, int precision, kv_krylov_exit* exit_info) {
    if (!ctx || !x) return KV_ERR_ARG;
    return guard(ctx, [&] {
        KV_REQUIRE(max_iter >= 2 && max_iter <= 64, kv::ERR_ARG, "max_iter out of range");
        kv::Tensor* w1 = nsite >= 1 ? need(W1, "W1") : nullptr;
        kv::Tensor* w2 = nsite == 2 ? need(W2, "W2") : nullptr;
        need(x, "x");
        kv::KrylovExit ex = kv::local_krylov(ctx->dev, nsite, *need(L, "L"), w1, w2,
                                             *need(R, "R"), x->t, dt, tol, max_iter,
                                             reverse != 0, prec_of(precision));
        if (exit_info) {
            exit_info->converged = ex.converged ? 1 : 0;
            exit_info->iterations = ex.iterations;
            exit_info->error = ex.error;
            exit_info->loss_norm = ex.loss_norm;
        }
    });
}

int kv_svd_truncate(kv_ctx* ctx, const kv_tensor* A, int r1, double cutoff, int maxdim,
                    int left_to_right, kv_tensor** left, kv_tensor** right, int* nkept) {
    if (!ctx || !left || !right) return KV_ERR_ARG;
    return guard(ctx, [&] {
        KV_REQUIRE(maxdim >= 1, kv::ERR_ARG, "maxdim must be positive");
        kv::SvdResult r = kv::svd_truncate(ctx->dev, *ctx->sol, *need(A, "A"), r1, cutoff,
                                           maxdim, left_to_right != 0);
        if (nkept) *nkept = r.kept;
        *left = wrap(ctx, std::move(r.left));
        *right = wrap(ctx, std::move(r.right));
    });
}

int kv_tdvp_create(kv_ctx* ctx, int nsites, kv_tensor* const* mps, kv_tensor* const* mpo,
                   const kv_tdvp_config* cfg, kv_tdvp** out) {
    if (!ctx || !out || !cfg || !mps || !mpo) return KV_ERR_ARG;
    return guard(ctx, [&] {
        std::vector<std::unique_ptr<kv::Tensor>> a, w;
        for (int i = 0; i < nsites; ++i) {
            a.push_back(need(mps[i], "mps")->clone());
            w.push_back(need(mpo[i], "mpo")->clone());
        }
        kv::TdvpConfig c;
        c.dt = cfg->dt;
        c.cutoff = cfg->cutoff;
        c.maxdim = cfg->maxdim;
        c.krylov_tol = cfg->krylov_tol;
        c.krylov_maxiter = cfg->krylov_maxiter;
        c.prec = prec_of(cfg->precision);
        KV_REQUIRE(c.maxdim >= 1 && c.krylov_maxiter >= 2 && c.krylov_maxiter <= 64, kv::ERR_ARG,
                   "bad TDVP configuration");
        kv_tdvp* h = new kv_tdvp;
        h->ctx = ctx;
        try {
            h->s.reset(new kv::TDVP(ctx->dev, std::move(a), std::move(w), c));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int kv_tdvp_destroy(kv_tdvp* s) {
    if (!s) return KV_ERR_ARG;
    delete s;
    return KV_OK;
}

int kv_tdvp_two_site(kv_tdvp* s) {
    if (!s || !s->s) return KV_ERR_ARG;
    return guard(s->ctx, [&] { s->s->two_site(); });
}

int kv_tdvp_one_site(kv_tdvp* s) {
    if (!s || !s->s) return KV_ERR_ARG;
    return guard(s->ctx, [&] { s->s->one_site(); });
}

int kv_tdvp_get_site(kv_tdvp* s, int i, kv_tensor** out) {
    if (!s || !s->s || !out) return KV_ERR_ARG;
    return guard(s->ctx, [&] {
        KV_REQUIRE(i >= 0 && i < s->s->nsites(), kv::ERR_ARG, "site index out of range");
        auto t = s->s->site(i).clone();
        s->ctx->dev->sync();
        *out = wrap(s->ctx, std::move(t));
    });
}

int kv_tdvp_sa_one_site(kv_tdvp* s, const double* betas, int nbetas, uint64_t seed) {
    if (!s || !s->s || !betas || nbetas < 0) return KV_ERR_ARG;
    return guard(s->ctx, [&] {
        KV_REQUIRE(nbetas >= s->s->nsites() - 1, kv::ERR_ARG, "SA-1TDVP needs a fitted width for every bond");
        std::vector<double> b(betas, betas + nbetas);
        for (double x : b) KV_REQUIRE(std::isfinite(x) && x > 0.0, kv::ERR_ARG, "SA-1TDVP: bad inverse width");
        s->s->one_site(&b, seed);
    });
}

int kv_tdvp_svs_open(kv_tdvp* s, const char* path) {
    if (!s || !s->s || !path) return KV_ERR_ARG;
    return guard(s->ctx, [&] {
        FILE* f = std::fopen(path, "w");
        KV_REQUIRE(f != nullptr, kv::ERR_ARG, std::string("cannot write the SVS log ") + path);
        std::fclose(f);
        s->s->open_svs(path);
    });
}

int kv_sa_fit_widths(const char* svs_path, double* widths, int max_bonds, int* nbonds) {
    if (!svs_path || !nbonds || max_bonds < 0 || (max_bonds > 0 && !widths)) return KV_ERR_ARG;
    FILE* f = std::fopen(svs_path, "rb");
    if (!f) return KV_ERR_ARG;
    std::string text;
    char buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    std::vector<double> w;
    if (!kv::sa_fit_widths(text, w)) return KV_ERR_FORMAT;
    *nbonds = (int)w.size();
    for (int i = 0; i < (int)w.size() && i < max_bonds; ++i) widths[i] = w[i];
    return KV_OK;
}

int kv_tdvp_global_krylov(kv_tdvp* s, int* iterations) {
    if (!s || !s->s) return KV_ERR_ARG;
    return guard(s->ctx, [&] {
        kv::GlobalKrylovExit ex = s->s->global_krylov();
        if (iterations) *iterations = ex.iterations;
    });
}

int kv_tdvp_bond_dims(kv_tdvp* s, int* dims) {
    if (!s || !s->s || !dims) return KV_ERR_ARG;
    auto d = s->s->bond_dims();
    for (size_t i = 0; i < d.size(); ++i) dims[i] = d[i];
    return KV_OK;
}

int kv_tdvp_stats(kv_tdvp* s, double* out, int n) {
    if (!s || !s->s || !out) return KV_ERR_ARG;
    const kv::TdvpTimers& t = s->s->timers();
    const kv::KernelStats& h = s->s->heff_kernels();
    const kv::KernelStats& e = s->s->env_kernels();
    const kv::TrdStats& r = s->s->trd_stats();
    const kv::BlockStats& bk = s->s->device().blk;
    const double v[36] = {t.evl2, t.evl1, t.evl0, t.svd, t.update, t.sweep,
                          (h.gemm_ms + h.wfold_ms) * 1e-3, h.gemm_flop + h.wfold_flop,
                          (e.gemm_ms + e.wfold_ms) * 1e-3, e.gemm_flop + e.wfold_flop,
                          (double)s->s->lanczos_stats().matvecs,
                          h.gemm_ms * 1e-3, h.gemm_flop, (double)h.gemm_launches,
                          h.wfold_ms * 1e-3, h.wfold_flop, (double)h.wfold_launches,
                          h.gemm_bytes, h.gemm32_ms * 1e-3, h.gemm32_flop,
                          (double)h.gemm32_launches, h.gemm32_bytes,
                          r.ms * 1e-3, r.bytes, (double)r.columns,
                          bk.gemm_ms * 1e-3, bk.gemm_flop, (double)bk.gemm_launches, bk.gemm_bytes,
                          bk.orth_ms * 1e-3, (double)bk.orth_calls, bk.eig_ms * 1e-3,
                          (double)bk.eig_calls, (double)bk.sectors, (double)bk.fallbacks,
                          (double)s->s->sa_added()};
    for (int i = 0; i < n && i < 36; ++i) out[i] = v[i];
    return KV_OK;
}

int kv_tdvp_reset_stats(kv_tdvp* s) {
    if (!s || !s->s) return KV_ERR_ARG;
    s->s->reset_stats();
    return KV_OK;
}

int kv_join(kv_ctx* ctx, int nsites, kv_tensor* const* mpo, kv_tensor* const* bra,
            kv_tensor* const* ket, double* re_im) {
    if (!ctx || !mpo || !bra || !ket || !re_im) return KV_ERR_ARG;
    return guard(ctx, [&] {
        std::vector<kv::Tensor*> w;
        std::vector<const kv::Tensor*> b, k;
        for (int i = 0; i < nsites; ++i) {
            w.push_back(need(mpo[i], "mpo"));
            b.push_back(need(bra[i], "bra"));
            k.push_back(need(ket[i], "ket"));
        }
        auto v = kv::join(ctx->dev, w, b, k);
        re_im[0] = v.real();
        re_im[1] = v.imag();
    });
}

int kv_overlap_states(kv_ctx* ctx, int nsites, kv_tensor* const* s1, kv_tensor* const* s2,
                      double* re_im) {
    if (!ctx || !s1 || !s2 || !re_im) return KV_ERR_ARG;
    return guard(ctx, [&] {
        std::vector<const kv::Tensor*> a, b;
        for (int i = 0; i < nsites; ++i) {
            a.push_back(need(s1[i], "s1"));
            b.push_back(need(s2[i], "s2"));
        }
        auto v = kv::overlap_states(ctx->dev, a, b);
        re_im[0] = v.real();
        re_im[1] = v.imag();
    });
}

static int load_common(kv_ctx* ctx, const char* path, int max_sites, int* nsites,
                       kv_tensor** sites, int rank, int magic) {
    if (!ctx || !path || !nsites) return KV_ERR_ARG;
    return guard(ctx, [&] {
        auto v = kv::load_train(ctx->dev, path, rank, magic);
        *nsites = (int)v.size();
        KV_REQUIRE(sites == nullptr || (int)v.size() <= max_sites, kv::ERR_ARG,
                   "site array too small");
        if (sites)
            for (size_t i = 0; i < v.size(); ++i) sites[i] = wrap(ctx, std::move(v[i]));
    });
}

int kv_mps_load(kv_ctx* ctx, const char* path, int max_sites, int* nsites, kv_tensor** sites) {
    return load_common(ctx, path, max_sites, nsites, sites, 3, 111);
}

int kv_mpo_load(kv_ctx* ctx, const char* path, int max_sites, int* nsites, kv_tensor** sites) {
    return load_common(ctx, path, max_sites, nsites, sites, 4, 112);
}

static int save_common(kv_ctx* ctx, const char* path, int nsites, kv_tensor* const* sites,
                       int rank, int magic) {
    if (!ctx || !path || !sites) return KV_ERR_ARG;
    return guard(ctx, [&] {
        std::vector<kv::Tensor*> v;
        for (int i = 0; i < nsites; ++i) v.push_back(need(sites[i], "site"));
        kv::save_train(path, v, rank, magic);
    });
}

int kv_mps_save(kv_ctx* ctx, const char* path, int nsites, kv_tensor* const* sites) {
    return save_common(ctx, path, nsites, sites, 3, 111);
}

int kv_mpo_save(kv_ctx* ctx, const char* path, int nsites, kv_tensor* const* sites) {
    return save_common(ctx, path, nsites, sites, 4, 112);
}

}  // extern "C"
