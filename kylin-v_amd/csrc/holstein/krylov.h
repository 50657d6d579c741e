// Global Krylov time step on the device: exp(-i dt H) applied to the whole MPS by a
// Lanczos recurrence whose vectors are MPS (holstein/krylov.h:206-284), the stage the
// reference runs before two-site TDVP to open charge sectors a local update cannot
// (holstein/tdvp.h:318-327, with H scaled by dt on its first site).
//
// Building blocks, each on device tensors:
//   stack / state_add / state_scale / canonicalize   holstein/tensor.h:262-319,
//                                                     holstein/state.h:52-80,166-185
//   apply_operator (zip-up MPO x MPS)                 holstein/operator.h:231-272
//   apply_ortho (first and later Krylov vectors)      holstein/krylov.h:126-205
//   apply_ortho_site                                  holstein/krylov.h:12-45
#pragma once

#include <complex>
#include <array>
#include <map>
#include <memory>
#include <set>
#include <vector>

#include "tdvp.h"

namespace kv {

using zc = std::complex<double>;
using State = std::vector<std::unique_ptr<Tensor>>;

inline double tensor_norm(Device& D, const Tensor& t) {
    DevBuf s(D.pool.get(), sizeof(double)), scr(D.pool.get(), 1024 * sizeof(double));
    launch_nrm2sq(D.stream, t.d(), t.lay.total, s.as<double>(), scr.as<double>());
    double h = 0;
    KV_HIP(hipMemcpyAsync(&h, s.get(), sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    return std::sqrt(h);
}

// overlap(ori, oth) = sum over blocks of equal charges of conj(oth) . ori
// (holstein/tensor.h:225-241).
inline zc tensor_overlap(Device& D, const Tensor& ori, const Tensor& oth) {
    DevBuf s(D.pool.get(), sizeof(double2)), scr(D.pool.get(), 2 * 1024 * sizeof(double));
    zc acc(0, 0);
    for (const auto& ba : ori.lay.blocks) {
        const int j = oth.lay.find(ba.qn);
        if (j < 0 || !qn_equal(oth.lay.blocks[j].shape, ba.shape, ori.lay.rank)) continue;
        launch_dotc(D.stream, oth.d() + oth.lay.blocks[j].off, ori.d() + ba.off, ba.size,
                    s.as<double2>(), scr.as<double>());
        double h[2];
        KV_HIP(hipMemcpyAsync(h, s.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, D.stream));
        D.sync();
        acc += zc(h[0], h[1]);
    }
    return acc;
}

inline std::unique_ptr<Tensor> scaled(const std::shared_ptr<Device>& dev, const Tensor& t, zc c) {
    auto out = Tensor::create(dev, t.lay, false);
    launch_zaxpby(dev->stream, out->d(), t.d(), nullptr, make_double2(c.real(), c.imag()),
                  make_double2(0.0, 0.0), t.lay.total);
    dev->sync();
    return out;
}

// Direct sum of two rank-3 tensors on legs 0 and/or 2 (mask bit a = leg a stacked):
// sector q of a stacked leg gets dimA(q) + dimB(q); A's blocks sit at offset 0, B's after
// A's part (tensor.h:262-319).  Leg 1 must agree.
inline std::unique_ptr<Tensor> stack(const std::shared_ptr<Device>& dev, const Tensor& A,
                                     const Tensor& B, unsigned mask) {
    KV_REQUIRE(A.lay.rank == 3 && B.lay.rank == 3 && !(mask & 2u), ERR_ARG,
               "stack: rank-3 tensors, legs 0 and 2 only");
    std::map<int, int> dA[3], dB[3];
    for (const auto& b : A.lay.blocks)
        for (int a = 0; a < 3; ++a) dA[a][b.qn[a]] = b.shape[a];
    for (const auto& b : B.lay.blocks)
        for (int a = 0; a < 3; ++a) dB[a][b.qn[a]] = b.shape[a];
    auto dim = [&](int a, int q) {
        auto ia = dA[a].find(q), ib = dB[a].find(q);
        return (ia == dA[a].end() ? 0 : ia->second) + (ib == dB[a].end() ? 0 : ib->second);
    };
    auto offB = [&](int a, int q) {
        auto ia = dA[a].find(q);
        return ia == dA[a].end() ? 0 : ia->second;
    };
    Layout L;
    L.rank = 3;
    L.dirs = A.lay.dirs;
    std::set<std::array<int, 3>> seen;
    auto add_block = [&](const BlockDesc& b) {
        Idx q{}, sh{};
        for (int a = 0; a < 3; ++a) {
            q[a] = b.qn[a];
            sh[a] = (mask & (1u << a)) ? dim(a, b.qn[a]) : b.shape[a];
        }
        if (!seen.insert({q[0], q[1], q[2]}).second) return;
        BlockDesc nb;
        nb.qn = q;
        nb.shape = sh;
        L.blocks.push_back(nb);
    };
    for (const auto& b : A.lay.blocks) add_block(b);
    for (const auto& b : B.lay.blocks) add_block(b);
    L.finalize();
    auto out = Tensor::create(dev, L, true);
    std::vector<Copy2D> cps;
    int maxrows = 1;
    auto place = [&](const Tensor& X, bool second) {
        for (const auto& b : X.lay.blocks) {
            const BlockDesc& o = L.blocks[L.find(b.qn)];
            const int oa = (second && (mask & 1u)) ? offB(0, b.qn[0]) : 0;
            const int ob = (second && (mask & 4u)) ? offB(2, b.qn[2]) : 0;
            Copy2D c{};
            c.src_off = b.off;
            c.dst_off = o.off + (int64_t)oa * o.shape[1] * o.shape[2] + ob;
            c.rows = b.shape[0] * b.shape[1];
            c.cols = b.shape[2];
            c.src_ld = b.shape[2];
            c.dst_ld = o.shape[2];
            maxrows = std::max(maxrows, c.rows);
            cps.push_back(c);
        }
    };
    place(A, false);
    const size_t na = cps.size();
    place(B, true);
    if (!cps.empty()) {
        // one launch per source tensor (the copy kernel reads a single source)
        DevBuf d = upload_vec(*dev, cps);
        if (na)
            launch_copy2d<double2>(dev->stream, A.d(), out->d(), d.as<Copy2D>(), (int)na, maxrows,
                                   nullptr);
        if (cps.size() > na)
            launch_copy2d<double2>(dev->stream, B.d(), out->d(), d.as<Copy2D>() + na,
                                   (int)(cps.size() - na), maxrows, nullptr);
        dev->sync();
    }
    return out;
}

inline State state_copy(const State& s) {
    State o;
    for (const auto& t : s) o.push_back(t->clone());
    return o;
}

// State += (state.h:52-63): site 0 on its right bond, middle sites on both, last on its left.
inline State state_add(const std::shared_ptr<Device>& dev, const State& a, const State& b) {
    const size_t ns = a.size();
    State o;
    o.push_back(stack(dev, *a[0], *b[0], 4u));
    for (size_t i = 1; i + 1 < ns; ++i) o.push_back(stack(dev, *a[i], *b[i], 5u));
    o.push_back(stack(dev, *a[ns - 1], *b[ns - 1], 1u));
    return o;
}

// State * c scales the first site (state.h:66-80).
inline State state_scale(const std::shared_ptr<Device>& dev, const State& s, zc c) {
    State o;
    o.push_back(scaled(dev, *s[0], c));
    for (size_t i = 1; i < s.size(); ++i) o.push_back(s[i]->clone());
    return o;
}

// state.h:166-185
inline void canonicalize(const std::shared_ptr<Device>& dev, Solver& sol, State& s, double tol,
                         int maxdim) {
    const int ns = (int)s.size();
    for (int i = 0; i < ns - 1; ++i) {
        SvdResult r = svd_truncate(dev, sol, *s[i], 2, tol, maxdim, true);
        s[i] = std::move(r.left);
        s[i + 1] = product(dev, *r.right, *s[i + 1], 1, 'N', 'N');
    }
    for (int i = ns - 1; i > 0; --i) {
        SvdResult r = svd_truncate(dev, sol, *s[i], 1, tol, maxdim, false);
        s[i] = std::move(r.right);
        s[i - 1] = product(dev, *s[i - 1], *r.left, 1, 'N', 'N');
    }
}

// lmo[a', s', wR, b] = sum W[s', wL, s, wR] L[a', wL, a] A[a, s, b] (env_mps_mpo,
// holstein/product.h:203-276): the first two stages of the environment update.
inline std::unique_ptr<Tensor> env_mps_mpo(const std::shared_ptr<Device>& dev, const Tensor& L,
                                           const Tensor& A, Tensor& W) {
    DevProduct s1;
    s1.build(*dev, plan_product(L.lay, A.lay, 1, 'N', 'N'));
    DevWFold s2;
    s2.build(*dev, plan_wfold(s1.p.out, 1, W.lay, W.host_data().data()));
    DevBuf t1(dev->pool.get(), (size_t)s1.p.out.total * sizeof(double2));
    auto out = Tensor::create(dev, s2.p.out, false);
    s1.run<double2>(*dev, L.d(), A.d(), t1.as<double2>());
    s2.run<double2>(*dev, t1.as<double2>(), out->d());
    dev->sync();
    return out;
}

// operator.h:231-272
inline State apply_operator(const std::shared_ptr<Device>& dev, Solver& sol,
                            const std::vector<Tensor*>& H, const State& wf, double tol,
                            int maxdim) {
    const int ns = (int)H.size();
    std::unique_ptr<Tensor> envl = left_boundary(dev);
    const int last_b = wf[ns - 1]->lay.blocks[0].qn[2];
    const int last_w = H[ns - 1]->lay.blocks[0].qn[3];
    State res(ns);
    for (int i = 0; i < ns; ++i) {
        auto lmo = env_mps_mpo(dev, *envl, *wf[i], *H[i]);
        if (i == ns - 1) {
            Layout l;
            l.rank = 3;
            l.dirs = Idx{-1, 1, 1, 0, 0, 0};
            BlockDesc b;
            b.qn = Idx{last_b, last_w, last_b, 0, 0, 0};
            b.shape = Idx{1, 1, 1, 0, 0, 0};
            l.blocks.push_back(b);
            l.finalize();
            auto envr = Tensor::create(dev, l, false);
            const double one[2] = {1.0, 0.0};
            envr->upload(one);
            res[i] = product(dev, *lmo, *envr, 2, 'N', 'T');
        } else {
            SvdResult r = svd_truncate(dev, sol, *lmo, 2, tol, maxdim, true);
            res[i] = std::move(r.left);
            envl = std::move(r.right);
        }
    }
    for (int i = ns - 1; i > 0; --i) {
        SvdResult r = svd_truncate(dev, sol, *res[i], 1, tol, maxdim, false);
        res[i] = std::move(r.right);
        res[i - 1] = product(dev, *res[i - 1], *r.left, 1, 'N', 'N');
    }
    return res;
}

// <s2|s1> with the reference's overlap(State, State) argument order (state.h:206-227).
inline zc state_overlap(const std::shared_ptr<Device>& dev, const State& s1, const State& s2) {
    std::vector<const Tensor*> a, b;
    for (auto& t : s2) a.push_back(t.get());
    for (auto& t : s1) b.push_back(t.get());
    return overlap_states(dev, a, b);
}

// krylov.h:12-45: Nh minus its La/Lb components from the 2x2 Gram system, as the
// reference writes it (its coefficients use overlap(La, Nh) = <Nh|La>, the exact
// projection for the real overlaps of a Lanczos recurrence).
inline std::unique_ptr<Tensor> apply_ortho_site(const std::shared_ptr<Device>& dev, const Tensor& Nh,
                                                const Tensor& La, const Tensor& Lb) {
    Device& D = *dev;
    const zc LNa = tensor_overlap(D, La, Nh), LNb = tensor_overlap(D, Lb, Nh);
    const double na = tensor_norm(D, La), nb = tensor_norm(D, Lb);
    const zc aa(na * na, 0), bb(nb * nb, 0);
    const zc ab = tensor_overlap(D, La, Lb), ba = std::conj(ab);
    const zc cc = ab * ba - aa * bb;
    std::unique_ptr<Tensor> res = Nh.clone();
    if (std::abs(cc) < 1e-17) {
        if (std::abs(aa) < 1e-17) return res;
        const zc beta = -1.0 * LNa / aa;
        return axpy(dev, La, *res, make_double2(beta.real(), beta.imag()));
    }
    const zc betaa = (bb * LNa - ab * LNb) / cc;
    const zc betab = (aa * LNb - ba * LNa) / cc;
    res = axpy(dev, La, *res, make_double2(betaa.real(), betaa.imag()));
    return axpy(dev, Lb, *res, make_double2(betab.real(), betab.imag()));
}

inline std::unique_ptr<Tensor> unit_env2(const std::shared_ptr<Device>& dev, bool left, int Q) {
    Layout l;
    l.rank = 2;
    l.dirs = left ? Idx{1, -1, 0, 0, 0, 0} : Idx{-1, 1, 0, 0, 0, 0};
    BlockDesc b;
    b.qn = left ? Idx{0, 0, 0, 0, 0, 0} : Idx{Q, Q, 0, 0, 0, 0};
    b.shape = Idx{1, 1, 0, 0, 0, 0};
    l.blocks.push_back(b);
    l.finalize();
    auto t = Tensor::create(dev, l, false);
    const double one[2] = {1.0, 0.0};
    t->upload(one);
    return t;
}

// two-leg overlap environments (krylov.h:66-88, 107-125)
inline std::unique_ptr<Tensor> env2_left(const std::shared_ptr<Device>& dev, const Tensor& E,
                                         const Tensor& s1, const Tensor& s2) {
    auto lm = product(dev, E, s2, 1, 'N', 'N');
    return product(dev, s1, *lm, 2, 'C', 'N');
}
inline std::unique_ptr<Tensor> env2_right(const std::shared_ptr<Device>& dev, const Tensor& E,
                                          const Tensor& s1, const Tensor& s2) {
    const Idx ax{2, 1, 0, 0, 0, 0};
    auto s1t = transpose(dev, s1, ax), s2t = transpose(dev, s2, ax);
    auto lm = product(dev, E, *s2t, 1, 'N', 'N');
    return product(dev, *s1t, *lm, 2, 'C', 'N');
}

// krylov.h:117-122
inline std::unique_ptr<Tensor> apply_double_state(const std::shared_ptr<Device>& dev,
                                                  const Tensor& envl, const Tensor& M,
                                                  const Tensor& envr) {
    auto lm = product(dev, envl, M, 1, 'N', 'N');
    return product(dev, *lm, envr, 1, 'N', 'T');
}

// apply_double (holstein/product.h:478-486) between environments whose bra and ket bonds
// differ, so input and output layouts differ too: the four stages run as planned, without
// the square operator's shared vector layout.
inline std::unique_ptr<Tensor> heff2(const std::shared_ptr<Device>& dev, const Tensor& L, Tensor* W1,
                                     Tensor* W2, const Tensor& R, const Tensor& x) {
    DevProduct s1, s4;
    DevWFold s2, s3;
    s1.build(*dev, plan_product(L.lay, x.lay, 1, 'N', 'N'));
    s2.build(*dev, plan_wfold(s1.p.out, 1, W1->lay, W1->host_data().data()));
    s3.build(*dev, plan_wfold(s2.p.out, 2, W2->lay, W2->host_data().data()));
    s4.build(*dev, plan_product(s3.p.out, R.lay, 2, 'N', 'T'));
    DevBuf t1(dev->pool.get(), (size_t)s1.p.out.total * sizeof(double2));
    DevBuf t2(dev->pool.get(), (size_t)s2.p.out.total * sizeof(double2));
    DevBuf t3(dev->pool.get(), (size_t)s3.p.out.total * sizeof(double2));
    auto y = Tensor::create(dev, s4.p.out, false);
    s1.run<double2>(*dev, L.d(), x.d(), t1.as<double2>());
    s2.run<double2>(*dev, t1.as<double2>(), t2.as<double2>());
    s3.run<double2>(*dev, t2.as<double2>(), t3.as<double2>());
    s4.run<double2>(*dev, t3.as<double2>(), R.d(), y->d());
    dev->sync();
    return y;
}

// First Krylov vector (krylov.h:126-134).
inline State apply_ortho1(const std::shared_ptr<Device>& dev, Solver& sol,
                          const std::vector<Tensor*>& H, const State& Va, double tol, int maxdim) {
    State gss = apply_operator(dev, sol, H, Va, tol, maxdim);
    const zc ovp = state_overlap(dev, gss, Va);
    State sum = state_add(dev, gss, state_scale(dev, Va, -ovp));
    canonicalize(dev, sol, sum, 0.0, 1000);
    return sum;
}

// Later Krylov vectors (krylov.h:136-204).
inline State apply_ortho2(const std::shared_ptr<Device>& dev, Solver& sol,
                          const std::vector<Tensor*>& H, const State& Va, const State& Vb,
                          double tol, int maxdim) {
    State gss = apply_operator(dev, sol, H, Va, tol, maxdim);
    const int ns = (int)H.size();
    const int Q = gss[ns - 1]->lay.blocks[0].qn[2];
    std::vector<std::unique_ptr<Tensor>> ELH, ERH, ELa, ERa, ELb, ERb;
    ELH.push_back(left_boundary(dev));
    ERH.push_back(right_boundary(dev, Q));
    ELa.push_back(unit_env2(dev, true, Q));
    ERa.push_back(unit_env2(dev, false, Q));
    ELb.push_back(unit_env2(dev, true, Q));
    ERb.push_back(unit_env2(dev, false, Q));
    for (int i = ns - 1; i > 1; --i) {
        ERH.push_back(env_update(dev, *ERH.back(), gss[i].get(), *H[i], *Va[i], false, PREC_F64));
        ERa.push_back(env2_right(dev, *ERa.back(), *gss[i], *Va[i]));
        ERb.push_back(env2_right(dev, *ERb.back(), *gss[i], *Vb[i]));
    }
    auto bond = [&](int i, bool l2r) {
        auto Ta = product(dev, *Va[i], *Va[i + 1], 1, 'N', 'N');
        auto Tb = product(dev, *Vb[i], *Vb[i + 1], 1, 'N', 'N');
        auto Nh = heff2(dev, *ELH.back(), H[i], H[i + 1], *ERH.back(), *Ta);
        auto La = apply_double_state(dev, *ELa.back(), *Ta, *ERa.back());
        auto Lb = apply_double_state(dev, *ELb.back(), *Tb, *ERb.back());
        auto Mgs = apply_ortho_site(dev, *Nh, *La, *Lb);
        SvdResult r = svd_truncate(dev, sol, *Mgs, 2, tol, maxdim, l2r);
        gss[i] = std::move(r.left);
        gss[i + 1] = std::move(r.right);
    };
    for (int i = 0; i < ns - 1; ++i) {
        bond(i, true);
        if (i < ns - 2) {
            ELH.push_back(env_update(dev, *ELH.back(), gss[i].get(), *H[i], *Va[i], true, PREC_F64));
            ERH.pop_back();
            ELa.push_back(env2_left(dev, *ELa.back(), *gss[i], *Va[i]));
            ERa.pop_back();
            ELb.push_back(env2_left(dev, *ELb.back(), *gss[i], *Vb[i]));
            ERb.pop_back();
        }
    }
    for (int i = ns - 1; i > 0; --i) {
        bond(i - 1, false);
        if (i > 1) {
            ERH.push_back(env_update(dev, *ERH.back(), gss[i].get(), *H[i], *Va[i], false, PREC_F64));
            ELH.pop_back();
            ERa.push_back(env2_right(dev, *ERa.back(), *gss[i], *Va[i]));
            ELa.pop_back();
            ERb.push_back(env2_right(dev, *ERb.back(), *gss[i], *Vb[i]));
            ELb.pop_back();
        }
    }
    return gss;
}

struct GlobalKrylovExit {
    int iterations = 0;
    double error = 0;
    double loss_norm = 0;
};

// krylov.h:206-284 (Hc = H * dt).
inline GlobalKrylovExit global_main(const std::shared_ptr<Device>& dev, Solver& sol,
                                    const std::vector<Tensor*>& Hc, State& mps, double tol,
                                    int maxdim, double krylov_tol, int max_iter) {
    Device& D = *dev;
    const double nrms = tensor_norm(D, *mps[0]);
    std::vector<State> V;
    V.push_back(state_scale(dev, mps, zc(1.0 / nrms, 0)));
    GlobalKrylovExit ex;
    int k = 0;
    double err = 1.0;
    while (k < max_iter - 1 && err > krylov_tol) {
        State hv = k > 0 ? apply_ortho2(dev, sol, Hc, V[k], V[k - 1], tol, maxdim)
                         : apply_ortho1(dev, sol, Hc, V[k], tol, maxdim);
        const double nmhv = tensor_norm(D, *hv[0]);
        err *= nmhv;
        if (err < krylov_tol || nmhv > 1.0) break;
        if (k == max_iter - 1) break;
        V.push_back(state_scale(dev, hv, zc(1.0 / nmhv, 0)));
        ++k;
    }
    const int n = (int)V.size();
    std::vector<zc> alpha(n), beta(n > 1 ? n - 1 : 0);
    auto ptrs = [](const State& s) {
        std::vector<const Tensor*> p;
        for (auto& t : s) p.push_back(t.get());
        return p;
    };
    std::vector<Tensor*> hv(Hc.begin(), Hc.end());
    for (int i = 0; i < n; ++i) {
        alpha[i] = join(dev, hv, ptrs(V[i]), ptrs(V[i]));
        if (i < n - 1) beta[i] = join(dev, hv, ptrs(V[i]), ptrs(V[i + 1]));
    }
    const std::vector<zc> c = krylov_coefficients(alpha, beta, false);
    State out = state_scale(dev, V[0], c[0]);
    for (int i = 1; i < n; ++i) {
        out = state_add(dev, out, state_scale(dev, V[i], c[i]));
        canonicalize(dev, sol, out, tol, maxdim);
    }
    ex.iterations = n;
    ex.error = err;
    ex.loss_norm = 1.0 - tensor_norm(D, *out[0]);
    mps = std::move(out);
    return ex;
}

// One global Krylov step of the whole MPS with H scaled by dt on its first site
// (holstein/tdvp.h:318-327), then fresh boundary and right environments.  The reference
// builds its right environments once before this stage and does not rebuild them, so its
// first two-site half-sweep after a Krylov stage sees environments of the pre-Krylov
// state; here they are rebuilt from the new MPS.
inline GlobalKrylovExit TDVP::global_krylov() {
    Clock c;
    const Tensor& W0 = *mpo_[0];
    std::vector<double> h = mpo_[0]->host_data();
    for (double& v : h) v *= cfg_.dt;
    auto w0 = Tensor::create(dev_, W0.lay, false);
    w0->upload(h.data());
    std::vector<Tensor*> Hc;
    Hc.push_back(w0.get());
    for (size_t i = 1; i < mpo_.size(); ++i) Hc.push_back(mpo_[i].get());
    GlobalKrylovExit ex = global_main(dev_, *sol_, Hc, mps_, cfg_.cutoff, cfg_.maxdim,
                                      cfg_.krylov_tol, cfg_.krylov_maxiter);
    rebuild_environments();
    tm_.global += c.s();
    return ex;
}

}  // namespace kv
