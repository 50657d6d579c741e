// TDVP sweep driver and observables on the device.
// Follows class TDVP (holstein/tdvp.h:34-386): boundary environments and the right
// stack (:300-316), two_site (:144-227) with its SVS log (:71-74, holstein/svd.h:390-400),
// one_site (:78-141) and sa_site (:230-294); Operator::join (holstein/operator.h:178-198)
// and overlap(State, State) (holstein/state.h:206-227).
#pragma once

#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "exponential.h"
#include "sa.h"
#include "svd.h"

namespace kv {

struct TdvpConfig {
    double dt = 0.05;
    double cutoff = 1e-8;
    int maxdim = 1000;
    double krylov_tol = 1e-8;
    int krylov_maxiter = 15;
    Prec prec = PREC_F64;
};

// Wall-clock module timers with the reference's Timer names (holstein/tdvp.h:85-214).
struct TdvpTimers {
    double evl2 = 0, evl1 = 0, evl0 = 0, svd = 0, update = 0, sweep = 0;
    double global = 0;  // global Krylov steps ("Krylov-Iter" + "Krylov-Diag")
};

struct GlobalKrylovExit;

inline std::unique_ptr<Tensor> left_boundary(const std::shared_ptr<Device>& dev) {
    Layout l;
    l.rank = 3;
    l.dirs = Idx{1, -1, -1, 0, 0, 0};
    BlockDesc b;
    b.qn = Idx{0, 0, 0, 0, 0, 0};
    b.shape = Idx{1, 1, 1, 0, 0, 0};
    l.blocks.push_back(b);
    l.finalize();
    auto t = Tensor::create(dev, l, false);
    const double one[2] = {1.0, 0.0};
    t->upload(one);
    return t;
}

inline std::unique_ptr<Tensor> right_boundary(const std::shared_ptr<Device>& dev, int Q) {
    Layout l;
    l.rank = 3;
    l.dirs = Idx{-1, 1, 1, 0, 0, 0};
    BlockDesc b;
    b.qn = Idx{Q, 0, Q, 0, 0, 0};
    b.shape = Idx{1, 1, 1, 0, 0, 0};
    l.blocks.push_back(b);
    l.finalize();
    auto t = Tensor::create(dev, l, false);
    const double one[2] = {1.0, 0.0};
    t->upload(one);
    return t;
}

class Clock {
  public:
    Clock() : t0_(std::chrono::steady_clock::now()) {}
    double s() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

  private:
    std::chrono::steady_clock::time_point t0_;
};

class TDVP {
  public:
    TDVP(std::shared_ptr<Device> dev, std::vector<std::unique_ptr<Tensor>> mps,
         std::vector<std::unique_ptr<Tensor>> mpo, const TdvpConfig& cfg)
        : dev_(std::move(dev)), sol_(new Solver(*dev_)), mps_(std::move(mps)),
          mpo_(std::move(mpo)), cfg_(cfg) {
        KV_REQUIRE(mps_.size() == mpo_.size() && mps_.size() >= 2, ERR_ARG,
                   "TDVP needs matching MPS/MPO of at least two sites");
        for (auto& w : mpo_) w->host_data();
        const char* v = std::getenv("KV_VERBOSE");
        verbose_ = v && v[0] == '1';
        rebuild_environments();
    }

    // Boundary environments and the right stack (holstein/tdvp.h:300-316).
    void rebuild_environments() {
        envL_.clear();
        envR_.clear();
        envL_.push_back(left_boundary(dev_));
        const Layout& last = mps_.back()->lay;
        KV_REQUIRE(!last.blocks.empty(), ERR_ARG, "empty MPS site");
        envR_.push_back(right_boundary(dev_, last.blocks[0].qn[2]));
        for (int i = (int)mps_.size() - 1; i > 0; --i)
            envR_.push_back(update(*envR_.back(), *mps_[i], *mpo_[i], false));
    }

    // One global Krylov step (defined in krylov.h).
    GlobalKrylovExit global_krylov();

    // One time step (holstein/tdvp.h:144-227).
    void two_site() {
        const int ns = (int)mps_.size();
        envR_.pop_back();
        for (int i = 0; i < ns - 1; ++i) {
            Clock c;
            auto mm = product(dev_, *mps_[i], *mps_[i + 1], 1, 'N', 'N');
            krylov(2, mm, mpo_[i].get(), mpo_[i + 1].get(), false);
            const double te = c.s();
            tm_.evl2 += te;
            Clock c2;
            SvdResult r = bond_svd(*mm, i, true);
            log_svs(r);
            const double ts = c2.s();
            tm_.svd += ts;
            if (verbose_)
                std::fprintf(stderr, "[kv] sweep L2R bond %d-%d: evl2 %.4f s (%d it) svd %.4f s kept %d/%d%s\n",
                             i, i + 1, te, last_iters_, ts, r.kept, r.total, r.detail.c_str());
            mps_[i] = std::move(r.left);
            if (i < ns - 2) {
                Clock c3;
                envL_.push_back(update(*envL_.back(), *mps_[i], *mpo_[i], true));
                tm_.update += c3.s();
                Clock c4;
                krylov(1, r.right, mpo_[i + 1].get(), nullptr, true);
                tm_.evl1 += c4.s();
                mps_[i + 1] = std::move(r.right);
                envR_.pop_back();
            } else {
                mps_[i + 1] = std::move(r.right);
            }
        }
        for (int i = ns - 1; i > 0; --i) {
            Clock c;
            auto mm = product(dev_, *mps_[i - 1], *mps_[i], 1, 'N', 'N');
            krylov(2, mm, mpo_[i - 1].get(), mpo_[i].get(), false);
            const double te = c.s();
            tm_.evl2 += te;
            Clock c2;
            SvdResult r = bond_svd(*mm, i - 1, false);
            log_svs(r);
            const double ts = c2.s();
            tm_.svd += ts;
            if (verbose_)
                std::fprintf(stderr, "[kv] sweep R2L bond %d-%d: evl2 %.4f s (%d it) svd %.4f s kept %d/%d%s\n",
                             i - 1, i, te, last_iters_, ts, r.kept, r.total, r.detail.c_str());
            mps_[i] = std::move(r.right);
            if (i > 1) {
                Clock c3;
                envR_.push_back(update(*envR_.back(), *mps_[i], *mpo_[i], false));
                tm_.update += c3.s();
                Clock c4;
                krylov(1, r.left, mpo_[i - 1].get(), nullptr, true);
                tm_.evl1 += c4.s();
                mps_[i - 1] = std::move(r.left);
                envL_.pop_back();
            } else {
                mps_[i - 1] = std::move(r.left);
            }
        }
        Clock c5;
        envR_.push_back(update(*envR_.back(), *mps_[1], *mpo_[1], false));
        tm_.update += c5.s();
    }

    // One time step of one-site TDVP (holstein/tdvp.h:78-141).  With betas (one per bond:
    // the inverse fitted widths) it is the SA-1TDVP step (holstein/tdvp.h:230-294): every
    // SVD is the enrichment SVD of bond i (sweeping right) or i-1 (sweeping left), the
    // sweep's SVDs numbered 0, 1, ... as the `call` of their draws under `seed`.
    void one_site(const std::vector<double>* betas = nullptr, uint64_t seed = 0) {
        const int ns = (int)mps_.size();
        KV_REQUIRE(!betas || (int)betas->size() >= ns - 1, ERR_ARG,
                   "SA-1TDVP needs a fitted width for every bond");
        uint64_t call = 0;
        auto split = [&](const Tensor& t, int r1, bool l2r, int bond) -> SvdResult {
            if (!betas) return svd_truncate(dev_, *sol_, t, r1, cfg_.cutoff, cfg_.maxdim, l2r);
            SaSvdResult e = svd_enrich(dev_, *sol_, t, r1, l2r, cfg_.cutoff, cfg_.maxdim, (*betas)[bond], seed, call++);
            sa_added_ += e.added;
            return std::move(e.svd);
        };
        for (int i = 0; i < ns; ++i) {
            Clock c;
            krylov(1, mps_[i], mpo_[i].get(), nullptr, false);
            tm_.evl1 += c.s();
            if (i < ns - 1) {
                Clock c2;
                SvdResult r = split(*mps_[i], 2, true, i);
                tm_.svd += c2.s();
                mps_[i] = std::move(r.left);
                Clock c3;
                envL_.push_back(update(*envL_.back(), *mps_[i], *mpo_[i], true));
                tm_.update += c3.s();
                Clock c4;
                krylov(0, r.right, nullptr, nullptr, true);
                tm_.evl0 += c4.s();
                Clock c5;
                mps_[i + 1] = product(dev_, *r.right, *mps_[i + 1], 1, 'N', 'N');
                tm_.sweep += c5.s();
                envR_.pop_back();
            }
        }
        for (int i = ns - 1; i >= 0; --i) {
            Clock c;
            krylov(1, mps_[i], mpo_[i].get(), nullptr, false);
            tm_.evl1 += c.s();
            if (i > 0) {
                Clock c2;
                SvdResult r = split(*mps_[i], 1, false, i - 1);
                tm_.svd += c2.s();
                mps_[i] = std::move(r.right);
                Clock c3;
                envR_.push_back(update(*envR_.back(), *mps_[i], *mpo_[i], false));
                tm_.update += c3.s();
                Clock c4;
                krylov(0, r.left, nullptr, nullptr, true);
                tm_.evl0 += c4.s();
                Clock c5;
                mps_[i - 1] = product(dev_, *mps_[i - 1], *r.left, 1, 'N', 'N');
                tm_.sweep += c5.s();
                envL_.pop_back();
            }
        }
    }

    int nsites() const { return (int)mps_.size(); }
    const Tensor& site(int i) const { return *mps_[i]; }
    std::vector<int> bond_dims() const {
        std::vector<int> d;
        for (int i = 0; i + 1 < (int)mps_.size(); ++i) d.push_back(mps_[i]->lay.leg_dim(2));
        return d;
    }
    const TdvpTimers& timers() const { return tm_; }
    const LanczosStats& lanczos_stats() const { return ls_; }
    const KernelStats& heff_kernels() const { return heff_k_; }
    const KernelStats& env_kernels() const { return env_k_; }
    double env_flop() const { return env_flop_; }
    Device& device() { return *dev_; }
    void reset_stats() {
        tm_ = TdvpTimers();
        ls_ = LanczosStats();
        env_flop_ = 0;
        heff_k_ = KernelStats();
        env_k_ = KernelStats();
        dev_->trd = TrdStats();
        dev_->blk = BlockStats();
    }
    const TrdStats& trd_stats() const { return dev_->trd; }
    // States added by the enrichment SVDs since construction.
    long long sa_added() const { return sa_added_; }
    // Starts the SVS log at path ("Total Site = N", holstein/tdvp.h:71-74); every later
    // two-site SVD appends its record.
    void open_svs(const std::string& path) {
        svs_path_ = path;
        svs_write_header(path, (int)mps_.size());
    }
    bool timing = true;

  private:
    void log_svs(const SvdResult& r) {
        if (svs_path_.empty()) return;
        std::vector<double> kept;
        for (const auto& v : r.svals) kept.insert(kept.end(), v.begin(), v.end());
        std::sort(kept.begin(), kept.end(), [](double a, double b) { return a > b; });
        svs_append(svs_path_, kept);
    }
    // The two-site bond SVD of bond (i, i+1) in one sweep direction (S to the right when l2r).
    SvdResult bond_svd(const Tensor& mm, int /*i*/, bool l2r) {
        return svd_truncate(dev_, *sol_, mm, 2, cfg_.cutoff, cfg_.maxdim, l2r, -1, svd_opt_);
    }
    // Environments stay FP64 in every precision mode: they enter every later matvec, so
    // their rounding error would bias the effective Hamiltonian itself (they cost ~5% of
    // the H_eff work).
    std::unique_ptr<Tensor> update(const Tensor& env, const Tensor& A, Tensor& W, bool l2r) {
        auto out = env_update(dev_, env, nullptr, W, A, l2r, PREC_F64, timing ? &log_ : nullptr,
                              &env_flop_);
        if (timing) log_.resolve(env_k_);
        return out;
    }
    void krylov(int nsite, std::unique_ptr<Tensor>& x, Tensor* W1, Tensor* W2, bool reverse) {
        KrylovExit ex = local_krylov(dev_, nsite, *envL_.back(), W1, W2, *envR_.back(), x, cfg_.dt,
                                     cfg_.krylov_tol, cfg_.krylov_maxiter, reverse, cfg_.prec,
                                     timing ? &log_ : nullptr, &ls_, &heff_k_);
        last_iters_ = ex.iterations;
    }

    std::shared_ptr<Device> dev_;
    std::unique_ptr<Solver> sol_;
    std::vector<std::unique_ptr<Tensor>> mps_, mpo_;
    std::vector<std::unique_ptr<Tensor>> envL_, envR_;
    TdvpConfig cfg_;
    TdvpTimers tm_;
    LanczosStats ls_;
    EventLog log_;
    KernelStats heff_k_, env_k_;
    double env_flop_ = 0;
    bool verbose_ = false;
    int last_iters_ = 0;
    long long sa_added_ = 0;
    std::string svs_path_;
    SvdOptions svd_opt_ = SvdOptions::from_env();
};

// Value of the single 1x1x1 (or 1x1) closing block of a finished environment pass.
inline std::complex<double> closing_value(Tensor& env) {
    const auto& h = env.host_data();
    for (const auto& b : env.lay.blocks) {
        bool diag = b.qn[0] == b.qn[env.lay.rank - 1];
        for (int r = 1; r + 1 < env.lay.rank; ++r) diag = diag && b.qn[r] == 0;
        if (diag && b.size == 1) return {h[2 * b.off], h[2 * b.off + 1]};
    }
    return {0.0, 0.0};
}

// <bra|O|ket> (holstein/operator.h:178-198)
inline std::complex<double> join(const std::shared_ptr<Device>& dev, std::vector<Tensor*>& mpo,
                                 const std::vector<const Tensor*>& bra,
                                 const std::vector<const Tensor*>& ket) {
    auto env = left_boundary(dev);
    for (size_t i = 0; i < mpo.size(); ++i)
        env = env_update(dev, *env, bra[i], *mpo[i], *ket[i], true, PREC_F64);
    return closing_value(*env);
}

// <s1|s2> (holstein/state.h:206-227)
inline std::complex<double> overlap_states(const std::shared_ptr<Device>& dev,
                                           const std::vector<const Tensor*>& s1,
                                           const std::vector<const Tensor*>& s2) {
    Layout l;
    l.rank = 2;
    l.dirs = Idx{1, -1, 0, 0, 0, 0};
    BlockDesc b;
    b.shape = Idx{1, 1, 0, 0, 0, 0};
    l.blocks.push_back(b);
    l.finalize();
    auto env = Tensor::create(dev, l, false);
    const double one[2] = {1.0, 0.0};
    env->upload(one);
    for (size_t i = 0; i < s1.size(); ++i) {
        auto lm = product(dev, *env, *s2[i], 1, 'N', 'N');
        env = product(dev, *s1[i], *lm, 2, 'C', 'N');
    }
    return closing_value(*env);
}

}  // namespace kv
