// Lanczos exponential of the effective Hamiltonian on the device, FP64 vector algebra.
// Follows local_krylov<Rank> (holstein/exponential.h:92-190) step by step:
//   v0 = x/|x| (:97-98); h = (dt/2) H_eff v_k (:113-125);
//   beta = <h|v_{k-1}> with the conjugate on h, as overlap(v_{k-1}, h) computes it, then
//   h -= beta v_{k-1} (:126-132); alpha = <h|v_k>, h -= alpha v_k (:134-137);
//   err *= |h|, stop when err < tol (:139-147); else v_{k+1} = h/|h| while
//   k < MaxIter-1 (:109, :156-159);  T = diag(alpha) + upper(beta), Hermitian
//   eigen-decomposition, c = V exp(-/+ i Lambda) V^H e_0 (:162-183);  x = sum c_i v_i
//   without renormalisation (:184-186).
// The H_eff contractions may run in complex FP32 (prec) from the second Krylov vector on;
// the first matvec and everything else in this file are FP64.
#pragma once

#include <chrono>
#include <complex>
#include <memory>
#include <vector>

#include "herm_eig.h"
#include "product.h"

namespace kv {

struct KrylovExit {
    bool converged = false;
    int iterations = 0;
    double error = 0;
    double loss_norm = 0;
};

struct LanczosStats {
    double heff_flop = 0;
    long matvecs = 0;
};

// x is replaced by exp(-/+ i dt/2 H_eff) x/|x| in H_eff's closed block layout.
inline KrylovExit local_krylov(const std::shared_ptr<Device>& dev, int nsite, const Tensor& L,
                               Tensor* W1, Tensor* W2, const Tensor& R,
                               std::unique_ptr<Tensor>& x, double dt, double tol, int max_iter,
                               bool reverse, Prec prec, EventLog* log = nullptr,
                               LanczosStats* stats = nullptr, KernelStats* kstats = nullptr) {
    Heff H;
    H.build(dev, nsite, L, W1, W2, R, x->lay, prec);
    const Layout& xl = H.layout();
    const int64_t n = xl.total;
    const int nv_max = std::max(max_iter, 2);
    Device& D = *dev;
    std::unique_ptr<Tensor> xin = relayout(dev, *x, xl);
    DevBuf V(D.pool.get(), (size_t)nv_max * (size_t)n * sizeof(double2));
    DevBuf hv(D.pool.get(), (size_t)n * sizeof(double2));
    // device scalars: [0..1] beta, [2..3] alpha, [4] |h|^2, [5] |x|^2 ; reduction scratch
    DevBuf sc(D.pool.get(), 8 * sizeof(double));
    DevBuf scratch(D.pool.get(), 2 * 1024 * sizeof(double));
    double* s = sc.as<double>();
    double2* Vb = V.as<double2>();
    launch_nrm2sq(D.stream, xin->d(), n, s + 5, scratch.as<double>());
    launch_scale_invsqrt_dev(D.stream, Vb, xin->d(), s + 5, n);

    std::vector<std::complex<double>> alp, bet_up;
    int k = 0;
    double err = 1.0;
    KrylovExit ex;
    bool conv = false;
    double host[6];
    while (k < max_iter - 1 && err > tol) {
        double2* vk = Vb + (int64_t)k * n;
        // Mixed precision: the first matvec runs in FP64.  Its rounding error enters the
        // result at first order in dt/2 |H v|; the errors of later Krylov vectors are
        // damped by further powers of the Lanczos coefficients, so FP32 suffices there.
        H.apply(vk, hv.as<double2>(), dt * 0.5, log, k == 0);
        if (stats) {
            stats->heff_flop += H.flop();
            stats->matvecs += 1;
        }
        // beta = <h|v_{k-1}>, h -= beta v_{k-1}, alpha = <h|v_k>, h -= alpha v_k, |h|^2: the
        // reference's order, each subtraction fused with the reduction after it
        double2* h = hv.as<double2>();
        double2* sb = reinterpret_cast<double2*>(s);
        if (k > 0) {
            double2* vkm = Vb + (int64_t)(k - 1) * n;
            launch_dotc(D.stream, h, vkm, n, sb, scratch.as<double>());
            launch_axpy_dotc(D.stream, h, vkm, sb, vk, n, sb + 1, scratch.as<double>());
        } else {
            launch_dotc(D.stream, h, vk, n, sb + 1, scratch.as<double>());
        }
        launch_axpy_nrm2sq(D.stream, h, vk, sb + 1, n, s + 4, scratch.as<double>());
        KV_HIP(hipMemcpyAsync(host, s, 5 * sizeof(double), hipMemcpyDeviceToHost, D.stream));
        D.sync();
        if (log) log->resolve(kstats ? *kstats : D.stats);
        if (k > 0) bet_up.emplace_back(host[0], host[1]);
        alp.emplace_back(host[2], host[3]);
        const double nm = std::sqrt(host[4]);
        err *= nm;
        if (err < tol) {
            conv = true;
            break;
        }
        launch_scale_invsqrt_dev(D.stream, Vb + (int64_t)(k + 1) * n, hv.as<double2>(), s + 4, n);
        ++k;
    }
    ex.converged = conv;
    ex.iterations = (int)alp.size();
    ex.error = err;
    std::vector<std::complex<double>> c = krylov_coefficients(alp, bet_up, reverse);
    DevBuf cd(D.pool.get(), c.size() * sizeof(double2));
    KV_HIP(hipMemcpyAsync(cd.get(), c.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice,
                          D.stream));
    auto out = Tensor::create(dev, xl, false);
    launch_lincomb(D.stream, out->d(), Vb, cd.as<double2>(), (int)c.size(), n);
    launch_nrm2sq(D.stream, out->d(), n, s + 5, scratch.as<double>());
    KV_HIP(hipMemcpyAsync(host + 5, s + 5, sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    ex.loss_norm = 1.0 - std::sqrt(host[5]);
    x = std::move(out);
    return ex;
}

}  // namespace kv
