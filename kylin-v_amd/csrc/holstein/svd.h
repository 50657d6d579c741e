// Bond SVD with truncation, FP64 on the device (rocSOLVER zgesvd per charge sector).
// Follows svd_impl<R1,R2> (holstein/svd.h:296-440); the sector structure and the
// truncation rule are in svd_rule.h.
//
// Each sector matrix M (rows x cols, row-major) is gathered from the blocks.  rocSOLVER is
// column-major, so the buffer is handed over as M^T (cols x rows).  With
// M^T = U' S V'^H, M = conj(V') S U'^T, so rocSOLVER's V^H output read row-major with
// ld = k is exactly U_M (rows x k), and its U output read row-major with ld = cols is
// exactly V_M^H (k x cols): no extra transposes.
#pragma once

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <memory>
#include <vector>

#include "svd_rule.h"
#include "tensor.h"

namespace kv {

struct Solver {
    rocblas_handle h = nullptr;
    explicit Solver(Device& dev) {
        if (rocblas_create_handle(&h) != rocblas_status_success)
            throw Error(ERR_SOLVER, "rocblas_create_handle failed");
        if (rocblas_set_stream(h, dev.stream) != rocblas_status_success)
            throw Error(ERR_SOLVER, "rocblas_set_stream failed");
    }
    ~Solver() {
        if (h) rocblas_destroy_handle(h);
    }
};

struct SvdResult {
    std::unique_ptr<Tensor> left, right;
    int kept = 0;
    int total = 0;  // number of singular values before truncation
    std::vector<std::vector<double>> svals;  // per sector pair, kept values
};

inline SvdResult svd_truncate(const std::shared_ptr<Device>& dev, Solver& sol, const Tensor& A,
                              int r1, double cutoff, int maxdim, bool left2right,
                              int left_dir = -1) {
    Device& D = *dev;
    const SvdStructure S = svd_structure(A.lay, r1);
    const int np = (int)S.pairs.size();
    std::vector<std::vector<double>> svals(np);
    std::vector<DevBuf> Ub(np), Vb(np), Sb(np);
    std::vector<int> kdim(np);
    for (int p = 0; p < np; ++p) {
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        const int64_t nr = rs.len, nc = cs.len;
        KV_REQUIRE(nr < (1 << 30) && nc < (1 << 30), ERR_ARG, "sector matrix too large");
        const int k = (int)std::min(nr, nc);
        kdim[p] = k;
        DevBuf M(D.pool.get(), (size_t)(nr * nc) * sizeof(double2));
        launch_zero(D.stream, M.get(), M.bytes());
        std::vector<Copy2D> cps;
        int maxrows = 1;
        for (const auto& rsub : rs.subs)
            for (const auto& csub : cs.subs) {
                Idx q{};
                for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
                for (int l = 0; l < S.r2; ++l) q[S.r1 + l] = csub.qn[l];
                const int bi = A.lay.find(q);
                if (bi < 0) continue;
                Copy2D c{};
                c.src_off = A.lay.blocks[bi].off;
                c.dst_off = rsub.pos * nc + csub.pos;
                c.rows = (int)rsub.size;
                c.cols = (int)csub.size;
                c.src_ld = (int)csub.size;
                c.dst_ld = (int)nc;
                maxrows = std::max(maxrows, c.rows);
                cps.push_back(c);
            }
        if (!cps.empty()) {
            DevBuf dc = upload_vec(D, cps);
            launch_copy2d<double2>(D.stream, A.d(), M.as<double2>(), dc.as<Copy2D>(),
                                   (int)cps.size(), maxrows, nullptr);
            D.sync();
        }
        Sb[p] = DevBuf(D.pool.get(), (size_t)k * sizeof(double));
        Ub[p] = DevBuf(D.pool.get(), (size_t)(nc * k) * sizeof(double2));  // V_M^H  (k x nc)
        Vb[p] = DevBuf(D.pool.get(), (size_t)(k * nr) * sizeof(double2));  // U_M    (nr x k)
        DevBuf E(D.pool.get(), (size_t)std::max(k, 1) * sizeof(double));
        DevBuf info(D.pool.get(), sizeof(rocblas_int));
        const rocblas_status st = rocsolver_zgesvd(
            sol.h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)nc,
            (rocblas_int)nr, reinterpret_cast<rocblas_double_complex*>(M.get()), (rocblas_int)nc,
            Sb[p].as<double>(), reinterpret_cast<rocblas_double_complex*>(Ub[p].get()),
            (rocblas_int)nc, reinterpret_cast<rocblas_double_complex*>(Vb[p].get()),
            (rocblas_int)k, E.as<double>(), rocblas_outofplace, info.as<rocblas_int>());
        KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, "rocsolver_zgesvd failed");
        rocblas_int hinfo = 0;
        svals[p].resize(k);
        KV_HIP(hipMemcpyAsync(&hinfo, info.get(), sizeof(rocblas_int), hipMemcpyDeviceToHost,
                              D.stream));
        KV_HIP(hipMemcpyAsync(svals[p].data(), Sb[p].get(), k * sizeof(double),
                              hipMemcpyDeviceToHost, D.stream));
        D.sync();
        KV_REQUIRE(hinfo == 0, ERR_SOLVER, "Problems in SVD (zgesvd did not converge)");
    }
    std::vector<int> kept = truncation_counts(svals, cutoff, maxdim);
    SvdResult res;
    // output layouts
    Layout Ll, Rl;
    Ll.rank = S.r1 + 1;
    Rl.rank = S.r2 + 1;
    for (int l = 0; l < S.r1; ++l) Ll.dirs[l] = A.lay.dirs[l];
    Ll.dirs[S.r1] = left_dir;
    Rl.dirs[0] = -left_dir;
    for (int l = 0; l < S.r2; ++l) Rl.dirs[1 + l] = A.lay.dirs[S.r1 + l];
    for (int p = 0; p < np; ++p) {
        if (kept[p] <= 0) continue;
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        for (const auto& rsub : rs.subs) {
            BlockDesc b;
            for (int l = 0; l < S.r1; ++l) {
                b.qn[l] = rsub.qn[l];
                b.shape[l] = rsub.shape[l];
            }
            b.qn[S.r1] = -left_dir * rs.q;
            b.shape[S.r1] = kept[p];
            Ll.blocks.push_back(b);
        }
        for (const auto& csub : cs.subs) {
            BlockDesc b;
            b.qn[0] = left_dir * cs.q;
            b.shape[0] = kept[p];
            for (int l = 0; l < S.r2; ++l) {
                b.qn[1 + l] = csub.qn[l];
                b.shape[1 + l] = csub.shape[l];
            }
            Rl.blocks.push_back(b);
        }
        res.kept += kept[p];
    }
    for (auto& v : svals) res.total += (int)v.size();
    Ll.finalize();
    Rl.finalize();
    res.left = Tensor::create(dev, Ll, false);
    res.right = Tensor::create(dev, Rl, false);
    for (int p = 0; p < np; ++p) {
        if (kept[p] <= 0) continue;
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        const int k = kdim[p], nk = kept[p];
        const int64_t nc = cs.len;
        std::vector<Copy2D> lc, rc;
        int maxr_l = 1, maxr_r = 1;
        for (const auto& rsub : rs.subs) {
            Idx q{};
            for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
            q[S.r1] = -left_dir * rs.q;
            const int bi = Ll.find(q);
            Copy2D c{};
            c.src_off = rsub.pos * (int64_t)k;  // U_M rows [pos, pos+size), ld k
            c.dst_off = Ll.blocks[bi].off;
            c.rows = (int)rsub.size;
            c.cols = nk;
            c.src_ld = k;
            c.dst_ld = nk;
            c.scale_mode = left2right ? 0 : 2;  // U S when S goes left
            c.scale_off = 0;
            maxr_l = std::max(maxr_l, c.rows);
            lc.push_back(c);
        }
        for (const auto& csub : cs.subs) {
            Idx q{};
            q[0] = left_dir * cs.q;
            for (int l = 0; l < S.r2; ++l) q[1 + l] = csub.qn[l];
            const int bi = Rl.find(q);
            Copy2D c{};
            c.src_off = csub.pos;  // V_M^H rows [0, nk), cols [pos, pos+size), ld nc
            c.dst_off = Rl.blocks[bi].off;
            c.rows = nk;
            c.cols = (int)csub.size;
            c.src_ld = (int)nc;
            c.dst_ld = (int)csub.size;
            c.scale_mode = left2right ? 1 : 0;  // S V^H when S goes right
            c.scale_off = 0;
            maxr_r = std::max(maxr_r, c.rows);
            rc.push_back(c);
        }
        DevBuf dl = upload_vec(D, lc), dr = upload_vec(D, rc);
        launch_copy2d<double2>(D.stream, Vb[p].as<double2>(), res.left->d(), dl.as<Copy2D>(),
                               (int)lc.size(), maxr_l, Sb[p].as<double>());
        launch_copy2d<double2>(D.stream, Ub[p].as<double2>(), res.right->d(), dr.as<Copy2D>(),
                               (int)rc.size(), maxr_r, Sb[p].as<double>());
        D.sync();
        svals[p].resize(nk);
    }
    res.svals = std::move(svals);
    return res;
}

}  // namespace kv
