// The enrichment SVD of SA-1TDVP on the device.
//
// Replaces sa_svd_impl<R1,R2> (holstein/svd.h:443-614) with gs_ortho (util/allocation.h:
// 260-327): every sector is decomposed in full (rocSOLVER's one-sided SVD, nothing
// truncated, S absorbed as svd_impl does), sa_extension_count decides how many states to
// add, and the sectors with room on the isometric side (rows when sweeping right, columns
// when sweeping left), in ascending row charge, receive seeded random vectors made
// orthonormal to the sector's singular vectors and to each other; the other factor gets
// zero rows (columns), so the product of the two factors is unchanged.
// The orthonormalisation is block classical Gram-Schmidt against the singular vectors
// (two passes, grouped GEMMs) followed by Cholesky QR of the new block (two passes): in
// exact arithmetic each new vector is the normalised part of its random start outside the
// span of everything before it, which is what the reference's vector-by-vector modified
// Gram-Schmidt computes.
#pragma once

#include <climits>
#include <memory>
#include <vector>

#include "sa_rule.h"
#include "svd.h"

namespace kv {

struct SaSvdResult {
    SvdResult svd;   // left, right with the new states; kept = values + added
    int values = 0;  // singular values of the full decomposition
    int added = 0;   // states placed
};

inline SaSvdResult svd_enrich(const std::shared_ptr<Device>& dev, Solver& sol, const Tensor& A, int r1,
                              bool left2right, double cutoff, int maxdim, double beta, uint64_t seed,
                              uint64_t call, int left_dir = -1) {
    Device& D = *dev;
    SvdOptions opt;
    opt.force_exact = true;
    opt.block = false;
    SaSvdResult out;
    SvdResult r = svd_truncate(dev, sol, A, r1, -1.0, 1 << 30, left2right, left_dir, opt);
    const SvdStructure S = svd_structure(A.lay, r1);
    const int np = (int)S.pairs.size();
    KV_REQUIRE((int)r.svals.size() == np, ERR_INTERNAL, "enrichment SVD: sector count mismatch");
    std::vector<double> all;
    for (const auto& v : r.svals) all.insert(all.end(), v.begin(), v.end());
    std::sort(all.begin(), all.end(), [](double a, double b) { return a > b; });
    out.values = (int)all.size();
    int rem = sa_extension_count(all, maxdim, cutoff, beta, seed, call);
    std::vector<int> k(np), add(np, 0);
    for (int p = 0; p < np; ++p) {
        k[p] = (int)r.svals[p].size();
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        const int64_t room = (left2right ? rs.len : cs.len) - k[p];
        if (rem > 0 && room > 0) {
            add[p] = (int)std::min<int64_t>(rem, room);
            rem -= add[p];
            out.added += add[p];
        }
    }
    if (out.added == 0) {
        out.svd = std::move(r);
        return out;
    }
    // layouts with the new bond dimensions
    Layout Ll, Rl;
    Ll.rank = S.r1 + 1;
    Rl.rank = S.r2 + 1;
    for (int l = 0; l < S.r1; ++l) Ll.dirs[l] = A.lay.dirs[l];
    Ll.dirs[S.r1] = left_dir;
    Rl.dirs[0] = -left_dir;
    for (int l = 0; l < S.r2; ++l) Rl.dirs[1 + l] = A.lay.dirs[S.r1 + l];
    for (int p = 0; p < np; ++p) {
        if (k[p] <= 0) continue;
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        for (const auto& rsub : rs.subs) {
            BlockDesc b;
            for (int l = 0; l < S.r1; ++l) {
                b.qn[l] = rsub.qn[l];
                b.shape[l] = rsub.shape[l];
            }
            b.qn[S.r1] = -left_dir * rs.q;
            b.shape[S.r1] = k[p] + add[p];
            Ll.blocks.push_back(b);
        }
        for (const auto& csub : cs.subs) {
            BlockDesc b;
            b.qn[0] = left_dir * cs.q;
            b.shape[0] = k[p] + add[p];
            for (int l = 0; l < S.r2; ++l) {
                b.qn[1 + l] = csub.qn[l];
                b.shape[1 + l] = csub.shape[l];
            }
            Rl.blocks.push_back(b);
        }
    }
    Ll.finalize();
    Rl.finalize();
    auto left = Tensor::create(dev, Ll, true);
    auto right = Tensor::create(dev, Rl, true);
    const Tensor& oL = *r.left;
    const Tensor& oR = *r.right;
    int v0 = 0;  // index of the call's next new vector
    for (int p = 0; p < np; ++p) {
        if (k[p] <= 0) continue;
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        const int kp = k[p], ap = add[p], kn = kp + ap;
        // the decomposition's own columns / rows into the leading part of every block
        std::vector<Copy2D> lc, rc;
        int mr_l = 1, mr_r = 1;
        for (const auto& rsub : rs.subs) {
            Idx q{};
            for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
            q[S.r1] = -left_dir * rs.q;
            const int bo = oL.lay.find(q), bn = Ll.find(q);
            KV_REQUIRE(bo >= 0 && bn >= 0, ERR_INTERNAL, "enrichment SVD: left block missing");
            lc.push_back(Copy2D{oL.lay.blocks[bo].off, Ll.blocks[bn].off, (int)rsub.size, kp, kp, kn, 0, 0});
            mr_l = std::max(mr_l, (int)rsub.size);
        }
        for (const auto& csub : cs.subs) {
            Idx q{};
            q[0] = left_dir * cs.q;
            for (int l = 0; l < S.r2; ++l) q[1 + l] = csub.qn[l];
            const int bo = oR.lay.find(q), bn = Rl.find(q);
            KV_REQUIRE(bo >= 0 && bn >= 0, ERR_INTERNAL, "enrichment SVD: right block missing");
            rc.push_back(Copy2D{oR.lay.blocks[bo].off, Rl.blocks[bn].off, kp, (int)csub.size, (int)csub.size,
                                (int)csub.size, 0, 0});
            mr_r = std::max(mr_r, kp);
        }
        {
            DevBuf dl = upload_vec(D, lc), dr = upload_vec(D, rc);
            launch_copy2d<double2>(D.stream, oL.d(), left->d(), dl.as<Copy2D>(), (int)lc.size(), mr_l, nullptr);
            launch_copy2d<double2>(D.stream, oR.d(), right->d(), dr.as<Copy2D>(), (int)rc.size(), mr_r, nullptr);
        }
        if (ap == 0) continue;
        KV_REQUIRE(ap <= kCholMax, ERR_ARG, "SA-1TDVP: more new states in one sector than the orthonormaliser takes");
        // Old (len x kp, row-major): the isometric side's vectors as columns (conjugated
        // rows of V^H when sweeping left; a real random start keeps the conjugation exact)
        const int len = (int)(left2right ? rs.len : cs.len);
        DevBuf Old(D.pool.get(), (size_t)len * kp * sizeof(double2));
        if (left2right) {
            std::vector<Copy2D> g;
            int mr = 1;
            for (const auto& rsub : rs.subs) {
                Idx q{};
                for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
                q[S.r1] = -left_dir * rs.q;
                const int bo = oL.lay.find(q);
                g.push_back(Copy2D{oL.lay.blocks[bo].off, rsub.pos * (int64_t)kp, (int)rsub.size, kp, kp, kp, 0, 0});
                mr = std::max(mr, (int)rsub.size);
            }
            DevBuf dg = upload_vec(D, g);
            launch_copy2d<double2>(D.stream, oL.d(), Old.as<double2>(), dg.as<Copy2D>(), (int)g.size(), mr, nullptr);
        } else {
            DevBuf Vs(D.pool.get(), (size_t)kp * len * sizeof(double2));
            std::vector<Copy2D> g;
            for (const auto& csub : cs.subs) {
                Idx q{};
                q[0] = left_dir * cs.q;
                for (int l = 0; l < S.r2; ++l) q[1 + l] = csub.qn[l];
                const int bo = oR.lay.find(q);
                g.push_back(Copy2D{oR.lay.blocks[bo].off, csub.pos, kp, (int)csub.size, (int)csub.size, len, 0, 0});
            }
            DevBuf dg = upload_vec(D, g);
            launch_copy2d<double2>(D.stream, oR.d(), Vs.as<double2>(), dg.as<Copy2D>(), (int)g.size(), kp, nullptr);
            launch_ctranspose(D.stream, Vs.as<double2>(), Old.as<double2>(), kp, len, len, kp);
        }
        // New (len x ap): entry (e, i) = the e-th draw of vector v0 + i, real
        std::vector<double> h((size_t)2 * len * ap, 0.0);
        for (int i = 0; i < ap; ++i)
            for (int e = 0; e < len; ++e)
                h[2 * ((size_t)e * ap + i)] = 2.0 * sa_u01(seed, call, 1 + (uint64_t)(v0 + i), (uint64_t)e) - 1.0;
        DevBuf New(D.pool.get(), h.size() * sizeof(double));
        D.upload(New.get(), h.data(), h.size() * sizeof(double));
        DevBuf C(D.pool.get(), (size_t)kp * ap * sizeof(double2)), P(D.pool.get(), (size_t)len * ap * sizeof(double2));
        for (int pass = 0; pass < 2; ++pass) {
            gemm_dense(D, 'C', 'N', Old.as<double2>(), kp, New.as<double2>(), ap, C.as<double2>(), ap, kp, ap, len);
            gemm_dense(D, 'N', 'N', Old.as<double2>(), kp, C.as<double2>(), ap, P.as<double2>(), ap, len, ap, kp);
            launch_zaxpby(D.stream, New.as<double2>(), New.as<double2>(), P.as<double2>(), double2{1.0, 0.0},
                          double2{-1.0, 0.0}, (int64_t)len * ap);
        }
        std::vector<DevBuf*> ys{&New};
        block_orth(D, ys, {len}, ap, 2, false, seed ^ call);
        // the new vectors into columns [kp, kn) of the left blocks, or rows [kp, kn) of the right
        if (left2right) {
            std::vector<Copy2D> s;
            int mr = 1;
            for (const auto& rsub : rs.subs) {
                Idx q{};
                for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
                q[S.r1] = -left_dir * rs.q;
                const int bn = Ll.find(q);
                s.push_back(Copy2D{rsub.pos * (int64_t)ap, Ll.blocks[bn].off + kp, (int)rsub.size, ap, ap, kn, 0, 0});
                mr = std::max(mr, (int)rsub.size);
            }
            DevBuf ds = upload_vec(D, s);
            launch_copy2d<double2>(D.stream, New.as<double2>(), left->d(), ds.as<Copy2D>(), (int)s.size(), mr, nullptr);
        } else {
            DevBuf NT(D.pool.get(), (size_t)ap * len * sizeof(double2));
            launch_ctranspose(D.stream, New.as<double2>(), NT.as<double2>(), len, ap, ap, len);
            std::vector<Copy2D> s;
            for (const auto& csub : cs.subs) {
                Idx q{};
                q[0] = left_dir * cs.q;
                for (int l = 0; l < S.r2; ++l) q[1 + l] = csub.qn[l];
                const int bn = Rl.find(q);
                s.push_back(Copy2D{csub.pos, Rl.blocks[bn].off + (int64_t)kp * csub.size, ap, (int)csub.size, len,
                                   (int)csub.size, 0, 0});
            }
            DevBuf ds = upload_vec(D, s);
            launch_copy2d<double2>(D.stream, NT.as<double2>(), right->d(), ds.as<Copy2D>(), (int)s.size(), ap, nullptr);
        }
        v0 += ap;
    }
    out.svd.left = std::move(left);
    out.svd.right = std::move(right);
    out.svd.kept = out.values + out.added;
    out.svd.total = out.values;
    out.svd.svals = std::move(r.svals);
    for (int p = 0; p < np; ++p) out.svd.kept_by_q[S.rows[S.pairs[p].first].q] = k[p] + add[p];
    return out;
}

}  // namespace kv
