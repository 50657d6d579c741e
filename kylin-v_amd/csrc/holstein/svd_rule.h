// Sector structure and the truncation rule of the bond SVD.  Host-only C++.
//
// Follows svd_impl<R1,R2> (holstein/svd.h:296-440): the row charge of a block is
// sum_{l<R1} dir_l qn_l and its column charge sum_{l>=R1} dir_l qn_l (fill_block,
// :110-205); a row sector pairs with the column sector of opposite charge (:314); all
// singular values go into one descending list (:351-353) and, walking from the
// smallest, entry i is dropped while i >= MaxDim or s_i <= SVDThres, never entry 0
// (:361-388).  The reference sorts with std::sort (unstable); ties here are broken by
// (sector order, index) so the result is deterministic.
#pragma once

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "layout.h"

namespace kv {

struct SvdSub {
    Idx qn{};     // the sub-block's charges on the legs of this side
    Idx shape{};  // and its dims
    int64_t pos = 0, size = 0;
};
struct SvdSector {
    int q = 0;
    std::vector<SvdSub> subs;  // ascending qn
    int64_t len = 0;
};
struct SvdStructure {
    int r1 = 0, r2 = 0;
    std::vector<SvdSector> rows, cols;
    std::vector<std::pair<int, int>> pairs;  // (row sector, col sector), ascending row charge
};

inline SvdStructure svd_structure(const Layout& A, int r1) {
    SvdStructure S;
    S.r1 = r1;
    S.r2 = A.rank - r1;
    KV_REQUIRE(r1 >= 1 && S.r2 >= 1, ERR_ARG, "svd split must leave legs on both sides");
    std::map<int, std::map<std::vector<int>, Idx>> rows, cols;
    for (const auto& b : A.blocks) {
        int qr = 0, qc = 0;
        for (int l = 0; l < r1; ++l) qr += A.dirs[l] * b.qn[l];
        for (int l = r1; l < A.rank; ++l) qc += A.dirs[l] * b.qn[l];
        std::vector<int> rk(b.qn.begin(), b.qn.begin() + r1), ck(b.qn.begin() + r1, b.qn.begin() + A.rank);
        Idx rs{}, cs{};
        for (int l = 0; l < r1; ++l) rs[l] = b.shape[l];
        for (int l = r1; l < A.rank; ++l) cs[l - r1] = b.shape[l];
        auto& rmap = rows[qr];
        auto itr = rmap.find(rk);
        if (itr == rmap.end()) rmap.emplace(rk, rs);
        else KV_REQUIRE(qn_equal(itr->second, rs, r1), ERR_ARG, "inconsistent leg dims in SVD input");
        auto& cmap = cols[qc];
        auto itc = cmap.find(ck);
        if (itc == cmap.end()) cmap.emplace(ck, cs);
        else KV_REQUIRE(qn_equal(itc->second, cs, S.r2), ERR_ARG, "inconsistent leg dims in SVD input");
    }
    auto build = [](const std::map<int, std::map<std::vector<int>, Idx>>& m, int nl,
                    std::vector<SvdSector>& out) {
        for (auto& kv1 : m) {
            SvdSector sec;
            sec.q = kv1.first;
            for (auto& kv2 : kv1.second) {
                SvdSub s;
                for (int l = 0; l < nl; ++l) s.qn[l] = kv2.first[l];
                s.shape = kv2.second;
                s.size = 1;
                for (int l = 0; l < nl; ++l) s.size *= s.shape[l];
                s.pos = sec.len;
                sec.len += s.size;
                sec.subs.push_back(s);
            }
            out.push_back(sec);
        }
    };
    build(rows, r1, S.rows);
    build(cols, S.r2, S.cols);
    for (int i = 0; i < (int)S.rows.size(); ++i)
        for (int j = 0; j < (int)S.cols.size(); ++j)
            if (S.rows[i].q + S.cols[j].q == 0) S.pairs.push_back({i, j});
    return S;
}

// kept[p] = number of singular values kept in sector pair p.
inline std::vector<int> truncation_counts(const std::vector<std::vector<double>>& svals,
                                          double cutoff, int maxdim) {
    std::vector<std::tuple<double, int, int>> all;
    std::vector<int> kept(svals.size());
    for (int p = 0; p < (int)svals.size(); ++p) {
        kept[p] = (int)svals[p].size();
        for (int i = 0; i < (int)svals[p].size(); ++i) all.emplace_back(svals[p][i], p, i);
    }
    std::sort(all.begin(), all.end(), [](const auto& a, const auto& b) {
        if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) > std::get<0>(b);
        if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) < std::get<1>(b);
        return std::get<2>(a) < std::get<2>(b);
    });
    for (int i = (int)all.size() - 1; i > 0; --i) {
        if (i >= maxdim || std::get<0>(all[i]) <= cutoff)
            kept[std::get<1>(all[i])] -= 1;
        else
            break;
    }
    return kept;
}

}  // namespace kv
