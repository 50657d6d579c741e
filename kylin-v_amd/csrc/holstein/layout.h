// Block structure of a U(1) block-sparse tensor: per-block charge tuple and shape, and
// the block's offset inside one contiguous slab.  Host-only C++ (no HIP).
//
// The reference keeps a vector of separately allocated blocks and finds a block by a
// linear (shape, qn) scan (holstein/tensor.h:173-186).  Here a tensor is ONE slab plus a
// table sorted by charge tuple: lookups are binary searches and every vector operation
// of the Lanczos recurrence is a flat pass over the slab.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <sstream>
#include <string>
#include <vector>

#include "../kv_error.h"

namespace kv {

constexpr int MAXR = 6;
using Idx = std::array<int, MAXR>;

struct BlockDesc {
    Idx qn{};
    Idx shape{};
    int64_t off = 0;
    int64_t size = 0;
};

inline bool qn_less(const Idx& a, const Idx& b, int rank) {
    for (int i = 0; i < rank; ++i)
        if (a[i] != b[i]) return a[i] < b[i];
    return false;
}
inline bool qn_equal(const Idx& a, const Idx& b, int rank) {
    for (int i = 0; i < rank; ++i)
        if (a[i] != b[i]) return false;
    return true;
}

struct Layout {
    int rank = 0;
    Idx dirs{};
    std::vector<BlockDesc> blocks;  // ascending qn after finalize()
    int64_t total = 0;

    // Sort blocks by qn, assign offsets, reject duplicate charge tuples.
    void finalize() {
        std::sort(blocks.begin(), blocks.end(), [this](const BlockDesc& a, const BlockDesc& b) {
            return qn_less(a.qn, b.qn, rank);
        });
        total = 0;
        for (size_t i = 0; i < blocks.size(); ++i) {
            if (i > 0 && qn_equal(blocks[i - 1].qn, blocks[i].qn, rank))
                throw Error(ERR_ARG, "duplicate block charge tuple");
            int64_t s = 1;
            for (int r = 0; r < rank; ++r) {
                KV_REQUIRE(blocks[i].shape[r] > 0, ERR_ARG, "block dimension must be positive");
                s *= blocks[i].shape[r];
            }
            blocks[i].size = s;
            blocks[i].off = total;
            total += s;
        }
    }
    int find(const Idx& q) const {
        size_t lo = 0, hi = blocks.size();
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (qn_less(blocks[mid].qn, q, rank))
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < blocks.size() && qn_equal(blocks[lo].qn, q, rank)) return (int)lo;
        return -1;
    }
    int64_t prod(int b, int from, int to) const {
        int64_t p = 1;
        for (int r = from; r < to; ++r) p *= blocks[b].shape[r];
        return p;
    }
    // every block conserves charge: sum(dir * qn) == 0 (holstein/tensor.h:201-211)
    bool check_sym() const {
        for (const auto& b : blocks) {
            int s = 0;
            for (int r = 0; r < rank; ++r) s += dirs[r] * b.qn[r];
            if (s != 0) return false;
        }
        return true;
    }
    // total bond dimension of leg ax: sum over distinct qn of the leg dimension
    int leg_dim(int ax) const {
        std::vector<std::pair<int, int>> seen;
        for (const auto& b : blocks) {
            bool f = false;
            for (auto& s : seen)
                if (s.first == b.qn[ax]) f = true;
            if (!f) seen.push_back({b.qn[ax], b.shape[ax]});
        }
        int d = 0;
        for (auto& s : seen) d += s.second;
        return d;
    }
    std::string describe() const {
        std::ostringstream os;
        os << "rank " << rank << " blocks " << blocks.size() << " total " << total;
        return os.str();
    }
};

inline bool same_layout(const Layout& a, const Layout& b) {
    if (a.rank != b.rank || a.blocks.size() != b.blocks.size()) return false;
    for (int r = 0; r < a.rank; ++r)
        if (a.dirs[r] != b.dirs[r]) return false;
    for (size_t i = 0; i < a.blocks.size(); ++i) {
        if (!qn_equal(a.blocks[i].qn, b.blocks[i].qn, a.rank)) return false;
        if (!qn_equal(a.blocks[i].shape, b.blocks[i].shape, a.rank)) return false;
    }
    return true;
}

// Union of the block sets of two layouts with equal rank and dirs.
inline Layout union_layout(const Layout& a, const Layout& b) {
    KV_REQUIRE(a.rank == b.rank, ERR_INTERNAL, "union of layouts of different rank");
    Layout u;
    u.rank = a.rank;
    u.dirs = a.dirs;
    u.blocks = a.blocks;
    for (const auto& bb : b.blocks) {
        const int i = a.find(bb.qn);
        if (i < 0) {
            u.blocks.push_back(bb);
        } else {
            KV_REQUIRE(qn_equal(a.blocks[i].shape, bb.shape, a.rank), ERR_ARG,
                       "blocks with equal charges but different shapes");
        }
    }
    u.finalize();
    return u;
}

}  // namespace kv
