// Reference binary MPS/MPO formats (holstein/state.h:107-157, holstein/operator.h:94-143):
//   INT magic (111 MPS, 112 MPO), INT nsites, then per site INT nblocks and per block
//   INT shape[rank], INT qn[rank], complex<double>[prod(shape)] row-major.
// INT is 32-bit (util/global.h:24, :44).  Save appends the trailing newline the reference
// writes (`ofs << endl`).
#pragma once

#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "tensor.h"

namespace kv {

inline std::vector<std::unique_ptr<Tensor>> load_train(const std::shared_ptr<Device>& dev,
                                                       const std::string& path, int rank,
                                                       int magic) {
    FILE* f = std::fopen(path.c_str(), "rb");
    KV_REQUIRE(f, ERR_ARG, "cannot open " + path);
    auto rd = [&](void* p, size_t n) {
        if (std::fread(p, 1, n, f) != n) {
            std::fclose(f);
            throw Error(ERR_FORMAT, "truncated file " + path);
        }
    };
    int32_t m = 0, ns = 0;
    rd(&m, 4);
    if (m != magic) {
        std::fclose(f);
        throw Error(ERR_FORMAT, magic == 111 ? "Illegal MPS file!" : "Illegal MPO file!");
    }
    rd(&ns, 4);
    KV_REQUIRE(ns > 0 && ns < (1 << 20), ERR_FORMAT, "bad site count");
    const Idx dirs_mps{1, 1, -1, 0, 0, 0}, dirs_mpo{1, 1, -1, -1, 0, 0};
    std::vector<std::unique_ptr<Tensor>> out;
    for (int i = 0; i < ns; ++i) {
        int32_t nb = 0;
        rd(&nb, 4);
        KV_REQUIRE(nb >= 0, ERR_FORMAT, "bad block count");
        std::vector<BlockDesc> blocks(nb);
        std::vector<std::vector<double>> data(nb);
        for (int j = 0; j < nb; ++j) {
            int32_t sh[4], qn[4];
            rd(sh, 4 * rank);
            rd(qn, 4 * rank);
            int64_t sz = 1;
            for (int r = 0; r < rank; ++r) {
                KV_REQUIRE(sh[r] > 0, ERR_FORMAT, "bad block shape");
                blocks[j].shape[r] = sh[r];
                blocks[j].qn[r] = qn[r];
                sz *= sh[r];
            }
            data[j].resize(2 * sz);
            rd(data[j].data(), 16 * sz);
        }
        // blocks with equal qn are summed (push_back_ch, holstein/tensor.h:173-186)
        Layout lay;
        lay.rank = rank;
        lay.dirs = rank == 3 ? dirs_mps : dirs_mpo;
        for (int j = 0; j < nb; ++j) {
            bool dup = false;
            for (const auto& b : lay.blocks)
                if (qn_equal(b.qn, blocks[j].qn, rank)) dup = true;
            if (!dup) lay.blocks.push_back(blocks[j]);
        }
        lay.finalize();
        KV_REQUIRE(lay.check_sym(), ERR_SYM, "block breaks charge conservation");
        std::vector<double> slab(2 * lay.total, 0.0);
        for (int j = 0; j < nb; ++j) {
            const int b = lay.find(blocks[j].qn);
            KV_REQUIRE(qn_equal(lay.blocks[b].shape, blocks[j].shape, rank), ERR_FORMAT,
                       "equal charges with different shapes");
            for (size_t e = 0; e < data[j].size(); ++e) slab[2 * lay.blocks[b].off + e] += data[j][e];
        }
        auto t = Tensor::create(dev, lay, false);
        t->upload(slab.data());
        out.push_back(std::move(t));
    }
    std::fclose(f);
    return out;
}

inline void save_train(const std::string& path, const std::vector<Tensor*>& sites, int rank,
                       int magic) {
    FILE* f = std::fopen(path.c_str(), "wb");
    KV_REQUIRE(f, ERR_ARG, "cannot open " + path);
    auto wr = [&](const void* p, size_t n) {
        if (std::fwrite(p, 1, n, f) != n) {
            std::fclose(f);
            throw Error(ERR_ARG, "write failed " + path);
        }
    };
    const int32_t m = magic, ns = (int32_t)sites.size();
    wr(&m, 4);
    wr(&ns, 4);
    for (Tensor* t : sites) {
        KV_REQUIRE(t->lay.rank == rank, ERR_ARG, "wrong tensor rank for this file type");
        const auto& h = t->host_data();
        const int32_t nb = (int32_t)t->lay.blocks.size();
        wr(&nb, 4);
        for (const auto& b : t->lay.blocks) {
            int32_t sh[4], qn[4];
            for (int r = 0; r < rank; ++r) {
                sh[r] = b.shape[r];
                qn[r] = b.qn[r];
            }
            wr(sh, 4 * rank);
            wr(qn, 4 * rank);
            wr(h.data() + 2 * b.off, 16 * (size_t)b.size);
        }
    }
    wr("\n", 1);
    std::fclose(f);
}

}  // namespace kv
