// Device-resident U(1) block-sparse tensor: a Layout (sorted block table) plus one FP64
// complex slab in HBM.  Replaces Block<Rank>/Tensor<Rank> (holstein/block.h:23-162,
// holstein/tensor.h:16-212), whose blocks are separate host mallocs.
#pragma once

#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "../kv_kernels.h"
#include "../kv_runtime.h"
#include "layout.h"

namespace kv {

struct Tensor {
    std::shared_ptr<Device> dev;
    Layout lay;
    DevBuf buf;                 // double2[lay.total]
    std::vector<double> host;   // host mirror (interleaved re, im), valid when host_valid
    bool host_valid = false;

    double2* d() const { return buf.as<double2>(); }

    static std::unique_ptr<Tensor> create(const std::shared_ptr<Device>& dev, Layout lay,
                                          bool zero = true) {
        std::unique_ptr<Tensor> t(new Tensor);
        t->dev = dev;
        t->lay = std::move(lay);
        t->buf = DevBuf(dev->pool.get(), (size_t)t->lay.total * sizeof(double2));
        if (zero && t->lay.total) launch_zero(dev->stream, t->buf.get(), t->buf.bytes());
        return t;
    }
    void upload(const double* data) {
        if (lay.total)
            KV_HIP(hipMemcpyAsync(buf.get(), data, (size_t)lay.total * sizeof(double2),
                                  hipMemcpyHostToDevice, dev->stream));
        host.assign(data, data + 2 * (size_t)lay.total);
        host_valid = true;
        dev->sync();
    }
    void download(double* data) const {
        if (lay.total)
            KV_HIP(hipMemcpyAsync(data, buf.get(), (size_t)lay.total * sizeof(double2),
                                  hipMemcpyDeviceToHost, dev->stream));
        dev->sync();
    }
    // Host copy of the payload (downloads once; MPO tensors keep theirs from upload).
    const std::vector<double>& host_data() {
        if (!host_valid) {
            host.resize(2 * (size_t)lay.total);
            download(host.data());
            host_valid = true;
        }
        return host;
    }
    void invalidate_host() { host_valid = false; }
    std::unique_ptr<Tensor> clone() const {
        auto t = create(dev, lay, false);
        if (lay.total)
            KV_HIP(hipMemcpyAsync(t->buf.get(), buf.get(), (size_t)lay.total * sizeof(double2),
                                  hipMemcpyDeviceToDevice, dev->stream));
        if (host_valid) {
            t->host = host;
            t->host_valid = true;
        }
        return t;
    }
};

// Host-to-device upload of a plan's descriptor array into a fresh device buffer.  The
// copy is ordered on the context stream (the pooled buffer may still be read by earlier
// work) and staged through the context's pinned ring, so the host vector may be freed at
// once and the host does not wait.
template <class T>
inline DevBuf upload_vec(Device& dev, const std::vector<T>& v) {
    DevBuf b(dev.pool.get(), v.size() * sizeof(T));
    dev.upload(b.get(), v.data(), v.size() * sizeof(T));
    return b;
}

// Several host arrays in one device buffer by one upload (each array 256-byte aligned),
// so a launch's descriptor lists cost one copy instead of one each.
struct PackedUpload {
    DevBuf buf;
    std::vector<size_t> off;
    template <class T>
    T* at(int i) const {
        return reinterpret_cast<T*>(static_cast<char*>(buf.get()) + off[i]);
    }
};

template <class... V>
inline PackedUpload upload_packed_on(Device& dev, hipStream_t on, const V&... vs) {
    PackedUpload p;
    size_t total = 0;
    auto place = [&](const auto& v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        const size_t o = (total + 255) / 256 * 256;
        p.off.push_back(o);
        total = o + v.size() * sizeof(T);
    };
    (place(vs), ...);
    std::vector<char> h(total);
    int i = 0;
    auto fill = [&](const auto& v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (!v.empty()) std::memcpy(h.data() + p.off[i], v.data(), v.size() * sizeof(T));
        ++i;
    };
    (fill(vs), ...);
    p.buf = DevBuf(dev.pool.get(), total);
    dev.upload(p.buf.get(), h.data(), total, on);
    return p;
}

template <class... V>
inline PackedUpload upload_packed(Device& dev, const V&... vs) {
    return upload_packed_on(dev, nullptr, vs...);
}

}  // namespace kv
