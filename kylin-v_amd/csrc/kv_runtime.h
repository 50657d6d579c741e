// Device runtime for the Holstein TDVP engine: errors, a caching device allocator,
// the per-context stream and HIP-event timers.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kv_error.h"

namespace kv {

#define KV_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::kv::Error(::kv::ERR_HIP, std::string(#call) + ": " +               \
                                                 hipGetErrorString(e_));               \
    } while (0)

// Caching allocator.  hipMalloc/hipFree synchronise the device and cost tens of
// microseconds, so buffers of the sweep (Krylov vectors, intermediates, environment
// slabs) are recycled by size class.  Size classes: 256 B granularity below 1 MiB,
// 1 MiB granularity above.
class Pool {
  public:
    explicit Pool(int device) : device_(device) {
        if (const char* c = std::getenv("KV_POOL_CACHE_GB")) cap_ = (size_t)(std::atof(c) * (1ull << 30));
    }
    ~Pool() { release_all(); }
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;

    // Size classes: 256-byte steps below 64 KiB, then four classes per power of two
    // (waste at most 25%).  Bond dimensions change from bond to bond, so exact-size classes
    // would keep every shape ever seen cached; geometric classes let neighbouring shapes
    // share buffers.
    static size_t size_class(size_t bytes) {
        if (bytes == 0) bytes = 1;
        if (bytes < (size_t(1) << 16)) return (bytes + 255) / 256 * 256;
        int k = 63 - __builtin_clzll((unsigned long long)bytes);
        const size_t step = size_t(1) << (k - 2);
        return (bytes + step - 1) / step * step;
    }
    void* get(size_t bytes) {
        const size_t sz = size_class(bytes);
        auto it = free_.find(sz);
        if (it != free_.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            cached_ -= sz;
            return p;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            // release cached blocks and retry once
            (void)hipGetLastError();
            release_cached();
            KV_HIP(hipMalloc(&p, sz));
        }
        live_ += sz;
        return p;
    }
    void put(void* p, size_t bytes) {
        if (!p) return;
        const size_t sz = size_class(bytes);
        free_[sz].push_back(p);
        cached_ += sz;
        // bounded cache: runs whose tensor sizes keep changing (chains of different shape
        // on one context) would otherwise pin device memory in ever new size classes
        if (cached_ > cap_) trim(cap_ / 2);
    }
    // Free cached buffers, largest first, until at most keep bytes stay cached.
    // hipFree synchronises the device, so freed buffers are no longer in use.
    void trim(size_t keep) {
        for (auto it = free_.rbegin(); it != free_.rend() && cached_ > keep; ++it) {
            while (!it->second.empty() && cached_ > keep) {
                (void)hipFree(it->second.back());
                it->second.pop_back();
                cached_ -= it->first;
                live_ -= it->first;
            }
        }
    }
    void release_cached() {
        for (auto& kvp : free_) {
            for (void* p : kvp.second) {
                (void)hipFree(p);
                live_ -= kvp.first;
            }
            kvp.second.clear();
        }
        cached_ = 0;
    }
    void release_all() { release_cached(); }
    size_t live_bytes() const { return live_; }

  private:
    int device_;
    std::map<size_t, std::vector<void*>> free_;
    size_t live_ = 0, cached_ = 0;
    size_t cap_ = (size_t)24 << 30;  // KV_POOL_CACHE_GB overrides
};

// Owning device buffer returned to the pool on destruction.
class DevBuf {
  public:
    DevBuf() = default;
    DevBuf(Pool* pool, size_t bytes) : pool_(pool), bytes_(bytes) {
        ptr_ = bytes ? pool->get(bytes) : nullptr;
    }
    ~DevBuf() { reset(); }
    DevBuf(DevBuf&& o) noexcept : pool_(o.pool_), ptr_(o.ptr_), bytes_(o.bytes_) {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            pool_ = o.pool_;
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void reset() {
        if (ptr_ && pool_) pool_->put(ptr_, bytes_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    void* get() const { return ptr_; }
    template <class T> T* as() const { return static_cast<T*>(ptr_); }
    size_t bytes() const { return bytes_; }

  private:
    Pool* pool_ = nullptr;
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Counters accumulated by the contraction executors.
struct KernelStats {
    double gemm_flop = 0, gemm_ms = 0, gemm_bytes = 0;  // every GEMM launch
    double wfold_flop = 0, wfold_ms = 0;
    long gemm_launches = 0, wfold_launches = 0;
    // the FP32-operand launches alone (FP32 and split-BF16 kernels)
    double gemm32_flop = 0, gemm32_ms = 0, gemm32_bytes = 0;
    long gemm32_launches = 0;
};

// The Householder tridiagonalisations of the bond SVD (hetrd_lower): HIP-event time, the
// Hermitian trailing triangle each column's matvec must read (16 B per element), columns.
struct TrdStats {
    double ms = 0, bytes = 0;
    long columns = 0;
};

// The bond SVD's block subspace route: HIP-event time of its GEMM launches with their
// algorithmic flops (8 m n k) and bytes (operands and result once each, 16 B per complex),
// of its Cholesky QR kernels (chol_upper + tri_inv) and of its Rayleigh-Ritz eigensolves;
// sectors solved by the route and sectors redone by the full eigensolver.
struct BlockStats {
    double gemm_ms = 0, gemm_flop = 0, gemm_bytes = 0;
    long gemm_launches = 0;
    double orth_ms = 0, eig_ms = 0;
    long orth_calls = 0, eig_calls = 0, sectors = 0, fallbacks = 0;
};

// HIP-event spans on the context stream whose durations are added to their counters when
// the stream is next synchronised, so timing a launch does not make the host wait for it.
// Counters must outlive the pending spans (fields of the Device).
class SpanLog {
  public:
    SpanLog() = default;
    SpanLog(const SpanLog&) = delete;
    SpanLog& operator=(const SpanLog&) = delete;
    ~SpanLog() {
        for (auto& p : pending_) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (hipEvent_t e : spare_) (void)hipEventDestroy(e);
    }
    hipEvent_t mark(hipStream_t s) {
        hipEvent_t e;
        if (!spare_.empty()) {
            e = spare_.back();
            spare_.pop_back();
        } else {
            KV_HIP(hipEventCreate(&e));
        }
        KV_HIP(hipEventRecord(e, s));
        return e;
    }
    // closes the span opened by mark(); its milliseconds go to *ms_counter
    void close(hipEvent_t a, hipStream_t s, double* ms_counter) { pending_.push_back({a, mark(s), ms_counter}); }
    // call only after the stream has been synchronised
    void resolve() {
        for (auto& p : pending_) {
            float ms = 0;
            KV_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            *p.dst += ms;
            spare_.push_back(p.a);
            spare_.push_back(p.b);
        }
        pending_.clear();
    }

  private:
    struct Span {
        hipEvent_t a, b;
        double* dst;
    };
    std::vector<Span> pending_;
    std::vector<hipEvent_t> spare_;
};

struct Device {
    int id = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<Pool> pool;
    SpanLog spans;
    bool timing = false;  // record HIP events around contraction kernels
    KernelStats stats;
    TrdStats trd;
    BlockStats blk;
    // GEMM tile autotuner: (precision, opA, opB, shape class) -> tile config
    bool autotune = true;
    std::map<uint64_t, int> gemm_tuned;
    explicit Device(int dev) : id(dev) {
        KV_HIP(hipSetDevice(dev));
        KV_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        pool.reset(new Pool(dev));
        if (const char* a = std::getenv("KV_AUTOTUNE")) autotune = a[0] != '0';
        KV_HIP(hipHostMalloc(&ring_, kRingBytes, hipHostMallocDefault));
    }
    ~Device() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            pool.reset();
            (void)hipStreamDestroy(stream);
        }
        if (ring_) (void)hipHostFree(ring_);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    // Waits for the stream; then the upload ring is free again and pending spans resolve.
    void sync() {
        KV_HIP(hipStreamSynchronize(stream));
        ring_off_ = 0;
        spans.resolve();
    }
    // Host-to-device copy of a small host array (work descriptors) that returns without
    // waiting: the bytes go through a pinned staging ring, whose slots are reused only after
    // the next sync().  Copies larger than a quarter of the ring wait for themselves.
    // `on`: the stream the copy is ordered on (default the context stream); work on another
    // stream must be joined back into the context stream before its next sync(), which
    // restarts the ring.
    void upload(void* dst, const void* src, size_t bytes, hipStream_t on = nullptr) {
        if (!bytes) return;
        if (!on) on = stream;
        if (bytes > kRingBytes / 4) {
            KV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, on));
            KV_HIP(hipStreamSynchronize(on));
            return;
        }
        if (ring_off_ + bytes > kRingBytes) sync();
        char* slot = static_cast<char*>(ring_) + ring_off_;
        std::memcpy(slot, src, bytes);
        KV_HIP(hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, on));
        ring_off_ = (ring_off_ + bytes + 255) / 256 * 256;
    }

  private:
    static constexpr size_t kRingBytes = (size_t)8 << 20;
    void* ring_ = nullptr;
    size_t ring_off_ = 0;
};

// Times a region of stream work with a pair of HIP events recorded on the stream the
// kernels run on.  ms() synchronises on the end event.
class EventTimer {
  public:
    EventTimer() {
        KV_HIP(hipEventCreate(&a_));
        KV_HIP(hipEventCreate(&b_));
    }
    ~EventTimer() {
        (void)hipEventDestroy(a_);
        (void)hipEventDestroy(b_);
    }
    void start(hipStream_t s) { KV_HIP(hipEventRecord(a_, s)); }
    void stop(hipStream_t s) { KV_HIP(hipEventRecord(b_, s)); }
    float ms() {
        KV_HIP(hipEventSynchronize(b_));
        float t = 0;
        KV_HIP(hipEventElapsedTime(&t, a_, b_));
        return t;
    }

  private:
    hipEvent_t a_{}, b_{};
};

}  // namespace kv
