// Minimal RAII device buffer and the host-side launch helpers shared by the .hip files.
#pragma once
#include "common.hpp"

#include <utility>
#include <vector>

namespace rtcd {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool view = false;  // bound to memory it does not own (bind): never allocates or frees
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), view(o.view) { o.p = nullptr; o.n = 0; o.view = false; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; view = o.view;
            o.p = nullptr; o.n = 0; o.view = false;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    // Makes this buffer a view of count elements at ptr (memory owned elsewhere).
    void bind(T* ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        view = true;
    }
    void alloc(size_t count) {
        if (view) throw std::runtime_error("a buffer bound to fixed memory cannot grow");
        release();
        if (count) RTCD_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    // Grow (never shrink); contents are not preserved.
    void reserve(size_t count) {
        if (count > n) alloc(count + count / 4);
    }
    void release() {
        if (p && !view) (void) hipFree(p);
        p = nullptr;
        n = 0;
        view = false;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        reserve(count);
        if (count) RTCD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    std::vector<T> download(size_t count, hipStream_t s) const {
        std::vector<T> h(count);
        if (count) RTCD_HIP(hipMemcpyAsync(h.data(), p, count * sizeof(T), hipMemcpyDeviceToHost, s));
        RTCD_HIP(hipStreamSynchronize(s));
        return h;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Pinned, coherent host memory that kernels read and write directly (h on the host, d on the
// device).  Small batches go through it: the pose kernel reads the caller's poses from it and the
// last kernel writes the status words (and small flag arrays) into it, so that a call needs no copy
// and one stream synchronisation.  Grows, never shrinks; contents are not preserved.
template <typename T>
struct HostMapped {
    T* h = nullptr;
    T* d = nullptr;
    size_t n = 0;
    HostMapped() = default;
    HostMapped(const HostMapped&) = delete;
    HostMapped& operator=(const HostMapped&) = delete;
    ~HostMapped() { release(); }
    // Device memory standing in for the pinned buffer (h stays null): kernels write there and the
    // host copies what it needs.  Such a buffer never allocates.
    void bind_device(T* ptr, size_t count) {
        release();
        d = ptr;
        n = count;
    }
    void reserve(size_t count) {
        if (count <= n) return;
        if (d && !h) throw std::runtime_error("a buffer bound to fixed memory cannot grow");
        release();
        const size_t c = count + count / 4 + 16;
        RTCD_HIP(hipHostMalloc((void**) &h, c * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
        void* dp = nullptr;
        const hipError_t e = hipHostGetDevicePointer(&dp, h, 0);
        if (e != hipSuccess) {
            (void) hipHostFree(h);
            h = nullptr;
            throw HipError(e, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
        }
        d = (T*) dp;
        n = c;
    }
    void release() {
        if (h) (void) hipHostFree(h);
        h = d = nullptr;
        n = 0;
    }
};

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned) ((n + bs - 1) / bs); }

// ---- scan.hip ----
// Exclusive prefix sum of n uint32 values.  tmp must hold scan_tmp_words(n) words.  If total is
// non-null the grand total is written to *total (device).  Safe under stream capture (no memset).
size_t scan_tmp_words(size_t n);
void exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, uint32_t* tmp, hipStream_t s);
// out[i] = value (fill kernel; used instead of hipMemsetAsync so that captured graphs stay valid).
void fill_u32(uint32_t* out, size_t n, uint32_t value, hipStream_t s);

// ---- lbvh.hip ----
struct BvhOut {
    DevBuf<BNode> nodes;
    DevBuf<int32_t> perm;  // leaf order -> original primitive index
    int n_nodes = 0;
    int depth = 0;         // max number of inner nodes on a root-to-leaf path
};
// boxes: 2 float4 per primitive (lo, hi).  Builds on the device: Morton codes, LDS radix sort,
// Karras hierarchy, bottom-up boxes, leaf collapse to at most leaf_size primitives.
void build_lbvh(const float4* boxes, int n, int leaf_size, hipStream_t s, BvhOut& out);

}  // namespace rtcd
