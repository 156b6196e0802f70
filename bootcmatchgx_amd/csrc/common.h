// common.h -- shared host/device infrastructure of the MI355X BootCMatchGX path.
//
// Errors never exit(): every failing HIP call throws bcmgx::Error, which the
// C ABI layer turns into a status code plus a message stored on the context
// (the reference exits from CHECK_DEVICE, src/utility/cudamacro.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bcmgx.h"

namespace bcmgx {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define BCMGX_HIP(call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw ::bcmgx::Error(BCMGX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                    " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define BCMGX_CHECK(cond, code, msg)                 \
    do {                                             \
        if (!(cond)) throw ::bcmgx::Error(code, msg); \
    } while (0)

// Kernel launch check (launch errors surface on the next API call otherwise).
#define BCMGX_LAUNCHED() BCMGX_HIP(hipGetLastError())

constexpr int WG = 256;  // default workgroup size (4 wave64s)

inline unsigned grid_for(long long items, int per_block) { return (unsigned)((items + per_block - 1) / per_block); }

// ---------------------------------------------------------------------------
// Caching device allocator: hipMalloc'ed blocks are recycled by exact size
// class (powers of two above 1 MiB granularity), so the setup's many
// temporaries and repeated solves do not pay hipMalloc/hipFree.
// ---------------------------------------------------------------------------
class DevicePool {
public:
    explicit DevicePool(int device) : device_(device) {}
    ~DevicePool() { release_all(); }
    void* alloc(size_t bytes);
    void free(void* p);
    void release_all();
    size_t bytes_in_use() const { return in_use_; }
    // Blocks are reused in the order of one stream.  While work is queued on a
    // second stream, the blocks that work frees are held back (set_defer(true))
    // until the first stream has been made to wait for it (release_deferred()).
    void set_defer(bool d) { defer_ = d; }
    void release_deferred();

private:
    static size_t size_class(size_t b);
    int device_;
    bool defer_ = false;
    std::vector<std::pair<size_t, void*>> deferred_;
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> live_;
    size_t in_use_ = 0;
};

struct Context;

// Typed device buffer owned by a pool (RAII).
template <typename T>
class DBuf {
public:
    DBuf() = default;
    DBuf(Context* c, size_t n) { reset(c, n); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept { swap(o); }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); swap(o); }
        return *this;
    }
    ~DBuf() { release(); }
    void reset(Context* c, size_t n);
    void release();
    T* get() const { return p_; }
    size_t size() const { return n_; }
    explicit operator bool() const { return p_ != nullptr; }
    void swap(DBuf& o) noexcept {
        std::swap(ctx_, o.ctx_);
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }

private:
    Context* ctx_ = nullptr;
    T* p_ = nullptr;
    size_t n_ = 0;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // a second stream (default priority) for setup work that overlaps, made on first use
    std::unique_ptr<DevicePool> pool;
    int num_cus = 256;
    int num_xcds = 8;
    std::string last_error;
    // pinned host staging for small device->host scalar reads
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // autotune cache: matrix signature -> lanes per row
    std::map<std::string, int> tuned_lanes;
    // this device's copy of the fixed scratch array of the allocation-free entry point
    void* scratch = nullptr;

    explicit Context(int dev);
    ~Context();
    void sync() { BCMGX_HIP(hipStreamSynchronize(stream)); }
    hipStream_t side_stream();
    template <typename T>
    T* dalloc(size_t n) { return static_cast<T*>(pool->alloc(n * sizeof(T))); }
    void dfree(void* p) { pool->free(p); }
    void* host_scratch(size_t bytes);
    // copy n elements device->host through pinned memory and wait
    template <typename T>
    void d2h(T* dst, const T* src, size_t n);
    // two small device->host reads with one synchronisation
    template <typename A, typename B>
    void d2h2(A* a, const A* da, size_t na, B* b, const B* db, size_t nb) {
        size_t oa = na * sizeof(A), ob = (oa + 15) / 16 * 16;
        char* h = (char*)host_scratch(ob + nb * sizeof(B));
        BCMGX_HIP(hipMemcpyAsync(h, da, oa, hipMemcpyDeviceToHost, stream));
        BCMGX_HIP(hipMemcpyAsync(h + ob, db, nb * sizeof(B), hipMemcpyDeviceToHost, stream));
        sync();
        std::memcpy(a, h, oa);
        std::memcpy(b, h + ob, nb * sizeof(B));
    }
    template <typename T>
    T read1(const T* src) {
        T v;
        d2h(&v, src, 1);
        return v;
    }
    template <typename T>
    void h2d(T* dst, const T* src, size_t n) {
        BCMGX_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, stream));
        sync();
    }
};

template <typename T>
void DBuf<T>::reset(Context* c, size_t n) {
    release();
    ctx_ = c;
    n_ = n;
    p_ = n ? c->dalloc<T>(n) : nullptr;
}
template <typename T>
void DBuf<T>::release() {
    if (p_ && ctx_) ctx_->dfree(p_);
    p_ = nullptr;
    n_ = 0;
}

template <typename T>
void Context::d2h(T* dst, const T* src, size_t n) {
    size_t b = n * sizeof(T);
    void* h = host_scratch(b);
    BCMGX_HIP(hipMemcpyAsync(h, src, b, hipMemcpyDeviceToHost, stream));
    sync();
    std::memcpy(dst, h, b);
}

// ---------------------------------------------------------------------------
// Device CSR: int32 row pointers/columns, FP64 values, optional FP32 copy.
// ---------------------------------------------------------------------------
// Row-tile index of the pattern (built on demand, kept with the matrix): for
// every tile of TILE_ROWS consecutive rows the sorted distinct columns its
// entries touch (tucol[tu[t] .. tu[t+1])) and, per entry, the 16-bit position
// of its column in that list (tlidx).  tu[t+1] == tu[t] for a tile with
// entries marks it unindexed (too many entries or columns): it is read by
// column index.
constexpr int TILE_ROWS = 64;
constexpr int TILE_MAX_ENTRIES = 2048;
constexpr int TILE_MAX_COLS = 1024;

struct DCsr {
    int n = 0, m = 0, nnz = 0;
    DBuf<int> row, col;
    DBuf<double> val;
    DBuf<float> valf;  // FP32 copy for the mixed-precision cycle (may be empty)
    DBuf<int> tu, tucol;            // row-tile index (empty until built)
    DBuf<unsigned short> tlidx;
    bool unit_rows = false;         // row i holds exactly entry i (row pointers 0, 1, .., n): a prolongator
    int maxrow = -1;                // entries in the longest row (-1: not known)
    bool tiled() const { return (bool)tu; }
    void alloc(Context* c, int n_, int m_, int nnz_) {
        n = n_; m = m_; nnz = nnz_;
        unit_rows = false;
        maxrow = -1;
        row.reset(c, (size_t)n + 1);
        col.reset(c, (size_t)(nnz > 0 ? nnz : 1));
        val.reset(c, (size_t)(nnz > 0 ? nnz : 1));
    }
};

}  // namespace bcmgx

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
#ifdef __HIPCC__
namespace bcmgx {
// Workgroups are dispatched round-robin over the 8 XCDs (WG b -> XCD b % 8),
// each with its own L2.  Remap so XCD x owns one contiguous range of logical
// blocks: rows that share x-vector gathers (neighbouring planes of a stencil)
// then hit the same L2.  Bijective on [0, G).
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int X = 8;
    int xcd = b % X, local = b / X;
    int per = G / X, rem = G % X;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}
// The other order: super-blocks of sb consecutive logical blocks dealt
// round-robin to the XCDs, so all XCDs work on one band of rows at a time
// (sb <= 0: xcd_remap).  Bijective on [0, G); the blocks of an incomplete last
// round keep their own index.
__device__ __forceinline__ int xcd_order(int b, int G, int sb) {
    if (sb <= 0) return xcd_remap(b, G);
    const int X = 8;
    int full = G / (X * sb) * (X * sb);
    if (b >= full) return b;
    int xcd = b % X, local = b / X;
    return ((local / sb) * X + xcd) * sb + local % sb;
}
template <typename T>
__device__ __forceinline__ T ldg_nt(const T* p) { return __builtin_nontemporal_load(p); }
}  // namespace bcmgx
#endif
