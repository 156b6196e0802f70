// prims.hip -- context, caching allocator and device primitives (scan,
// fill, conversions).  The reference gets these from CUB/Thrust
// (src/utility/devicePrefixSum.h, deviceMap.h, cuCompactor.cuh); here they are
// written for wave64 directly.
#include "kernels.h"

#include <algorithm>
#include <cstdio>

namespace bcmgx {

// ------------------------------------------------------------------ pool ---
size_t DevicePool::size_class(size_t b) {
    if (b == 0) b = 1;
    if (b <= (1u << 20)) {  // small: round to 4 KiB
        return (b + 4095) & ~size_t(4095);
    }
    size_t c = size_t(1) << 20;
    while (c < b) c <<= 1;
    if (c - b > (c >> 2)) {  // large and far below the power of two: 1 MiB granularity
        return (b + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
    }
    return c;
}

void* DevicePool::alloc(size_t bytes) {
    size_t c = size_class(bytes);
    auto it = free_.find(c);
    void* p = nullptr;
    if (it != free_.end()) {
        p = it->second;
        free_.erase(it);
    } else {
        (void)hipSetDevice(device_);  // blocks live on the pool's device
        hipError_t e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            // drop the cached blocks and retry once
            (void)hipGetLastError();
            for (auto& f : free_) (void)hipFree(f.second);
            free_.clear();
            e = hipMalloc(&p, c);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(BCMGX_ERR_ALLOC, "hipMalloc of " + std::to_string(c) + " bytes failed");
            }
        }
    }
    live_[p] = c;
    in_use_ += c;
    return p;
}

void DevicePool::free(void* p) {
    if (!p) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    in_use_ -= it->second;
    if (defer_) deferred_.push_back({it->second, p});
    else free_.insert({it->second, p});
    live_.erase(it);
}

void DevicePool::release_deferred() {
    for (auto& d : deferred_) free_.insert(d);
    deferred_.clear();
}

void DevicePool::release_all() {
    release_deferred();
    if (!free_.empty() || !live_.empty()) (void)hipSetDevice(device_);
    for (auto& f : free_) (void)hipFree(f.second);
    free_.clear();
    for (auto& l : live_) (void)hipFree(l.first);
    live_.clear();
    in_use_ = 0;
}

// --------------------------------------------------------------- context ---
Context::Context(int dev) : device(dev) {
    BCMGX_HIP(hipSetDevice(dev));
    BCMGX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    pool.reset(new DevicePool(dev));
    hipDeviceProp_t prop;
    BCMGX_HIP(hipGetDeviceProperties(&prop, dev));
    num_cus = prop.multiProcessorCount;
    scratch = scratch_address(dev);
}

Context::~Context() {
    if (side) (void)hipStreamSynchronize(side);
    if (stream) (void)hipStreamSynchronize(stream);
    pool.reset();
    if (pinned) (void)hipHostFree(pinned);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
}

hipStream_t Context::side_stream() {
    if (!side) BCMGX_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    return side;
}

void* Context::host_scratch(size_t bytes) {
    if (bytes > pinned_bytes) {
        if (pinned) BCMGX_HIP(hipHostFree(pinned));
        pinned = nullptr;
        size_t b = std::max<size_t>(bytes, 1 << 16);
        BCMGX_HIP(hipHostMalloc(&pinned, b, hipHostMallocDefault));
        pinned_bytes = b;
    }
    return pinned;
}

// ------------------------------------------------------------------ scan ---
// Three-phase exclusive scan: per-block totals, scan of the totals in one
// workgroup, then per-block rescan with the block offset.  1024 items per
// block (256 threads x 4).
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_BLOCK = WG * SCAN_ITEMS;

template <typename T>
__device__ T block_exclusive_scan(T v, T* lds_wave, T& total) {
    // wave64 inclusive scan with shuffles, then across the 4 waves through LDS
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds_wave[wid] = x;
    __syncthreads();
    T off = 0;
    for (int w = 0; w < wid; w++) off += lds_wave[w];
    total = lds_wave[0] + lds_wave[1] + lds_wave[2] + lds_wave[3];
    __syncthreads();
    return off + x - v;
}

template <typename TI, typename T>
__global__ __launch_bounds__(WG) void k_scan_partials(const TI* __restrict__ in, long long n, T* __restrict__ partials) {
    __shared__ T lw[4];
    long long base = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    T s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) s += (T)in[base + k];
    T tot;
    block_exclusive_scan<T>(s, lw, tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(WG) void k_scan_single(T* __restrict__ data, long long n, T* __restrict__ total_out) {
    __shared__ T lw[4];
    T carry = 0;
    for (long long base = 0; base < n; base += SCAN_BLOCK) {
        long long b = base + threadIdx.x * SCAN_ITEMS;
        T v[SCAN_ITEMS];
        T s = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            v[k] = (b + k < n) ? data[b + k] : (T)0;
            s += v[k];
        }
        T tot;
        T ex = block_exclusive_scan<T>(s, lw, tot);
        T run = carry + ex;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            if (b + k < n) data[b + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <typename TI, typename T>
__global__ __launch_bounds__(WG) void k_scan_final(const TI* __restrict__ in, long long n, const T* __restrict__ offs, T* __restrict__ out) {
    __shared__ T lw[4];
    long long base = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? (T)in[base + k] : (T)0;
        s += v[k];
    }
    T tot;
    T run = offs[blockIdx.x] + block_exclusive_scan<T>(s, lw, tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

template <typename TI, typename T>
static T scan_impl(Context* c, const TI* in, T* out, long long n, bool read_total = true) {
    // out[0..n) exclusive prefix sums, out[n] = total; returns the total
    // (read back only when asked: the other callers leave it on the device).
    if (n <= 0) {
        fill<T>(c, out, 1, (T)0);
        return 0;
    }
    long long nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    DBuf<T> part(c, nb + 1);
    k_scan_partials<TI, T><<<(unsigned)nb, WG, 0, c->stream>>>(in, n, part.get());
    BCMGX_LAUNCHED();
    k_scan_single<T><<<1, WG, 0, c->stream>>>(part.get(), nb, out + n);
    BCMGX_LAUNCHED();
    k_scan_final<TI, T><<<(unsigned)nb, WG, 0, c->stream>>>(in, n, part.get(), out);
    BCMGX_LAUNCHED();
    return read_total ? c->read1(out + n) : (T)0;
}

int scan_exclusive(Context* c, const int* in, int* out, long long n) { return scan_impl<int, int>(c, in, out, n); }
void scan_exclusive_async(Context* c, const int* in, int* out, long long n) { scan_impl<int, int>(c, in, out, n, false); }
void scan_exclusive_async(Context* c, const long long* in, long long* out, long long n) {
    scan_impl<long long, long long>(c, in, out, n, false);
}
long long scan_exclusive(Context* c, const long long* in, long long* out, long long n) {
    return scan_impl<long long, long long>(c, in, out, n);
}
long long scan_exclusive_i32_to_i64(Context* c, const int* in, long long* out, long long n) {
    return scan_impl<int, long long>(c, in, out, n);
}

// ------------------------------------------------------------ elementwise ---
template <typename T>
__global__ void k_fill(T* __restrict__ p, long long n, T v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
template <typename T>
void fill(Context* c, T* p, long long n, T v) {
    if (n <= 0) return;
    unsigned g = std::min<long long>(grid_for(n, WG), 65536);
    k_fill<T><<<g, WG, 0, c->stream>>>(p, n, v);
    BCMGX_LAUNCHED();
}
template void fill<int>(Context*, int*, long long, int);
template void fill<double>(Context*, double*, long long, double);
template void fill<float>(Context*, float*, long long, float);
template void fill<long long>(Context*, long long*, long long, long long);

template <typename A, typename B>
__global__ void k_convert(const A* __restrict__ s, B* __restrict__ d, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = (B)s[i];
}
template <typename A, typename B>
void convert(Context* c, const A* s, B* d, long long n) {
    if (n <= 0) return;
    unsigned g = std::min<long long>(grid_for(n, WG), 65536);
    k_convert<A, B><<<g, WG, 0, c->stream>>>(s, d, n);
    BCMGX_LAUNCHED();
}
template void convert<double, float>(Context*, const double*, float*, long long);
template void convert<float, double>(Context*, const float*, double*, long long);
template void convert<double, double>(Context*, const double*, double*, long long);

__global__ __launch_bounds__(WG) void k_boot_random(double* __restrict__ x, long long n, unsigned long long seed) {
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n; i += (long long)gridDim.x * WG) {
        unsigned long long z = seed + (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        double u01 = (double)(z >> 11) * 0x1.0p-53;
        x[i] = 2.0 * u01 - 1.0;
    }
}

void boot_random(Context* c, double* x, long long n, unsigned long long seed) {
    if (n <= 0) return;
    k_boot_random<<<grid_for(n, WG), WG, 0, c->stream>>>(x, n, seed);
    BCMGX_LAUNCHED();
}

// block maxima of |x|, then one block takes their maximum and scales x by it in
// a second launch (max is exact in any order, so the result is the oracle's)
__global__ __launch_bounds__(WG) void k_absmax_partial(const double* __restrict__ x, long long n, double* __restrict__ part) {
    __shared__ double s[WG / 64];
    double m = 0.0;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n; i += (long long)gridDim.x * WG) m = fmax(m, fabs(x[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WG / 64; w++) m = fmax(m, s[w]);
        part[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(WG) void k_scale_by_max(double* __restrict__ x, long long n, const double* __restrict__ part, int nb) {
    double m = 0.0;
    for (int b = 0; b < nb; b++) m = fmax(m, part[b]);
    if (!(m > 0.0)) return;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n; i += (long long)gridDim.x * WG) x[i] = x[i] / m;
}

void scale_by_absmax(Context* c, double* x, long long n) {
    if (n <= 0) return;
    int nb = (int)std::min<long long>(1024, (n + WG - 1) / WG);
    DBuf<double> part(c, nb);
    k_absmax_partial<<<nb, WG, 0, c->stream>>>(x, n, part.get());
    BCMGX_LAUNCHED();
    k_scale_by_max<<<nb, WG, 0, c->stream>>>(x, n, part.get(), nb);
    BCMGX_LAUNCHED();
}

__global__ __launch_bounds__(WG) void k_add_to(double* __restrict__ y, const double* __restrict__ x, long long n,
                                               const int* __restrict__ stop) {
    if (stop && *stop) return;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n; i += (long long)gridDim.x * WG) y[i] = y[i] + x[i];
}

void add_to(Context* c, double* y, const double* x, long long n, const int* stop) {
    if (n <= 0) return;
    k_add_to<<<grid_for(n, WG), WG, 0, c->stream>>>(y, x, n, stop);
    BCMGX_LAUNCHED();
}

void copy_d2d(Context* c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    BCMGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
}

// column-major (ld) <-> interleaved row-major [n][k], through an LDS tile of
// TR rows x k columns (row stride k + 1, so neither side's LDS accesses meet
// the same bank), so that the column-major side and the interleaved side are
// both read and written with consecutive lanes on consecutive addresses.
constexpr int TILE_ELEMS = 2048;
static int tile_rows(int k) { return std::max(1, TILE_ELEMS / k); }
static size_t tile_bytes(int k) { return (size_t)tile_rows(k) * (k + 1) * sizeof(double); }

// zero_flag (may be null) is cleared when any element read is nonzero; the
// block goes to alt instead of d when alt_if is set and nonzero
__global__ __launch_bounds__(WG) void k_cm_to_rm(const double* __restrict__ s, int ld, double* __restrict__ d, int n,
                                                 int k, int TR, int* __restrict__ zero_flag, const int* __restrict__ alt_if,
                                                 double* __restrict__ alt) {
    extern __shared__ double tile[];
    if (alt_if && *alt_if) d = alt;
    long long i0 = (long long)blockIdx.x * TR;
    int rows = (int)std::min<long long>(TR, n - i0);
    bool nz = false;
    for (int idx = threadIdx.x; idx < TR * k; idx += WG) {
        int r = idx % TR, c = idx / TR;
        if (r < rows) {
            double v = s[(long long)c * ld + i0 + r];
            nz |= v != 0.0;
            tile[r * (k + 1) + c] = v;
        }
    }
    if (zero_flag && nz) *zero_flag = 0;
    __syncthreads();
    for (int idx = threadIdx.x; idx < rows * k; idx += WG) {
        int r = idx / k, c = idx % k;
        d[(i0 + r) * k + c] = tile[r * (k + 1) + c];
    }
}

// st/p (may be null): columns whose FCG step waits (pad1) get x + pend * p on the way out
__global__ __launch_bounds__(WG) void k_rm_to_cm(const double* __restrict__ s, double* __restrict__ d, int ld, int n,
                                                 int k, int TR, const FcgState* __restrict__ st,
                                                 const double* __restrict__ p) {
    extern __shared__ double tile[];
    __shared__ double pend[64];
    __shared__ int waits[64];
    if (threadIdx.x < k) {
        waits[threadIdx.x] = st ? st[threadIdx.x].pad1 : 0;
        pend[threadIdx.x] = st ? st[threadIdx.x].pend : 0.0;
    }
    __syncthreads();
    long long i0 = (long long)blockIdx.x * TR;
    int rows = (int)std::min<long long>(TR, n - i0);
    for (int idx = threadIdx.x; idx < rows * k; idx += WG) {
        int r = idx / k, c = idx % k;
        long long o = (i0 + r) * k + c;
        double v = s[o];
        if (waits[c]) v = pend[c] * p[o] + v;  // the flush of k_flush_pending
        tile[r * (k + 1) + c] = v;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < TR * k; idx += WG) {
        int r = idx % TR, c = idx / TR;
        if (r < rows) d[(long long)c * ld + i0 + r] = tile[r * (k + 1) + c];
    }
}

void colmajor_to_interleaved(Context* c, const double* s, int ld, double* d, int n, int k, int* zero_flag,
                             const int* alt_if, double* alt) {
    if ((long long)n * k == 0) return;
    int TR = tile_rows(k);
    k_cm_to_rm<<<(unsigned)((n + TR - 1) / TR), WG, tile_bytes(k), c->stream>>>(s, ld, d, n, k, TR, zero_flag, alt_if,
                                                                                  alt);
    BCMGX_LAUNCHED();
}
void interleaved_to_colmajor(Context* c, const double* s, double* d, int ld, int n, int k, const FcgState* st,
                             const double* p) {
    if ((long long)n * k == 0) return;
    BCMGX_CHECK(k <= 64, BCMGX_ERR_ARG, "at most 64 columns");
    int TR = tile_rows(k);
    k_rm_to_cm<<<(unsigned)((n + TR - 1) / TR), WG, tile_bytes(k), c->stream>>>(s, d, ld, n, k, TR, st, p);
    BCMGX_LAUNCHED();
}

// sustained copy bandwidth (STREAM-copy style, 16 B per lane per access)
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(WG) void k_copy16(const dvec2* __restrict__ s, dvec2* __restrict__ d, long long n) {
    long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {  // four 16-byte loads in flight per lane
        dvec2 v0 = __builtin_nontemporal_load(&s[i]);
        dvec2 v1 = __builtin_nontemporal_load(&s[i + stride]);
        dvec2 v2 = __builtin_nontemporal_load(&s[i + 2 * stride]);
        dvec2 v3 = __builtin_nontemporal_load(&s[i + 3 * stride]);
        __builtin_nontemporal_store(v0, &d[i]);
        __builtin_nontemporal_store(v1, &d[i + stride]);
        __builtin_nontemporal_store(v2, &d[i + 2 * stride]);
        __builtin_nontemporal_store(v3, &d[i + 3 * stride]);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(&s[i]), &d[i]);
}
double stream_copy_gbs(Context* c, size_t bytes, int reps) {
    long long n = (long long)(bytes / 16);
    DBuf<dvec2> a(c, n), b(c, n);
    BCMGX_HIP(hipMemsetAsync(a.get(), 0, n * 16, c->stream));
    unsigned g = std::min<long long>(grid_for(n, WG), (long long)c->num_cus * 32);
    hipEvent_t e0, e1;
    BCMGX_HIP(hipEventCreate(&e0));
    BCMGX_HIP(hipEventCreate(&e1));
    k_copy16<<<g, WG, 0, c->stream>>>(a.get(), b.get(), n);
    BCMGX_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) k_copy16<<<g, WG, 0, c->stream>>>(a.get(), b.get(), n);
    BCMGX_HIP(hipEventRecord(e1, c->stream));
    BCMGX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    BCMGX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 2.0 * n * 16.0 * reps / (ms * 1e-3) / 1e9;
}

}  // namespace bcmgx
