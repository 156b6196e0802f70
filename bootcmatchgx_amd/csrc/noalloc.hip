// noalloc.hip -- the allocation-free k-column SpMM (bcmgx_spmm_noalloc).
//
// bcmgx_spmm stages its two operands in the interleaved layout of the
// multi-vector kernels (x[i*k + c]) in buffers the library allocates.  This
// variant stages them in one file-scope device array, `scratch`, of
// SCRATCH_BYTES instead: it calls no hipMalloc/hipFree and no allocating
// library routine, for callers that cannot allocate (graph capture, tight
// serving loops).  Each device holds its own copy of the array until the
// process ends.  Host code never names the array where the device address is
// needed: the address of each device's copy comes from hipGetSymbolAddress,
// taken once per device when the first context on it is created.
//
// Every call on one device shares that device's array, which keeps the last
// call's contents (nothing in it needs to start at zero: both staged operands
// are written before they are read).  Calls that overlap on one device (the
// streams of two contexts, or two host threads) race on it.
#include <map>
#include <mutex>

#include "kernels.h"

namespace bcmgx {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device -> address of its copy of `scratch`
}  // namespace

void* scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    int cur = -1;
    BCMGX_HIP(hipGetDevice(&cur));
    BCMGX_CHECK(cur == device, BCMGX_ERR_STATE, "scratch_address: the device must be current");
    void* p = nullptr;
    BCMGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BCMGX_CHECK(p != nullptr, BCMGX_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_addr[device] = p;
    return p;
}

size_t scratch_bytes() { return SCRATCH_BYTES; }

static size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

void spmm_noalloc(Context* c, const DCsr& A, const double* x, double* y, int k, int ld, double alpha, double beta, int op,
                  int kchunk, int lanes) {
    BCMGX_CHECK(c->scratch != nullptr, BCMGX_ERR_STATE, "spmm_noalloc: no scratch address for this device");
    BCMGX_CHECK(k >= 1 && k <= 64, BCMGX_ERR_ARG, "spmm_noalloc: k must be in [1, 64]");
    BCMGX_CHECK(ld >= std::max(A.n, A.m), BCMGX_ERR_ARG, "spmm_noalloc: leading dimension smaller than the matrix");
    size_t xb = round256((size_t)std::max(A.m, 1) * k * sizeof(double));
    size_t yb = round256((size_t)std::max(A.n, 1) * k * sizeof(double));
    BCMGX_CHECK(xb + yb <= SCRATCH_BYTES, BCMGX_ERR_ARG,
                "spmm_noalloc: (rows + columns) x k x 8 bytes = " + std::to_string(xb + yb) +
                    " bytes exceeds the " + std::to_string(SCRATCH_BYTES) + "-byte scratch array");
    if (A.n == 0) return;
    double* X = static_cast<double*>(c->scratch);
    double* Y = reinterpret_cast<double*>(static_cast<unsigned char*>(c->scratch) + xb);
    if (k == 1) {
        copy_d2d(c, X, x, sizeof(double) * (size_t)A.m);
        if (op != 1) copy_d2d(c, Y, y, sizeof(double) * (size_t)A.n);
    } else {
        colmajor_to_interleaved(c, x, ld, X, A.m, k);
        if (op != 1) colmajor_to_interleaved(c, y, ld, Y, A.n, k);
    }
    SpmvArgs a;
    a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
    a.x = X; a.y = Y; a.b = Y; a.alpha = alpha; a.beta = beta; a.k = k; a.kchunk = kchunk;
    a.lanes = k == 1 ? lanes : 0;
    SpOp o = op == 1 ? SpOp::AX : (op == 2 ? SpOp::RES : SpOp::AXPBY);
    spmv<double, double>(c, o, a);
    if (k == 1) copy_d2d(c, y, Y, sizeof(double) * (size_t)A.n);
    else interleaved_to_colmajor(c, Y, y, ld, A.n, k);
}

}  // namespace bcmgx
