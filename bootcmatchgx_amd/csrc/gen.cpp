// gen.cpp -- host problem generators and Matrix Market input.
//
// generateLocalLaplacian3D_7p / _27p (src/generator/laplacian.cu:119-334,
// :336-777) for the whole P x Q x R process grid at once: global rows in the
// block order of internal_index (:14-25), diagonal 6 (26), every in-range
// neighbour -1 (Dirichlet by omission), columns sorted per row.
// readMTXDouble + check_and_fix_order (src/datastruct/CSR.cu:1973-2152,
// :1895-1918).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <utility>
#include <vector>

#include "../../include/bcmgx.h"

namespace {

template <typename T>
T* dup(const std::vector<T>& v) {
    T* p = (T*)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

long long block_index(long long gi, long long gj, long long gk, int nx, int ny, int nz, int P, int Q) {
    long long i = gi % nx, j = gj % ny, k = gk % nz;
    long long p = gi / nx, q = gj / ny, r = gk / nz;
    return (r * P * Q + q * P + p) * (long long)nx * ny * nz + (k * (long long)nx * ny + j * nx + i);
}

}  // namespace

extern "C" {

void bcmgx_free(void* p) { std::free(p); }

int bcmgx_gen_laplacian3d(int points, int nx, int ny, int nz, int P, int Q, int R, int* n_out, int* nnz_out, int** row_out,
                          int** col_out, double** val_out) {
    if ((points != 7 && points != 27) || nx <= 0 || ny <= 0 || nz <= 0 || P <= 0 || Q <= 0 || R <= 0) return BCMGX_ERR_ARG;
    long long gnx = (long long)nx * P, gny = (long long)ny * Q, gnz = (long long)nz * R;
    long long N = gnx * gny * gnz;
    long long per = points == 7 ? 7 : 27;
    if (N * per >= (1ll << 31)) return BCMGX_ERR_UNSUPPORTED;
    std::vector<int> row((size_t)N + 1, 0), col;
    std::vector<double> val;
    col.reserve((size_t)(N * per));
    val.reserve((size_t)(N * per));
    const double dg = points == 7 ? 6.0 : 26.0;
    bool natural = (P == 1 && Q == 1 && R == 1);
    std::vector<std::pair<int, double>> ent;
    long long cnt = 0;
    for (int r = 0; r < R; r++)
        for (int q = 0; q < Q; q++)
            for (int p = 0; p < P; p++)
                for (int k = 0; k < nz; k++)
                    for (int j = 0; j < ny; j++)
                        for (int i = 0; i < nx; i++) {
                            long long gi = (long long)p * nx + i, gj = (long long)q * ny + j, gk = (long long)r * nz + k;
                            ent.clear();
                            for (int dz = -1; dz <= 1; dz++)
                                for (int dy = -1; dy <= 1; dy++)
                                    for (int dx = -1; dx <= 1; dx++) {
                                        int nzc = (dx != 0) + (dy != 0) + (dz != 0);
                                        if (points == 7 && nzc > 1) continue;
                                        long long a = gi + dx, b = gj + dy, c = gk + dz;
                                        if (a < 0 || a >= gnx || b < 0 || b >= gny || c < 0 || c >= gnz) continue;
                                        ent.push_back({(int)block_index(a, b, c, nx, ny, nz, P, Q), nzc ? -1.0 : dg});
                                    }
                            // lexicographic (dz, dy, dx) is already column order when P=Q=R=1
                            if (!natural)
                                std::sort(ent.begin(), ent.end(),
                                          [](const std::pair<int, double>& x, const std::pair<int, double>& y) { return x.first < y.first; });
                            for (auto& e : ent) {
                                col.push_back(e.first);
                                val.push_back(e.second);
                            }
                            row[++cnt] = (int)col.size();
                        }
    *n_out = (int)N;
    *nnz_out = (int)col.size();
    *row_out = dup(row);
    *col_out = dup(col);
    *val_out = dup(val);
    return BCMGX_OK;
}

int bcmgx_gen_diffusion7p(int nx, int ny, int nz, int kind, double p1, double p2, uint64_t seed, int* n_out, int* nnz_out,
                          int** row_out, int** col_out, double** val_out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || (kind != 0 && kind != 1)) return BCMGX_ERR_ARG;
    // face conductivities in the order: x-faces (k, j, i = 0..nx), y-faces, z-faces
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(p1, p2);
    auto draw = [&]() { return kind == 0 ? std::exp(p1 * nd(rng)) : ud(rng); };
    size_t sx = (size_t)nx + 1, sy = (size_t)ny + 1;
    std::vector<double> kx(sx * ny * nz), ky((size_t)nx * sy * nz), kz((size_t)nx * ny * (nz + 1));
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (size_t i = 0; i < sx; i++) kx[((size_t)k * ny + j) * sx + i] = draw();
    for (int k = 0; k < nz; k++)
        for (size_t j = 0; j < sy; j++)
            for (int i = 0; i < nx; i++) ky[((size_t)k * sy + j) * nx + i] = draw();
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) kz[((size_t)k * ny + j) * nx + i] = draw();
    long long N = (long long)nx * ny * nz;
    if (N * 7 >= (1ll << 31)) return BCMGX_ERR_UNSUPPORTED;
    std::vector<int> row((size_t)N + 1, 0), col;
    std::vector<double> val;
    col.reserve((size_t)N * 7);
    val.reserve((size_t)N * 7);
    auto id = [&](int i, int j, int k) { return (int)(((long long)k * ny + j) * nx + i); };
    long long cnt = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double xm = kx[((size_t)k * ny + j) * sx + i], xp = kx[((size_t)k * ny + j) * sx + i + 1];
                double ym = ky[((size_t)k * sy + j) * nx + i], yp = ky[((size_t)k * sy + j + 1) * nx + i];
                double zm = kz[((size_t)k * ny + j) * nx + i], zp = kz[((size_t)(k + 1) * ny + j) * nx + i];
                double diag = xm + xp + ym + yp + zm + zp;
                if (k > 0) { col.push_back(id(i, j, k - 1)); val.push_back(-zm); }
                if (j > 0) { col.push_back(id(i, j - 1, k)); val.push_back(-ym); }
                if (i > 0) { col.push_back(id(i - 1, j, k)); val.push_back(-xm); }
                col.push_back(id(i, j, k)); val.push_back(diag);
                if (i < nx - 1) { col.push_back(id(i + 1, j, k)); val.push_back(-xp); }
                if (j < ny - 1) { col.push_back(id(i, j + 1, k)); val.push_back(-yp); }
                if (k < nz - 1) { col.push_back(id(i, j, k + 1)); val.push_back(-zp); }
                row[++cnt] = (int)col.size();
            }
    *n_out = (int)N;
    *nnz_out = (int)col.size();
    *row_out = dup(row);
    *col_out = dup(col);
    *val_out = dup(val);
    return BCMGX_OK;
}

int bcmgx_read_mtx(const char* path, int* n_out, int* m_out, int* nnz_out, int** row_out, int** col_out, double** val_out) {
    if (!path) return BCMGX_ERR_ARG;
    FILE* fp = std::fopen(path, "r");
    if (!fp) return BCMGX_ERR_IO;
    char banner[64], mtx[64], crd[64], dtype[64], scheme[64], buf[1025];
    int status = BCMGX_ERR_IO;
    std::vector<long> ci, cj;
    std::vector<double> cv;
    unsigned long nr = 0, nc = 0, nz = 0;
    bool sym = false;
    do {
        if (std::fscanf(fp, "%63s %63s %63s %63s %63s\n", banner, mtx, crd, dtype, scheme) != 5) break;
        if (!std::fgets(buf, 1024, fp)) break;
        bool ok = true;
        while (buf[0] == '%')
            if (!std::fgets(buf, 1024, fp)) { ok = false; break; }
        if (!ok) break;
        if (std::sscanf(buf, "%lu %lu %lu", &nr, &nc, &nz) != 3) break;
        if (std::strcmp(dtype, "real") != 0) { status = BCMGX_ERR_UNSUPPORTED; break; }
        sym = std::strcmp(scheme, "symmetric") == 0;
        if (!sym && std::strcmp(scheme, "general") != 0) { status = BCMGX_ERR_UNSUPPORTED; break; }
        ci.reserve(nz * (sym ? 2 : 1));
        bool bad = false;
        for (unsigned long e = 0; e < nz; e++) {
            unsigned long r, c;
            double v;
            if (!std::fgets(buf, 1024, fp) || std::sscanf(buf, "%lu %lu %le", &r, &c, &v) != 3 || r < 1 || c < 1 || r > nr || c > nc) {
                bad = true;
                break;
            }
            ci.push_back((long)r - 1); cj.push_back((long)c - 1); cv.push_back(v);
            if (sym && r != c) { ci.push_back((long)c - 1); cj.push_back((long)r - 1); cv.push_back(v); }
        }
        if (bad) break;
        status = BCMGX_OK;
    } while (0);
    std::fclose(fp);
    if (status != BCMGX_OK) return status;
    if (ci.size() >= (1ul << 31)) return BCMGX_ERR_UNSUPPORTED;
    std::vector<int> row(nr + 1, 0), col(ci.size());
    std::vector<double> val(ci.size());
    for (long r : ci) row[r + 1]++;
    for (unsigned long i = 0; i < nr; i++) row[i + 1] += row[i];
    std::vector<int> pos(row.begin(), row.end() - 1);
    for (size_t e = 0; e < ci.size(); e++) {
        int p = pos[ci[e]]++;
        col[p] = (int)cj[e];
        val[p] = cv[e];
    }
    // check_and_fix_order: stable sort of rows found out of order
    for (unsigned long i = 0; i < nr; i++) {
        int s = row[i], e = row[i + 1];
        bool sorted = true;
        for (int j = s + 1; j < e; j++)
            if (col[j] < col[j - 1]) { sorted = false; break; }
        if (sorted) continue;
        std::vector<std::pair<int, double>> t;
        for (int j = s; j < e; j++) t.push_back({col[j], val[j]});
        std::stable_sort(t.begin(), t.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
        for (int j = s; j < e; j++) { col[j] = t[j - s].first; val[j] = t[j - s].second; }
    }
    *n_out = (int)nr;
    *m_out = (int)nc;
    *nnz_out = (int)col.size();
    *row_out = dup(row);
    *col_out = dup(col);
    *val_out = dup(val);
    return BCMGX_OK;
}

}  // extern "C"
