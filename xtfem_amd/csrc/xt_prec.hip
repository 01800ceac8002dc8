// Kronecker ILUT preconditioner on the GPU: Y = (LU)^-1 X Minv_r^T.
//
// Restates kronlusol (kronecker.f90:130-178) and lusol (ilut.f:2383-2432).  The
// reference inverts the q x q temporal matrix A, runs q scalar forward/backward
// sweeps over the MSR factors of ilut and mixes the q results with inv(A).
// Here the temporal mix is applied first (it commutes with the spatial solve),
// and the q*R columns of every row are solved together: rows are grouped into
// dependency levels (level(i) = 1 + max level of the rows it reads), and each
// level is one launch in which lane (row, column) performs the reference's
// row update x_i = y_i - sum_k alu(k) x(jlu(k)) (backward: times the stored
// inverse diagonal) in the reference's entry order.
#include <algorithm>
#include <numeric>

#include "xt_common.h"

namespace xt {

template <typename T>
__global__ void __launch_bounds__(256)
trsv_level_kernel(int nrows, const int *__restrict__ rows, const int *__restrict__ rp,
                  const int *__restrict__ ci, const T *__restrict__ v,
                  const T *__restrict__ dinv, int c, const T *__restrict__ b, T *__restrict__ x) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nrows * c) return;
    const int k = (int)(idx / c);
    const int col = (int)(idx - (long long)k * c);
    const int row = rows[k];
    const size_t o = (size_t)row * c + col;
    T s = b ? b[o] : x[o];
    const int e1 = rp[row + 1];
    for (int e = rp[row]; e < e1; ++e) s -= v[e] * x[(size_t)ci[e] * c + col];
    if (dinv) s *= dinv[row];
    x[o] = s;
}

// Both sweeps in one workgroup (small factors): the levels are walked in order with a
// barrier between levels instead of one launch per level, then the upper levels.
// Same per-row arithmetic and order as trsv_level_kernel.  XL: the solution vector
// lives in LDS during the sweeps (dynamic shared memory of n*c elements).
template <typename T, bool XL>
__global__ void __launch_bounds__(1024)
trsv_onewg_kernel(int n_, int nlevL, const int *__restrict__ lpL, const int *__restrict__ rowsL,
                  const int *__restrict__ rpL, const int *__restrict__ ciL,
                  const T *__restrict__ vL, int nlevU, const int *__restrict__ lpU,
                  const int *__restrict__ rowsU, const int *__restrict__ rpU,
                  const int *__restrict__ ciU, const T *__restrict__ vU,
                  const T *__restrict__ dinv, int c, const T *__restrict__ b, T *xg) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *x = XL ? reinterpret_cast<T *>(smem_raw) : xg;
    for (int l = 0; l < nlevL; ++l) {
        const int r0 = lpL[l], n = (lpL[l + 1] - r0) * c;
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
            const int k = idx / c, col = idx - k * c;
            const int row = rowsL[r0 + k];
            const size_t o = (size_t)row * c + col;
            T s = b[o];
            const int e1 = rpL[row + 1];
            for (int e = rpL[row]; e < e1; ++e) s -= vL[e] * x[(size_t)ciL[e] * c + col];
            x[o] = s;
        }
        __syncthreads();
    }
    for (int l = 0; l < nlevU; ++l) {
        const int r0 = lpU[l], n = (lpU[l + 1] - r0) * c;
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
            const int k = idx / c, col = idx - k * c;
            const int row = rowsU[r0 + k];
            const size_t o = (size_t)row * c + col;
            T s = x[o];
            const int e1 = rpU[row + 1];
            for (int e = rpU[row]; e < e1; ++e) s -= vU[e] * x[(size_t)ciU[e] * c + col];
            x[o] = dinv[row] * s;
        }
        __syncthreads();
    }
    if (XL)
        for (int i = threadIdx.x; i < n_ * c; i += blockDim.x) xg[i] = x[i];
}

// Level schedule of a triangular CSR: forward (lower) or backward (upper).
static void build_levels(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                         bool forward, std::vector<int> &lev_ptr, std::vector<int> &lev_rows) {
    std::vector<int> lev(n, 0);
    int nlev = 0;
    for (int t = 0; t < n; ++t) {
        const int i = forward ? t : n - 1 - t;
        int l = 0;
        for (int e = rp[i]; e < rp[i + 1]; ++e) l = std::max(l, lev[ci[e]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    lev_ptr.assign(nlev + 1, 0);
    for (int i = 0; i < n; ++i) lev_ptr[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) lev_ptr[l + 1] += lev_ptr[l];
    lev_rows.resize(n);
    std::vector<int> fill(lev_ptr.begin(), lev_ptr.end() - 1);
    for (int t = 0; t < n; ++t) {
        const int i = forward ? t : n - 1 - t;
        lev_rows[fill[lev[i]]++] = i;
    }
}

struct Tri {
    DevCSR m;                    // strict triangle, 0-based, reference entry order
    DevBuf lev_rows;             // rows sorted by level
    std::vector<int> lev_ptr;    // host copy of level boundaries
    DevBuf lev_ptr_dev;          // device copy (single-workgroup sweeps)
    std::vector<long long> lev_nnz;  // entries per level (algorithmic bytes)
};

}  // namespace xt

using namespace xt;

struct xt_prec_s {
    int p = 0;
    Tri L, U;                  // ILUT factors (level-scheduled solves)
    DevBuf dinv64, dinv32;
    DevBuf work;
    BandPrec *band = nullptr;  // or: exact banded factor (block sweeps)
    ~xt_prec_s() { if (band) band_delete(band); }
};

namespace xt {

static int upload_tri(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                      const std::vector<double> &v, bool forward, Tri &T) {
    T.m.n = n;
    T.m.nnz = rp[n];
    XT_TRY(T.m.rowptr.upload(rp.data(), sizeof(int) * (n + 1)));
    XT_TRY(T.m.col.upload(ci.data(), sizeof(int) * ci.size()));
    XT_TRY(T.m.val64.upload(v.data(), sizeof(double) * v.size()));
    std::vector<float> v32(v.begin(), v.end());
    XT_TRY(T.m.val32.upload(v32.data(), sizeof(float) * v32.size()));
    std::vector<int> lrows;
    build_levels(n, rp, ci, forward, T.lev_ptr, lrows);
    XT_TRY(T.lev_ptr_dev.upload(T.lev_ptr.data(), sizeof(int) * T.lev_ptr.size()));
    T.lev_nnz.assign(T.lev_ptr.size() - 1, 0);
    for (size_t l = 0; l + 1 < T.lev_ptr.size(); ++l)
        for (int k = T.lev_ptr[l]; k < T.lev_ptr[l + 1]; ++k)
            T.lev_nnz[l] += rp[lrows[k] + 1] - rp[lrows[k]];
    XT_TRY(T.lev_rows.upload(lrows.data(), sizeof(int) * n));
    XT_HIP(hipDeviceSynchronize());
    return XT_OK;
}

template <typename T>
static int run_tri(const Tri &tri, const T *vals, const T *dinv, int c, const T *b, T *x,
                   hipStream_t s) {
    const int nlev = (int)tri.lev_ptr.size() - 1;
    const int *rows = tri.lev_rows.as<int>();
    for (int l = 0; l < nlev; ++l) {
        const int r0 = tri.lev_ptr[l], nr = tri.lev_ptr[l + 1] - r0;
        if (!nr) continue;
        long long n = (long long)nr * c;
        const bool prof = prof_on(XT_PROF_TRSV);
        if (prof) prof_begin(XT_PROF_TRSV, s);
        hipLaunchKernelGGL(trsv_level_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, nr,
                           rows + r0, tri.m.rowptr.as<int>(), tri.m.col.as<int>(), vals, dinv,
                           c, b, x);
        if (prof) {   // level rows: values + column indices of their entries, row data, x
            const double nnz = (double)(tri.lev_nnz[l]);
            prof_end(XT_PROF_TRSV, s, nnz * (sizeof(T) + 4) + nr * 12.0 +
                                          (double)nr * c * sizeof(T) * (b ? 2 : 1) +
                                          nnz * c * sizeof(T), 2.0 * nnz * c);
        }
    }
    XT_HIP(hipGetLastError());
    return XT_OK;
}

// Sweeps whose work (factor entries + diagonal, times columns) is at most this run in a
// single workgroup.
constexpr long long kOneWgWork = 4LL << 20;
// ... and keep the solution vector in LDS when it fits this many bytes.
constexpr size_t kOneWgLds = 128 * 1024;

template <typename T>
int prec_apply_t(xt_prec pr, int R, int q, const T *minv, const T *x, T *y, bool f32,
                 hipStream_t s) {
    if (pr->band) return band_apply<T>(pr->band, R, q, minv, x, y, s);
    const int c = q * R;
    const T *b = x;
    if (minv) {
        XT_TRY(pr->work.reserve(sizeof(T) * (size_t)pr->p * c));
        XT_TRY(temporal_mix<T, T>(pr->p, q, R, minv, x, pr->work.as<T>(), s));
        b = pr->work.as<T>();
    }
    const T *lv = f32 ? (const T *)pr->L.m.val32.ptr : (const T *)pr->L.m.val64.ptr;
    const T *uv = f32 ? (const T *)pr->U.m.val32.ptr : (const T *)pr->U.m.val64.ptr;
    const T *dv = f32 ? (const T *)pr->dinv32.ptr : (const T *)pr->dinv64.ptr;
    // Small factors: both sweeps in one workgroup (a barrier per level instead of a
    // launch per level); large ones: one launch per level over the whole device.
    const long long work = (pr->L.m.nnz + pr->U.m.nnz + 2LL * pr->p) * c;
    if (work <= kOneWgWork) {
        const bool prof = prof_on(XT_PROF_TRSV);
        if (prof) prof_begin(XT_PROF_TRSV, s);
        const size_t xl = sizeof(T) * (size_t)pr->p * c;
#define XT_ONEWG(XLDS)                                                                      \
    hipLaunchKernelGGL((trsv_onewg_kernel<T, XLDS>), dim3(1), dim3(1024), XLDS ? xl : 0, s,  \
                       pr->p, (int)pr->L.lev_ptr.size() - 1, pr->L.lev_ptr_dev.as<int>(),     \
                       pr->L.lev_rows.as<int>(), pr->L.m.rowptr.as<int>(),                    \
                       pr->L.m.col.as<int>(), lv, (int)pr->U.lev_ptr.size() - 1,              \
                       pr->U.lev_ptr_dev.as<int>(), pr->U.lev_rows.as<int>(),                 \
                       pr->U.m.rowptr.as<int>(), pr->U.m.col.as<int>(), uv, dv, c, b, y)
        if (xl <= kOneWgLds) XT_ONEWG(true);
        else XT_ONEWG(false);
#undef XT_ONEWG
        if (prof)
            prof_end(XT_PROF_TRSV, s,
                     (double)(pr->L.m.nnz + pr->U.m.nnz) * (sizeof(T) + 4 + c * sizeof(T)) +
                         (double)pr->p * c * sizeof(T) * 3,
                     2.0 * (pr->L.m.nnz + pr->U.m.nnz) * c);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    XT_TRY(run_tri<T>(pr->L, lv, nullptr, c, b, y, s));
    XT_TRY(run_tri<T>(pr->U, uv, dv, c, nullptr, y, s));
    return XT_OK;
}
template int prec_apply_t<double>(xt_prec, int, int, const double *, const double *, double *,
                                  bool, hipStream_t);
template int prec_apply_t<float>(xt_prec, int, int, const float *, const float *, float *, bool,
                                 hipStream_t);

}  // namespace xt

extern "C" {

int xt_prec_create(int p, const double *alu, const int *jlu, const int *ju, xt_prec *out) {
    XT_ARG(out && p > 0 && alu && jlu && ju, "prec arguments");
    XT_HIP(hipFuncSetAttribute((const void *)trsv_onewg_kernel<double, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOneWgLds));
    XT_HIP(hipFuncSetAttribute((const void *)trsv_onewg_kernel<float, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOneWgLds));
    XT_ARG(jlu[0] == p + 2, "MSR factors must come from ilut (jlu(1) = n+2)");
    std::vector<int> lrp(p + 1, 0), urp(p + 1, 0), lci, uci;
    std::vector<double> lv, uv, dinv(p);
    for (int i = 0; i < p; ++i) {
        const int l0 = jlu[i] - 1, l1 = ju[i] - 1, u1 = jlu[i + 1] - 1;
        XT_ARG(l0 <= l1 && l1 <= u1, "inconsistent MSR row pointers");
        for (int k = l0; k < l1; ++k) {
            const int cidx = jlu[k] - 1;
            XT_ARG(cidx >= 0 && cidx < i, "L entry not strictly lower");
            lci.push_back(cidx);
            lv.push_back(alu[k]);
        }
        for (int k = l1; k < u1; ++k) {
            const int cidx = jlu[k] - 1;
            XT_ARG(cidx > i && cidx < p, "U entry not strictly upper");
            uci.push_back(cidx);
            uv.push_back(alu[k]);
        }
        lrp[i + 1] = (int)lci.size();
        urp[i + 1] = (int)uci.size();
        dinv[i] = alu[i];
    }
    xt_prec pr = new (std::nothrow) xt_prec_s;
    if (!pr) return XT_E_NOMEM;
    pr->p = p;
    int rc = upload_tri(p, lrp, lci, lv, true, pr->L);
    if (rc == XT_OK) rc = upload_tri(p, urp, uci, uv, false, pr->U);
    if (rc == XT_OK) rc = pr->dinv64.upload(dinv.data(), sizeof(double) * p);
    if (rc == XT_OK) {
        std::vector<float> d32(dinv.begin(), dinv.end());
        rc = pr->dinv32.upload(d32.data(), sizeof(float) * p);
        if (rc == XT_OK && hipDeviceSynchronize() != hipSuccess) rc = XT_E_HIP;
    }
    if (rc != XT_OK) { delete pr; return rc; }
    *out = pr;
    return XT_OK;
}

void xt_prec_destroy(xt_prec pr) { delete pr; }

int xt_prec_create_band(int p, const int *ia, const int *ja, const double *a, int block,
                        xt_prec *out) {
    XT_ARG(out && p > 0 && ia && ja && a, "prec_create_band arguments");
    xt_prec pr = new (std::nothrow) xt_prec_s;
    if (!pr) return XT_E_NOMEM;
    pr->p = p;
    int rc = band_build(p, ia, ja, a, block, &pr->band);
    if (rc != XT_OK) { delete pr; return rc; }
    *out = pr;
    return XT_OK;
}

int xt_prec_band_info(xt_prec pr, int *block, int *nblocks, int *bandwidth) {
    XT_ARG(pr && pr->band, "not a banded-factor preconditioner");
    band_dims(pr->band, block, nblocks, bandwidth);
    return XT_OK;
}

int xt_prec_levels(xt_prec pr, int *nlev_lower, int *nlev_upper) {
    XT_ARG(pr, "prec");
    if (pr->band) {   // sequential block steps of each sweep
        int nb = 0;
        band_dims(pr->band, nullptr, &nb, nullptr);
        if (nlev_lower) *nlev_lower = nb;
        if (nlev_upper) *nlev_upper = nb;
        return XT_OK;
    }
    if (nlev_lower) *nlev_lower = (int)pr->L.lev_ptr.size() - 1;
    if (nlev_upper) *nlev_upper = (int)pr->U.lev_ptr.size() - 1;
    return XT_OK;
}

int xt_prec_apply(xt_prec pr, int R, int q, const void *minv, const void *x, void *y, int fp32,
                  void *stream) {
    XT_ARG(pr && R > 0 && q > 0 && x && y && x != y, "prec_apply arguments");
    hipStream_t s = (hipStream_t)stream;
    if (fp32)
        return prec_apply_t<float>(pr, R, q, (const float *)minv, (const float *)x, (float *)y,
                                   true, s);
    return prec_apply_t<double>(pr, R, q, (const double *)minv, (const double *)x, (double *)y,
                                false, s);
}

}  // extern "C"
