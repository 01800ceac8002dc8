// Kronecker space-time operator Y = alpha * sum_k S_k X T_k^T + beta * Y on the GPU.
//
// Restates kronspmv / kronspgemv (kronecker.f90:19-127): the reference forms
// u_j = S x_j column by column with amux and then mixes y_i += T(i,j) u_j with
// q^2 dense vector passes.  Here one pass over the CSR rows produces, for each
// row, the q*R partial products Z_k(row, :) = sum_e S_k(row, e) X(e, :) in
// registers, exchanges them through LDS inside the workgroup and applies the
// q x q temporal mix before the single store of Y(row, :).  X and Y use the
// interleaved layout X[dof][r*q + j] so every gathered row of X is one
// contiguous run of q*R values.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include <new>

#include "xt_common.h"

namespace xt {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char *get_error() { return g_err; }

// ---------------------------------------------------------------------------
// CSR upload: 1-based host arrays -> 0-based device arrays (+ FP32 copy).
// ---------------------------------------------------------------------------
int csr_upload_1based(int n, const int *ia, const int *ja, const double *a, DevCSR &out,
                      bool want32, hipStream_t s) {
    XT_ARG(n >= 0 && ia && (n == 0 || (ja && a)), "csr arrays");
    long long nnz = (long long)ia[n] - ia[0];
    XT_ARG(nnz >= 0 && ia[0] == 1, "CSR must be 1-based with ia(1) = 1");
    std::vector<int> rp(n + 1), cj((size_t)nnz);
    for (int i = 0; i <= n; ++i) rp[i] = ia[i] - 1;
    for (long long k = 0; k < nnz; ++k) {
        int c = ja[k] - 1;
        XT_ARG(c >= 0 && c < n, "column index out of range");
        cj[k] = c;
    }
    out.n = n;
    out.nnz = nnz;
    XT_TRY(out.rowptr.upload(rp.data(), sizeof(int) * (n + 1), s));
    XT_TRY(out.col.upload(cj.data(), sizeof(int) * nnz, s));
    XT_TRY(out.val64.upload(a, sizeof(double) * nnz, s));
    if (want32) {
        std::vector<float> v32((size_t)nnz);
        for (long long k = 0; k < nnz; ++k) v32[k] = (float)a[k];
        XT_TRY(out.val32.upload(v32.data(), sizeof(float) * nnz, s));
    }
    // host vectors die at return: make the async copies complete first
    XT_HIP(hipStreamSynchronize(s));
    return XT_OK;
}

// ---------------------------------------------------------------------------
// Fused multi-term CSR SpMM + temporal mix.
//   block = 256 lanes = rows_per_block rows x ct columns (ct = column tile,
//   a multiple of q holding whole right-hand sides); lane (lr, col) accumulates
//   Z_k(row, col) for every term k, the workgroup exchanges Z through LDS and
//   lane (lr, col = r*q + i) writes Y(row, col) = sum_k sum_j T_{k,r}(i,j) Z_k(row, r*q + j).
// ---------------------------------------------------------------------------
template <typename T, int NT>
struct KronArgs {
    const int *rp[NT];
    const int *ci[NT];
    const T *v[NT];
};

template <typename T, int NT>
__global__ void __launch_bounds__(256)
kron_csr_kernel(KronArgs<T, NT> A, int p, int c, int q, int col0, int ct, int rpb,
                const T *__restrict__ tmat, T alpha, T beta, const T *__restrict__ x,
                T *__restrict__ y) {
    extern __shared__ unsigned char smem_raw[];
    T *z = reinterpret_cast<T *>(smem_raw);  // [NT][rpb][ct]
    const int lr = threadIdx.x / ct;
    const int col = threadIdx.x - lr * ct;
    const int row = blockIdx.x * rpb + lr;
    const bool inblk = lr < rpb;
    const bool active = inblk && row < p;
    T acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = T(0);
    if (active) {
        const T *xc = x + col0 + col;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int e0 = A.rp[t][row], e1 = A.rp[t][row + 1];
            T s = T(0);
            for (int e = e0; e < e1; ++e) s += A.v[t][e] * xc[(size_t)A.ci[t][e] * c];
            acc[t] = s;
        }
    }
    if (inblk) {
#pragma unroll
        for (int t = 0; t < NT; ++t) z[(t * rpb + lr) * ct + col] = acc[t];
    }
    __syncthreads();
    if (active) {
        const int gcol = col0 + col;
        const int r = gcol / q;
        const int i = gcol - r * q;
        const int zb = col - i;  // first column of this right-hand side in the tile
        T out = T(0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const T *tm = tmat + ((size_t)(r * NT + t) * q + i) * q;
            const T *zr = z + (t * rpb + lr) * ct + zb;
            for (int j = 0; j < q; ++j) out += tm[j] * zr[j];
        }
        const size_t o = (size_t)row * c + gcol;
        y[o] = (beta == T(0)) ? alpha * out : alpha * out + beta * y[o];
    }
}

template <typename T, int NT>
static int launch_kron(const DevCSR *S, int p, int q, int R, const T *tmat, double alpha,
                       double beta, const T *x, T *y, bool fp32, hipStream_t s) {
    KronArgs<T, NT> A;
    for (int t = 0; t < NT; ++t) {
        A.rp[t] = S[t].rowptr.as<int>();
        A.ci[t] = S[t].col.as<int>();
        A.v[t] = fp32 ? (const T *)S[t].val32.ptr : (const T *)S[t].val64.ptr;
    }
    const int c = q * R;
    const int tile = (256 / q) * q;  // whole right-hand sides per tile
    for (int col0 = 0; col0 < c; col0 += tile) {
        const int ct = (c - col0) < tile ? (c - col0) : tile;
        const int rpb = 256 / ct;
        const size_t lds = sizeof(T) * NT * rpb * ct;
        const int grid = ceil_div(p, rpb);
        if (grid == 0) break;
        const bool prof = prof_on(XT_PROF_SPMM);
        if (prof) prof_begin(XT_PROF_SPMM, s);
        hipLaunchKernelGGL((kron_csr_kernel<T, NT>), dim3(grid), dim3(256), lds, s, A, p, c, q,
                           col0, ct, rpb, tmat, (T)alpha, (T)beta, x, y);
        if (prof) {   // CSR values + column indices + row pointers per term, X, Y
            double nnz = 0.0;
            for (int t = 0; t < NT; ++t) nnz += (double)S[t].nnz;
            const double sv = sizeof(T), vec = (double)p * ct * sv;
            const double bytes = nnz * (sv + 4) + NT * 4.0 * (p + 1) + vec * (beta != 0 ? 3 : 2);
            prof_end(XT_PROF_SPMM, s, bytes, 2.0 * nnz * ct + 2.0 * NT * q * (double)p * ct);
        }
        XT_HIP(hipGetLastError());
    }
    return XT_OK;
}

// ---------------------------------------------------------------------------
// Temporal mix Y[dof][r*q+i] = sum_j M_r(i,j) X[dof][r*q+j]
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ void __launch_bounds__(256)
mix_kernel(long long n, int c, int q, const TI *__restrict__ M, const TI *__restrict__ x,
           TO *__restrict__ y) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const long long row = idx / c;
    const int col = (int)(idx - row * c);
    const int r = col / q, i = col - r * q;
    const TI *m = M + ((size_t)r * q + i) * q;
    const TI *xr = x + row * c + r * q;
    TI s = TI(0);
    for (int j = 0; j < q; ++j) s += m[j] * xr[j];
    y[idx] = (TO)s;
}

// The same mix for the temporal orders of the slabs (Q = 6 mechanical, 3 thermal): one
// thread per (DOF, case) group g = dof * R + r, whose Q inputs x[g Q .. g Q + Q) and Q
// outputs are contiguous; the R matrices sit in LDS.  The sums run in the same order as
// mix_kernel's (bit-identical), without its per-element 64-bit index division.
template <typename TI, typename TO, int Q>
__global__ void __launch_bounds__(256)
mix_group_kernel(int ngroups, int R, const TI *__restrict__ M, const TI *__restrict__ x,
                 TO *__restrict__ y) {
    __shared__ TI Ms[256 * Q];   // R Q <= 256
    for (int e = threadIdx.x; e < R * Q * Q; e += 256) Ms[e] = M[e];
    __syncthreads();
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ngroups) return;
    const int r = g % R;
    const TI *xr = x + (size_t)g * Q;
    TI v[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) v[j] = xr[j];
    const TI *m = Ms + r * Q * Q;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        TI sum = TI(0);
#pragma unroll
        for (int j = 0; j < Q; ++j) sum += m[i * Q + j] * v[j];
        y[(size_t)g * Q + i] = (TO)sum;
    }
}

template <typename TI, typename TO>
int temporal_mix(int p, int q, int R, const TI *M, const TI *x, TO *y, hipStream_t s) {
    long long n = (long long)p * q * R;
    if (n == 0) return XT_OK;
    const long long ng = (long long)p * R;
    if ((q == 6 || q == 3) && q * R <= 256 && ng < (1LL << 31)) {
        if (q == 6)
            hipLaunchKernelGGL((mix_group_kernel<TI, TO, 6>), dim3(ceil_div(ng, 256)), dim3(256), 0,
                               s, (int)ng, R, M, x, y);
        else
            hipLaunchKernelGGL((mix_group_kernel<TI, TO, 3>), dim3(ceil_div(ng, 256)), dim3(256), 0,
                               s, (int)ng, R, M, x, y);
    } else {
        hipLaunchKernelGGL((mix_kernel<TI, TO>), dim3(ceil_div(n, 256)), dim3(256), 0, s, n,
                           q * R, q, M, x, y);
    }
    XT_HIP(hipGetLastError());
    return XT_OK;
}
template int temporal_mix<double, double>(int, int, int, const double *, const double *,
                                          double *, hipStream_t);
template int temporal_mix<float, float>(int, int, int, const float *, const float *, float *,
                                        hipStream_t);

// ---------------------------------------------------------------------------
// Slab right-hand side (main3d.f90:803-823): b += coef (x) rext, fixed DOFs zeroed.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
slab_rhs_kernel(long long n, int c, int q, const double *__restrict__ coef,
                const double *__restrict__ rext, long long rext_ld,
                const unsigned char *__restrict__ fixed, double *__restrict__ b) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const long long dof = idx / c;
    const int col = (int)(idx - dof * c);
    const double re = rext[(long long)(col / q) * rext_ld + dof];
    b[idx] = (fixed && fixed[dof]) ? 0.0 : b[idx] + coef[col] * re;
}

// ---------------------------------------------------------------------------
// Layout conversion: temporal-major x[r][j*p + dof] <-> interleaved X[dof][r*q + j]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
to_inter_kernel(long long n, int p, int q, int R, const double *__restrict__ src,
                double *__restrict__ dst) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = q * R;
    const long long dof = idx / c;
    const int col = (int)(idx - dof * c);
    const int r = col / q, j = col - r * q;
    dst[idx] = src[(size_t)r * q * p + (size_t)j * p + dof];
}

__global__ void __launch_bounds__(256)
from_inter_kernel(long long n, int p, int q, int R, const double *__restrict__ src,
                  double *__restrict__ dst) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // temporal-major index
    if (idx >= n) return;
    const long long per = (long long)q * p;
    const int r = (int)(idx / per);
    const long long rem = idx - r * per;
    const int j = (int)(rem / p);
    const long long dof = rem - (long long)j * p;
    dst[idx] = src[dof * (q * R) + r * q + j];
}

}  // namespace xt

using namespace xt;

namespace xt {
struct BsrPlan;
int bsr_build(int p, const int *krp, const int *kci, const double *kv, const int *mrp,
              const int *mci, const double *mv, BsrPlan &P);
template <typename T>
int bsr_apply(const BsrPlan &P, int R, const T *tmat, double alpha, double beta, const T *x,
              T *y, bool f32, hipStream_t s);
BsrPlan *bsr_new();
void bsr_delete(BsrPlan *);
long long bsr_blocks(const BsrPlan *, int real);
int bsr_set_values(BsrPlan &P, const double *kv, const double *mv, hipStream_t s);

__global__ void __launch_bounds__(256)
copy_values_kernel(long long n, const double *__restrict__ v, double *__restrict__ d64,
                   float *__restrict__ d32) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        d64[i] = v[i];
        if (d32) d32[i] = (float)v[i];
    }
}
}  // namespace xt

struct xt_op_s {
    int p = 0, q = 0, nterms = 0;
    DevCSR S[3];
    xt::BsrPlan *bsr = nullptr;  // node-block form (3-DOF mechanics, q = 6), if applicable
    int format = 0;              // 0: node blocks when available, 1: CSR
    xt::SolverWorkspace *ws = nullptr;   // slab-solver buffers, created on first use
    ~xt_op_s() {
        if (bsr) xt::bsr_delete(bsr);
        delete ws;
    }
};

namespace xt {
SolverWorkspace *op_workspace(xt_op op) {
    if (!op->ws) op->ws = new (std::nothrow) SolverWorkspace;
    return op->ws;
}
int op_dims(xt_op op, int *p, int *q, int *nterms) {
    XT_ARG(op, "null operator");
    *p = op->p;
    *q = op->q;
    *nterms = op->nterms;
    return XT_OK;
}
}  // namespace xt

extern "C" {

const char *xt_last_error(void) { return get_error(); }

int xt_device_sync(void) {
    XT_HIP(hipDeviceSynchronize());
    return XT_OK;
}

int xt_op_create(int p, int q, int nterms, const int *const *rowptr, const int *const *colidx,
                 const double *const *vals, xt_op *out) {
    XT_ARG(out && p > 0 && q > 0 && q <= 8 && nterms >= 1 && nterms <= 3, "op sizes");
    XT_ARG(rowptr && colidx && vals, "op arrays");
    xt_op op = new (std::nothrow) xt_op_s;
    if (!op) return XT_E_NOMEM;
    op->p = p;
    op->q = q;
    op->nterms = nterms;
    for (int t = 0; t < nterms; ++t) {
        int rc = csr_upload_1based(p, rowptr[t], colidx[t], vals[t], op->S[t], true, 0);
        if (rc != XT_OK) { delete op; return rc; }
    }
    if (nterms == 2 && q == 6 && p % 3 == 0) {
        std::vector<int> krp(p + 1), mrp(p + 1);
        for (int i = 0; i <= p; ++i) { krp[i] = rowptr[0][i] - 1; mrp[i] = rowptr[1][i] - 1; }
        const long long nk = krp[p], nm = mrp[p];
        std::vector<int> kci(nk), mci(nm);
        for (long long e = 0; e < nk; ++e) kci[e] = colidx[0][e] - 1;
        for (long long e = 0; e < nm; ++e) mci[e] = colidx[1][e] - 1;
        op->bsr = bsr_new();
        int rc = bsr_build(p, krp.data(), kci.data(), vals[0], mrp.data(), mci.data(), vals[1],
                           *op->bsr);
        if (rc == XT_E_UNSUPPORTED) {
            bsr_delete(op->bsr);
            op->bsr = nullptr;
        } else if (rc != XT_OK) {
            delete op;
            return rc;
        }
    }
    *out = op;
    return XT_OK;
}

int xt_op_set_values(xt_op op, const double *const *vals, void *stream) {
    XT_ARG(op && vals, "op_set_values arguments");
    hipStream_t s = (hipStream_t)stream;
    for (int t = 0; t < op->nterms; ++t) {
        XT_ARG(vals[t], "op_set_values: null value array");
        const long long n = op->S[t].nnz;
        if (n)
            hipLaunchKernelGGL(copy_values_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n,
                               vals[t], op->S[t].val64.as<double>(), op->S[t].val32.as<float>());
    }
    if (op->bsr) XT_TRY(bsr_set_values(*op->bsr, vals[0], vals[1], s));
    XT_HIP(hipGetLastError());
    return XT_OK;
}

int xt_op_set_format(xt_op op, int format) {
    XT_ARG(op && (format == 0 || format == 1), "op format");
    op->format = format;
    return XT_OK;
}

int xt_op_format(xt_op op) {
    if (!op) return -1;
    return (op->bsr && op->format == 0) ? 0 : 1;
}

long long xt_op_blocks(xt_op op, int real) {
    if (!op || !op->bsr) return 0;
    return bsr_blocks(op->bsr, real);
}

void xt_op_destroy(xt_op op) { delete op; }

int xt_slab_rhs(int p, int q, int R, const double *coef, const double *rext, long long rext_ld,
                const unsigned char *fixed, double *b, void *stream) {
    XT_ARG(p > 0 && q > 0 && R > 0 && coef && rext && b, "slab_rhs arguments");
    XT_ARG(rext_ld == 0 || rext_ld >= p, "slab_rhs: rext_ld must be 0 or >= p");
    const long long n = (long long)p * q * R;
    hipLaunchKernelGGL(slab_rhs_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, q * R, q, coef, rext, rext_ld, fixed, b);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

long long xt_op_nnz(xt_op op, int k) {
    if (!op || k < 0 || k >= op->nterms) return -1;
    return op->S[k].nnz;
}

int xt_op_apply(xt_op op, int R, const void *tmat, double alpha, double beta, const void *x,
                void *y, int fp32, void *stream) {
    XT_ARG(op && R > 0 && tmat && x && y, "op_apply arguments");
    hipStream_t s = (hipStream_t)stream;
    const int nt = op->nterms;
    if (op->bsr && op->format == 0) {
        if (fp32)
            return bsr_apply<float>(*op->bsr, R, (const float *)tmat, alpha, beta,
                                    (const float *)x, (float *)y, true, s);
        return bsr_apply<double>(*op->bsr, R, (const double *)tmat, alpha, beta,
                                 (const double *)x, (double *)y, false, s);
    }
    if (fp32) {
        const float *tm = (const float *)tmat;
        const float *xx = (const float *)x;
        float *yy = (float *)y;
        if (nt == 1) return launch_kron<float, 1>(op->S, op->p, op->q, R, tm, alpha, beta, xx, yy, true, s);
        if (nt == 2) return launch_kron<float, 2>(op->S, op->p, op->q, R, tm, alpha, beta, xx, yy, true, s);
        return launch_kron<float, 3>(op->S, op->p, op->q, R, tm, alpha, beta, xx, yy, true, s);
    }
    const double *tm = (const double *)tmat;
    const double *xx = (const double *)x;
    double *yy = (double *)y;
    if (nt == 1) return launch_kron<double, 1>(op->S, op->p, op->q, R, tm, alpha, beta, xx, yy, false, s);
    if (nt == 2) return launch_kron<double, 2>(op->S, op->p, op->q, R, tm, alpha, beta, xx, yy, false, s);
    return launch_kron<double, 3>(op->S, op->p, op->q, R, tm, alpha, beta, xx, yy, false, s);
}

int xt_to_interleaved(int p, int q, int R, const double *src, double *dst, void *stream) {
    XT_ARG(p >= 0 && q > 0 && R > 0 && src && dst && src != dst, "layout arguments");
    long long n = (long long)p * q * R;
    if (!n) return XT_OK;
    hipLaunchKernelGGL(to_inter_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, p, q, R, src, dst);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

int xt_from_interleaved(int p, int q, int R, const double *src, double *dst, void *stream) {
    XT_ARG(p >= 0 && q > 0 && R > 0 && src && dst && src != dst, "layout arguments");
    long long n = (long long)p * q * R;
    if (!n) return XT_OK;
    hipLaunchKernelGGL(from_inter_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, p, q, R, src, dst);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

}  // extern "C"
