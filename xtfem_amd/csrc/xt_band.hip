// Exact factor preconditioner K^-1 for banded SPD K, applied as block sweeps.
//
// The reference's live solver preconditions GMRES with the exact factor of K (MUMPS
// LDL^T of the lower triangle, mumps.f90:22-145, applied to all q temporal blocks by
// mpilusol, kronecker_mumps.f90:120-183).  ILUT with lfil >= n and droptol = 0 is the
// same exact factor (ilut.f:2169-2170).  Here K (bandwidth w after the mesh numbering)
// is cut into nb diagonal blocks of b >= w rows, so it is block tridiagonal:
//
//   K = L L^T,  L = blockbidiag(L_ii, E_i = L_{i+1,i})     (block Cholesky, FP64, device)
//
// With D_i = L_ii^-1 the two triangular solves become dense block recurrences
//
//   forward   y_0 = D_0 x_0,      y_{i+1} = D_{i+1} x_{i+1} - G_i y_i,   G_i = D_{i+1} E_i
//   backward  z_nb = D^T y,       z_i     = D_i^T y_i     - H_i z_{i+1}, H_i = (E_i D_i)^T
//
// The D x and D^T y products of all blocks run in one batched launch; each recurrence
// step is one launch over b/8 row tiles.  All four matrix sets are stored row-major
// (the row of the operator a tile needs is contiguous) in FP64 and FP32.
//
// The factorization runs once per K on the device with rocSOLVER potrf/trtri and
// rocBLAS syrk/trsm/gemm (setup, not on the per-iteration path).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "xt_common.h"

namespace xt {

struct BandPrec {
    int p = 0, b = 0, nb = 0, bw = 0;
    // [0] FP64, [1] FP32.  Dl: D_i row-major (lower), Du: D_i^T row-major (upper),
    // G, H: recurrence matrices (nb - 1 of each).
    DevBuf Dl[2], Du[2], G[2], H[2];
    DevBuf w1, w2;
};

// ---------------------------------------------------------------------------
// out_blk[r][col] = (add ? add_blk[r][col] - s : s),  s = sum_k A_blk[r][k] V_blk[k][col]
// A_blk: b x b row-major; V, add, out: b x c row-major (c columns of one row contiguous).
// tri = 1: A lower triangular (k <= r), tri = 2: upper (k >= r); skipped parts are zero.
// Grid: x = b / 8 row tiles, y = column chunks of cw (<= 64), z = blocks of the batch.
// The k range is walked in chunks of KC rows.  Each thread loads its share of a chunk
// (the 8 x KC slice of A and the KC x cw slice of V) into registers with every load in
// flight at once, the chunk goes to LDS, and the loads of the next chunk are issued
// before the current one is consumed.  Thread t accumulates column t % cw over the
// k-slice t / cw; the slices are summed through LDS at the end.
// ---------------------------------------------------------------------------
constexpr int kTM = 8;

template <typename T>
struct BandTile {
    static constexpr int KC = sizeof(T) == 4 ? 256 : 128;   // k rows per chunk
    static constexpr int VEC = 16 / sizeof(T);               // elements per 16-byte load
    static constexpr int AV = KC * kTM / VEC / 256;          // A vectors per thread
    static constexpr int VV = KC * 64 / VEC / 256;           // V vectors per thread (max)
    static constexpr size_t smem = sizeof(T) * ((size_t)KC * kTM + (size_t)KC * 64 + kTM * 256);
};

template <typename T>
__global__ void __launch_bounds__(256)
band_gemm_kernel(int b, int c, int cw, int tri, const T *__restrict__ A, long long a_stride,
                 const T *__restrict__ V, long long v_stride, const T *add, long long add_stride,
                 T *out, long long out_stride) {
    using BT = BandTile<T>;
    constexpr int KC = BT::KC, VEC = BT::VEC, AV = BT::AV, VV = BT::VV;
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *As = reinterpret_cast<T *>(smem_raw);   // [KC][kTM]
    T *Vs = As + KC * kTM;                      // [KC][cw]
    T *red = Vs + KC * 64;                      // [kTM][ks][cw]
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * kTM;
    const int n0 = blockIdx.y * cw;
    const long long z = blockIdx.z;
    A += z * a_stride;
    V += z * v_stride;
    out += z * out_stride;
    if (add) add += z * add_stride;

    int k0 = 0, k1 = b;
    if (tri == 1) k1 = min(b, r0 + kTM);
    if (tri == 2) k0 = r0;
    // V layouts: one contiguous span when the chunk holds every column (c <= cw);
    // otherwise cw-wide row segments, vector loads when rows stay 16-byte aligned.
    const bool span = c <= cw;
    const bool vrow = !span && (c % VEC == 0);
    const int segv = cw / VEC;                  // vectors per row segment

    vec_t ra[AV], rv[VV];
    T rs[VV * VEC];
    auto load = [&](int kc0) {
        const int nk = min(KC, k1 - kc0);      // a multiple of 8
        const int nvr = nk / VEC;
#pragma unroll
        for (int j = 0; j < AV; ++j) {
            const int e = tid + j * 256;
            if (e < kTM * nvr) {
                const int r = e / nvr, kv = e - r * nvr;
                ra[j] = *reinterpret_cast<const vec_t *>(A + (size_t)(r0 + r) * b + kc0 + kv * VEC);
            }
        }
        if (span) {
            const T *src = V + (size_t)kc0 * c;
            const int nv = nk * c / VEC;
#pragma unroll
            for (int j = 0; j < VV; ++j) {
                const int e = tid + j * 256;
                if (e < nv) rv[j] = *reinterpret_cast<const vec_t *>(src + (size_t)e * VEC);
            }
        } else if (vrow) {
#pragma unroll
            for (int j = 0; j < VV; ++j) {
                const int e = tid + j * 256;
                const int kk = e / segv, cv = e - kk * segv;
                if (kk < nk && n0 + cv * VEC < c)
                    rv[j] = *reinterpret_cast<const vec_t *>(V + (size_t)(kc0 + kk) * c + n0 + cv * VEC);
            }
        } else {
#pragma unroll
            for (int j = 0; j < VV * VEC; ++j) {
                const int e = tid + j * 256;
                const int kk = e / cw, cl = e - kk * cw;
                rs[j] = (kk < nk && n0 + cl < c) ? V[(size_t)(kc0 + kk) * c + n0 + cl] : T(0);
            }
        }
    };
    auto store = [&](int kc0) {
        const int nk = min(KC, k1 - kc0);
        const int nvr = nk / VEC;
#pragma unroll
        for (int j = 0; j < AV; ++j) {
            const int e = tid + j * 256;
            if (e < kTM * nvr) {
                const int r = e / nvr, kv = e - r * nvr;
#pragma unroll
                for (int u = 0; u < VEC; ++u) As[(kv * VEC + u) * kTM + r] = ra[j][u];
            }
        }
        if (span) {
            const int nv = nk * c / VEC;
#pragma unroll
            for (int j = 0; j < VV; ++j) {
                const int e = tid + j * 256;
                if (e < nv) {
#pragma unroll
                    for (int u = 0; u < VEC; ++u) {
                        const int i = e * VEC + u, kk = i / c;
                        Vs[kk * cw + (i - kk * c)] = rv[j][u];
                    }
                }
            }
        } else if (vrow) {
#pragma unroll
            for (int j = 0; j < VV; ++j) {
                const int e = tid + j * 256;
                const int kk = e / segv, cv = e - kk * segv;
                if (kk < nk) {
                    const bool ok = n0 + cv * VEC < c;
#pragma unroll
                    for (int u = 0; u < VEC; ++u) Vs[kk * cw + cv * VEC + u] = ok ? rv[j][u] : T(0);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < VV * VEC; ++j) {
                const int e = tid + j * 256;
                if (e < nk * cw) Vs[e] = rs[j];
            }
        }
    };

    const int ks_n = 256 / cw;
    const int col = tid % cw, ks = tid / cw;
    const int cc = n0 + col;
    T acc[kTM];
#pragma unroll
    for (int r = 0; r < kTM; ++r) acc[r] = T(0);

    if (k0 < k1) load(k0);
    for (int kc0 = k0; kc0 < k1; kc0 += KC) {
        store(kc0);
        __syncthreads();
        if (kc0 + KC < k1) load(kc0 + KC);     // next chunk in flight during the compute
        const int nk = min(KC, k1 - kc0);
        if (cc < c) {
#pragma unroll 4
            for (int k = ks; k < nk; k += ks_n) {
                const T v = Vs[k * cw + col];
                const T *a = As + k * kTM;
#pragma unroll
                for (int r = 0; r < kTM; ++r) acc[r] = fma(a[r], v, acc[r]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kTM; ++r) red[(r * ks_n + ks) * cw + col] = acc[r];
    __syncthreads();
    for (int o = tid; o < kTM * cw; o += 256) {
        const int r = o / cw, cl = o - r * cw;
        if (n0 + cl >= c) continue;
        T s = T(0);
        for (int j = 0; j < ks_n; ++j) s += red[(r * ks_n + j) * cw + cl];
        const size_t idx = (size_t)(r0 + r) * c + n0 + cl;
        out[idx] = add ? add[idx] - s : s;
    }
}

static int col_chunk(int c) {
    int cw = 1;
    while (cw < c && cw < 64) cw *= 2;
    return cw;
}

template <typename T>
static int launch_gemm(const BandPrec &bp, int c, int tri, const T *A, long long as,
                       const T *V, long long vs, const T *add, long long ads, T *out,
                       long long os, int nblk, hipStream_t s) {
    const int cw = col_chunk(c);
    const int b = bp.b;
    dim3 grid(b / kTM, ceil_div(c, cw), nblk);
    const bool prof = prof_on(XT_PROF_BAND);
    if (prof) prof_begin(XT_PROF_BAND, s);
    hipLaunchKernelGGL(band_gemm_kernel<T>, grid, dim3(256), BandTile<T>::smem, s, b, c, cw, tri,
                       A, as, V, vs, add, ads, out, os);
    if (prof) {
        // matrix elements the tiles stage (triangular: rows r0..r0+7 up to/from the
        // diagonal tile), V read, out written, add read
        const double mat = tri ? (double)b * (b + kTM) / 2 : (double)b * b;
        const double vec = (double)b * c;
        const double bytes = sizeof(T) * nblk * (mat + vec * (add ? 3 : 2));
        prof_end(XT_PROF_BAND, s, bytes, 2.0 * nblk * mat * c);
    }
    return XT_OK;
}

// ---------------------------------------------------------------------------
// Setup kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
zero_upper_kernel(int b, double *L) {   // strict upper triangle of each col-major block
    const long long blk = blockIdx.y;
    double *a = L + blk * (long long)b * b;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < (long long)b * b;
         e += (long long)gridDim.x * 256) {
        const int j = (int)(e / b), i = (int)(e - (long long)j * b);
        if (i < j) a[e] = 0.0;
    }
}

// dst_blk = src_blk^T (b x b blocks), 32 x 32 tiles through LDS.
__global__ void __launch_bounds__(256)
transpose_kernel(int b, const double *__restrict__ src, double *__restrict__ dst) {
    __shared__ double tile[32][33];
    const long long off = (long long)blockIdx.z * b * b;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int j = ty; j < 32; j += 8) tile[j][tx] = src[off + (size_t)(j0 + j) * b + i0 + tx];
    __syncthreads();
    for (int i = ty; i < 32; i += 8) dst[off + (size_t)(i0 + i) * b + j0 + tx] = tile[tx][i];
}

__global__ void __launch_bounds__(256)
to_f32_kernel(long long n, const double *__restrict__ s, float *__restrict__ d) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        d[i] = (float)s[i];
}

template <typename T>
__global__ void __launch_bounds__(256)
fill_zero_kernel(long long n, T *d) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        d[i] = T(0);
}

#define XT_RB(...)                                                                      \
    do {                                                                                \
        rocblas_status st_ = (__VA_ARGS__);                                             \
        if (st_ != rocblas_status_success) {                                            \
            ::xt::set_error("%s:%d %s -> rocblas status %d", __FILE__, __LINE__,        \
                            #__VA_ARGS__, (int)st_);                                    \
            return XT_E_HIP;                                                            \
        }                                                                               \
    } while (0)

struct RbHandle {
    rocblas_handle h = nullptr;
    ~RbHandle() { if (h) rocblas_destroy_handle(h); }
};

static int factor(BandPrec &bp, std::vector<double> &kd, std::vector<double> &ko,
                  hipStream_t s) {
    const int b = bp.b, nb = bp.nb;
    const long long bb = (long long)b * b;
    DevBuf L, E, Dl, G, H, info;
    XT_TRY(L.upload(kd.data(), sizeof(double) * kd.size(), s));
    if (nb > 1) XT_TRY(E.upload(ko.data(), sizeof(double) * ko.size(), s));
    XT_TRY(info.reserve(sizeof(int) * (nb + 1)));
    XT_HIP(hipMemsetAsync(info.ptr, 0, sizeof(int) * (nb + 1), s));
    RbHandle rb;
    XT_RB(rocblas_create_handle(&rb.h));
    XT_RB(rocblas_set_stream(rb.h, s));
    XT_RB(rocblas_set_pointer_mode(rb.h, rocblas_pointer_mode_host));
    const double one = 1.0, mone = -1.0, zero = 0.0;
    double *Lp = L.as<double>(), *Ep = E.as<double>();
    int *ip = info.as<int>();
    for (int i = 0; i < nb; ++i) {
        double *Li = Lp + i * bb;
        if (i > 0)   // K_ii - E_{i-1} E_{i-1}^T
            XT_RB(rocblas_dsyrk(rb.h, rocblas_fill_lower, rocblas_operation_none, b, b, &mone,
                                Ep + (i - 1) * bb, b, &one, Li, b));
        XT_RB(rocsolver_dpotrf(rb.h, rocblas_fill_lower, b, Li, b, ip + i));
        if (i + 1 < nb)  // E_i = K_{i+1,i} L_ii^-T
            XT_RB(rocblas_dtrsm(rb.h, rocblas_side_right, rocblas_fill_lower,
                                rocblas_operation_transpose, rocblas_diagonal_non_unit, b, b,
                                &one, Li, b, Ep + i * bb, b));
    }
    std::vector<int> hinfo(nb + 1);
    XT_HIP(hipMemcpyAsync(hinfo.data(), ip, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    XT_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nb; ++i)
        if (hinfo[i] != 0) {
            set_error("K is not positive definite (block %d, minor %d)", i, hinfo[i]);
            return XT_E_SINGULAR;
        }
    hipLaunchKernelGGL(zero_upper_kernel, dim3(ceil_div(bb, 256 * 8), nb), dim3(256), 0, s, b, Lp);
    XT_HIP(hipGetLastError());
    // D_i = L_ii^-1 in place (column-major lower = D_i^T row-major: the backward set)
    XT_RB(rocsolver_dtrtri_strided_batched(rb.h, rocblas_fill_lower, rocblas_diagonal_non_unit,
                                           b, Lp, b, bb, ip, nb));
    XT_HIP(hipMemcpyAsync(hinfo.data(), ip, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    XT_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nb; ++i)
        if (hinfo[i] != 0) {
            set_error("singular diagonal factor block %d", i);
            return XT_E_SINGULAR;
        }
    XT_TRY(Dl.reserve(sizeof(double) * bb * nb));
    hipLaunchKernelGGL(transpose_kernel, dim3(b / 32, b / 32, nb), dim3(256), 0, s, b, Lp,
                       Dl.as<double>());
    XT_HIP(hipGetLastError());
    if (nb > 1) {
        XT_TRY(G.reserve(sizeof(double) * bb * (nb - 1)));
        XT_TRY(H.reserve(sizeof(double) * bb * (nb - 1)));
        // G_i row-major = (D_{i+1} E_i)^T column-major = E_i^T D_{i+1}^T
        XT_RB(rocblas_dgemm_strided_batched(rb.h, rocblas_operation_transpose,
                                            rocblas_operation_transpose, b, b, b, &one, Ep, b,
                                            bb, Lp + bb, b, bb, &zero, G.as<double>(), b, bb,
                                            nb - 1));
        // H_i row-major = (E_i D_i)^T row-major = E_i D_i column-major
        XT_RB(rocblas_dgemm_strided_batched(rb.h, rocblas_operation_none,
                                            rocblas_operation_none, b, b, b, &one, Ep, b, bb,
                                            Lp, b, bb, &zero, H.as<double>(), b, bb, nb - 1));
    }
    bp.Du[0] = std::move(L);
    bp.Dl[0] = std::move(Dl);
    bp.G[0] = std::move(G);
    bp.H[0] = std::move(H);
    for (int k = 0; k < 4; ++k) {
        DevBuf *src = k == 0 ? bp.Dl : k == 1 ? bp.Du : k == 2 ? bp.G : bp.H;
        if (!src[0].ptr) continue;
        const long long n = (long long)(src[0].bytes / sizeof(double));
        XT_TRY(src[1].reserve(sizeof(float) * n));
        hipLaunchKernelGGL(to_f32_kernel, dim3(4096), dim3(256), 0, s, n, src[0].as<double>(),
                           src[1].as<float>());
        XT_HIP(hipGetLastError());
    }
    XT_HIP(hipStreamSynchronize(s));
    return XT_OK;
}

int band_build(int p, const int *ia, const int *ja, const double *a, int block, BandPrec **out) {
    XT_ARG(p > 0 && ia && ja && a && out, "band prec arguments");
    XT_HIP(hipFuncSetAttribute((const void *)band_gemm_kernel<double>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)BandTile<double>::smem));
    XT_HIP(hipFuncSetAttribute((const void *)band_gemm_kernel<float>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)BandTile<float>::smem));
    XT_ARG(ia[0] == 1, "CSR must be 1-based");
    int bw = 0;
    for (int i = 0; i < p; ++i)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
            const int j = ja[k] - 1;
            XT_ARG(j >= 0 && j < p, "column index out of range");
            if (j <= i) bw = std::max(bw, i - j);
        }
    int b = std::max(std::max(block, bw), 64);
    b = (b + 63) / 64 * 64;
    if (b > 2048) {
        set_error("bandwidth %d exceeds the supported block size 2048 (renumber the mesh)", bw);
        return XT_E_UNSUPPORTED;
    }
    const int nb = (p + b - 1) / b;
    const long long bb = (long long)b * b;
    std::vector<double> kd((size_t)bb * nb, 0.0), ko((size_t)bb * std::max(nb - 1, 0), 0.0);
    for (int i = 0; i < p; ++i)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
            const int j = ja[k] - 1;
            if (j > i) continue;                     // lower triangle only (symmetric K)
            const int bi = i / b, bj = j / b, li = i - bi * b, lj = j - bj * b;
            if (bi == bj) kd[(size_t)bi * bb + (size_t)lj * b + li] = a[k];
            else ko[(size_t)bj * bb + (size_t)lj * b + li] = a[k];   // bi == bj + 1
        }
    for (int i = p; i < nb * b; ++i) {               // identity on the padding rows
        const int li = i - (nb - 1) * b;
        kd[(size_t)(nb - 1) * bb + (size_t)li * b + li] = 1.0;
    }
    BandPrec *bp = new (std::nothrow) BandPrec;
    if (!bp) return XT_E_NOMEM;
    bp->p = p;
    bp->b = b;
    bp->nb = nb;
    bp->bw = bw;
    hipStream_t s = nullptr;
    int rc = factor(*bp, kd, ko, s);
    if (rc != XT_OK) { delete bp; return rc; }
    *out = bp;
    return XT_OK;
}

void band_delete(BandPrec *bp) { delete bp; }

void band_dims(const BandPrec *bp, int *b, int *nb, int *bw) {
    if (b) *b = bp->b;
    if (nb) *nb = bp->nb;
    if (bw) *bw = bp->bw;
}

template <typename T>
int band_apply(BandPrec *bp, int R, int q, const T *minv, const T *x, T *y, hipStream_t s) {
    const int c = q * R, b = bp->b, nb = bp->nb;
    const int pi = sizeof(T) == 4 ? 1 : 0;
    const long long npad = (long long)nb * b * c, nreal = (long long)bp->p * c;
    XT_TRY(bp->w1.reserve(sizeof(T) * npad));
    XT_TRY(bp->w2.reserve(sizeof(T) * npad));
    T *w1 = bp->w1.as<T>(), *w2 = bp->w2.as<T>();
    if (minv) XT_TRY(temporal_mix<T, T>(bp->p, q, R, minv, x, w1, s));
    else XT_HIP(hipMemcpyAsync(w1, x, sizeof(T) * nreal, hipMemcpyDeviceToDevice, s));
    if (npad > nreal)
        hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(ceil_div(npad - nreal, 256)), dim3(256), 0,
                           s, npad - nreal, w1 + nreal);
    const long long bb = (long long)b * b, bc = (long long)b * c;
    const T *Dl = bp->Dl[pi].as<T>(), *Du = bp->Du[pi].as<T>();
    const T *G = bp->G[pi].as<T>(), *H = bp->H[pi].as<T>();
    // forward: w2 = D w1 (all blocks), w2_{i+1} -= G_i w2_i
    XT_TRY(launch_gemm<T>(*bp, c, 1, Dl, bb, w1, bc, nullptr, 0, w2, bc, nb, s));
    for (int i = 0; i + 1 < nb; ++i)
        XT_TRY(launch_gemm<T>(*bp, c, 0, G + i * bb, 0, w2 + i * bc, 0, w2 + (i + 1) * bc, 0,
                              w2 + (i + 1) * bc, 0, 1, s));
    // backward: w1 = D^T w2 (all blocks), w1_i -= H_i w1_{i+1}
    XT_TRY(launch_gemm<T>(*bp, c, 2, Du, bb, w2, bc, nullptr, 0, w1, bc, nb, s));
    for (int i = nb - 2; i >= 0; --i)
        XT_TRY(launch_gemm<T>(*bp, c, 0, H + i * bb, 0, w1 + (i + 1) * bc, 0, w1 + i * bc, 0,
                              w1 + i * bc, 0, 1, s));
    XT_HIP(hipGetLastError());
    XT_HIP(hipMemcpyAsync(y, w1, sizeof(T) * nreal, hipMemcpyDeviceToDevice, s));
    return XT_OK;
}
template int band_apply<double>(BandPrec *, int, int, const double *, const double *, double *,
                                hipStream_t);
template int band_apply<float>(BandPrec *, int, int, const float *, const float *, float *,
                               hipStream_t);

}  // namespace xt
