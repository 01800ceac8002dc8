// Batched right-preconditioned restarted GMRES for space-time slabs.
//
// Restates kronpgmres (kronecker.f90:181-365) for R independent right-hand sides
// that share the spatial matrices (load cases with their own temporal matrices
// A_r, B_r).  Per right-hand side the algorithm is the reference's: residual
// r0 = b - A x0, eps1 = eps*||r0||, Arnoldi with the preconditioner applied on
// the right, Givens rotations on the Hessenberg column, stop when i = im or
// |rs(i+1)| <= eps1, back substitution, x += P^-1 (V y), restart from the
// recombined residual V rs'.  Orthogonalisation is classical Gram-Schmidt with
// one re-orthogonalisation pass (CGS2) instead of the reference's modified
// Gram-Schmidt; both are backward stable, CGS2 needs two fused passes over
// the basis instead of i dependent ones.  The basis, dots and Givens
// arithmetic are FP64.
//
// mode 1 (GMRES-IR): the inner GMRES applies the operator and the ILUT solves
// in FP32 (FP32 values, FP32 accumulation) while keeping the FP64 basis; the
// outer loop recomputes the true residual b - A x with the FP64 operator and
// adds the inner correction, until ||b - A x|| <= eps * ||b - A x0||.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "xt_common.h"

struct xt_op_s;
struct xt_prec_s;

namespace xt {

static constexpr int JB = 8;  // basis vectors per multidot block row

// partial[b][j][col] = sum over the rows of chunk b of V_j[row][col] * w[row][col]; chunk b
// covers rows chunks[b].x .. chunks[b].y - 1 (chunks never straddle two segments).
// 256 / c row lanes of c columns; the block has at least those threads, at most 256.
// skip (device flag, may be NULL): nonzero makes the launch a no-op.  frozen (device,
// per instance, may be NULL): the columns of frozen instances are taken as zero.
__global__ void __launch_bounds__(256)
multidot_partial(int c, int k, const int2 *__restrict__ chunks, const double *__restrict__ V,
                 size_t ldv, const double *__restrict__ w, double *__restrict__ partial,
                 const int *skip, int q, int R, const int *__restrict__ seg_of,
                 const int *__restrict__ frozen) {
    if (skip && *skip) return;   // a second Gram-Schmidt pass found unnecessary
    __shared__ double sh[256];
    const int rpl = 256 / c;
    const int lr = threadIdx.x / c;
    const int col = threadIdx.x - lr * c;
    const int j0 = blockIdx.y * JB;
    const int rbeg = chunks[blockIdx.x].x;
    const int rend = chunks[blockIdx.x].y;
    // a frozen instance's columns are zero: no loads (a chunk lies in one segment)
    const bool off = frozen && frozen[(seg_of ? seg_of[rbeg] * R : 0) + col / q];
    double acc[JB];
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) acc[jj] = 0.0;
    if (lr < rpl && !off) {
        // four rows' loads issued before their products (the row loop is otherwise one
        // latency-bound load pair per iteration); the sums keep the row order
        constexpr int U = 4;
        int row = rbeg + lr;
        for (; row + (U - 1) * rpl < rend; row += U * rpl) {
            double wv[U], vv[U][JB];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const size_t o = (size_t)(row + u * rpl) * c + col;
                wv[u] = w[o];
#pragma unroll
                for (int jj = 0; jj < JB; ++jj)
                    vv[u][jj] = j0 + jj < k ? V[(size_t)(j0 + jj) * ldv + o] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int jj = 0; jj < JB; ++jj)
                    if (j0 + jj < k) acc[jj] += vv[u][jj] * wv[u];
        }
        for (; row < rend; row += rpl) {
            const size_t o = (size_t)row * c + col;
            const double wv = w[o];
#pragma unroll
            for (int jj = 0; jj < JB; ++jj)
                if (j0 + jj < k) acc[jj] += V[(size_t)(j0 + jj) * ldv + o] * wv;
        }
    }
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) {
        if (j0 + jj >= k) break;
        sh[threadIdx.x] = (lr < rpl) ? acc[jj] : 0.0;
        __syncthreads();
        if (threadIdx.x < c) {
            double s = 0.0;
            for (int l = 0; l < rpl; ++l) s += sh[l * c + threadIdx.x];
            partial[((size_t)blockIdx.x * k + j0 + jj) * c + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// out[j][g] = sum over the chunks b of segment s, sum_{i<q} partial[b][j][r*q + i] for
// instance g = s*R + r (segment s = a mesh's rows, r = load case): one workgroup per
// (j, g), its threads stride over the terms, then a tree sum in LDS.  A nonzero skip
// flag writes zeros instead.
__global__ void __launch_bounds__(256)
reduce_partials(int k, int c, int q, int R, int I, const int *__restrict__ seg_chunk,
                const double *__restrict__ partial, double *__restrict__ out,
                const int *skip) {
    __shared__ double sh[256];
    const int j = blockIdx.x, g = blockIdx.y;
    if (skip && *skip) {   // skipped pass: zero coefficients
        if (threadIdx.x == 0) out[(size_t)j * I + g] = 0.0;
        return;
    }
    const int sg = g / R, r = g - sg * R;
    const int b0 = seg_chunk[sg];
    const int n = (seg_chunk[sg + 1] - b0) * q;
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) {
        const int b = b0 + t / q, i = t % q;
        s += partial[((size_t)b * k + j) * c + r * q + i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)j * I + g] = sh[0];
}

// Instance of element idx of an interleaved vector: segment of its row (seg_of, NULL for
// one segment) times R plus its load case.
__device__ __forceinline__ int instance_of(long long idx, int c, int q, int R,
                                           const int *__restrict__ seg_of) {
    const long long row = idx / c;
    const int r = (int)(idx - row * c) / q;
    return (seg_of ? seg_of[row] * R : 0) + r;
}

// y = (yin ? yin : 0) + alpha * sum_j coef[j][g] V_j, g the element's instance; frozen
// (device, may be NULL, only with y == yin): frozen instances' elements are left alone
__global__ void __launch_bounds__(256)
multiaxpy(long long n, int c, int q, int R, int I, const int *__restrict__ seg_of, int k,
          double alpha, const double *__restrict__ coef, const double *__restrict__ V,
          size_t ldv, const double *yin, double *y, const int *skip,
          const int *__restrict__ frozen) {
    if (skip && *skip) return;   // (in place: y keeps its values)
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int g = instance_of(idx, c, q, R, seg_of);
    if (frozen && frozen[g]) return;   // in place on a frozen (zero) column: nothing to do
    double s = yin ? yin[idx] : 0.0;
    for (int j = 0; j < k; ++j) s += (alpha * coef[j * I + g]) * V[(size_t)j * ldv + idx];
    y[idx] = s;
}

// Whether the second Gram-Schmidt pass of a step can be skipped ("twice is enough",
// Kahan-Parlett): after the first pass w' = w - sum_j h_j V_j, |w'|^2 = |w|^2 - sum_j
// h_j^2; when every instance keeps at least half of |w|^2 there was no cancellation
// worth a second pass.  h: k first-pass coefficients then <w, w> as row k, per
// instance.  Sets *skip to 1 (skip) or 0; one workgroup.
__global__ void __launch_bounds__(256)
reorth_check_kernel(int k, int I, const double *__restrict__ h, int *skip) {
    __shared__ int need;
    if (threadIdx.x == 0) need = 0;
    __syncthreads();
    for (int g = threadIdx.x; g < I; g += 256) {
        const double ww = h[(size_t)k * I + g];
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += h[(size_t)j * I + g] * h[(size_t)j * I + g];
        if (ww > 0.0 && !(ww - s >= 0.5 * ww)) atomicOr(&need, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *skip = need ? 0 : 1;
}

// v *= 1/sqrt(nrm2[g]) (0 if the instance is frozen or the norm zero)
__global__ void __launch_bounds__(256)
scale_by_norm(long long n, int c, int q, int R, const int *__restrict__ seg_of,
              const double *__restrict__ nrm2, const int *__restrict__ frozen, double *v) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int g = instance_of(idx, c, q, R, seg_of);
    const double t = sqrt(nrm2[g]);
    v[idx] = (frozen[g] || t == 0.0) ? 0.0 : v[idx] * (1.0 / t);
}

// One Arnoldi step's Givens update per right-hand side (kronecker.f90:281-316), on
// the device so the host checks convergence only every few steps: column i of the
// Hessenberg matrix from the two Gram-Schmidt passes and the new norm, the previous
// rotations applied, a new rotation, the least-squares right-hand side rotated, the
// residual estimate |rs(i+1)| recorded; the right-hand side freezes when its cycle
// ends (i + 1 = K or |rs(i+1)| <= eps1).  Frozen right-hand sides are left alone, so
// steps run past a convergence change nothing.  Same arithmetic as the reference,
// products and sums rounded one by one.
#pragma clang fp contract(off)
__global__ void __launch_bounds__(64)
givens_kernel(int R, int i, int K, const double *__restrict__ hcol,
              const double *__restrict__ hcol2, const double *__restrict__ nrm2, double *hh,
              double *cs, double *sn, double *rs, int *frozen, const double *__restrict__ eps1,
              int *its, int *icyc, double *ro_last, double *hist, double hist_ro0,
              int hist_cap) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R || frozen[r]) return;
    const int it = ++its[r];
    double *H = hh + (size_t)r * (K + 1) * K;
    auto h = [&](int a, int b) -> double & { return H[a + (size_t)b * (K + 1)]; };
    double *c_ = cs + (size_t)r * K, *s_ = sn + (size_t)r * K, *rs_ = rs + (size_t)r * (K + 1);
    for (int j = 0; j <= i; ++j) h(j, i) = hcol[j * R + r] + hcol2[j * R + r];
    h(i + 1, i) = sqrt(nrm2[r]);
    for (int kk = 1; kk <= i; ++kk) {
        const double t = h(kk - 1, i);
        h(kk - 1, i) = c_[kk - 1] * t + s_[kk - 1] * h(kk, i);
        h(kk, i) = -s_[kk - 1] * t + c_[kk - 1] * h(kk, i);
    }
    double gam = sqrt(h(i, i) * h(i, i) + h(i + 1, i) * h(i + 1, i));
    if (gam == 0.0) gam = 1.0e-16;
    c_[i] = h(i, i) / gam;
    s_[i] = h(i + 1, i) / gam;
    rs_[i + 1] = -s_[i] * rs_[i];
    rs_[i] = c_[i] * rs_[i];
    h(i, i) = c_[i] * h(i, i) + s_[i] * h(i + 1, i);
    const double ro = fabs(rs_[i + 1]);
    ro_last[r] = ro;
    if (hist && it < hist_cap) hist[it] = ro / hist_ro0;
    icyc[r] = i + 1;
    if (!(i + 1 < K && ro > eps1[r])) frozen[r] = 1;
}
#pragma clang fp contract(on)

__global__ void __launch_bounds__(256) cvt_f64_f32(long long n, const double *a, float *b) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) b[idx] = (float)a[idx];
}
__global__ void __launch_bounds__(256) cvt_f32_f64(long long n, const float *a, double *b) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) b[idx] = (double)a[idx];
}
template <typename T>
__global__ void __launch_bounds__(256) add_into(long long n, const T *a, double *b) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) b[idx] += (double)a[idx];
}
__global__ void __launch_bounds__(256) sub_from(long long n, const double *b, double *r) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) r[idx] = b[idx] - r[idx];
}

// Partial-pivot LU inverse of a q x q row-major matrix (q <= 8). 0 on success.
static int small_inverse(int q, const double *A, double *Ainv) {
    double lu[64];
    int piv[8];
    std::memcpy(lu, A, sizeof(double) * q * q);
    for (int k = 0; k < q; ++k) {
        int pr = k;
        double mx = std::fabs(lu[k * q + k]);
        for (int i = k + 1; i < q; ++i)
            if (std::fabs(lu[i * q + k]) > mx) { mx = std::fabs(lu[i * q + k]); pr = i; }
        piv[k] = pr;
        if (lu[pr * q + k] == 0.0) return 1;
        if (pr != k)
            for (int j = 0; j < q; ++j) std::swap(lu[k * q + j], lu[pr * q + j]);
        const double rk = 1.0 / lu[k * q + k];
        for (int i = k + 1; i < q; ++i) lu[i * q + k] *= rk;
        for (int i = k + 1; i < q; ++i)
            for (int j = k + 1; j < q; ++j) lu[i * q + j] -= lu[i * q + k] * lu[k * q + j];
    }
    for (int cc = 0; cc < q; ++cc) {
        double e[8];
        for (int i = 0; i < q; ++i) e[i] = (i == cc) ? 1.0 : 0.0;
        for (int k = 0; k < q; ++k) if (piv[k] != k) std::swap(e[k], e[piv[k]]);
        for (int i = 0; i < q; ++i) for (int k = 0; k < i; ++k) e[i] -= lu[i * q + k] * e[k];
        for (int i = q - 1; i >= 0; --i) {
            for (int k = i + 1; k < q; ++k) e[i] -= lu[i * q + k] * e[k];
            e[i] /= lu[i * q + i];
        }
        for (int i = 0; i < q; ++i) Ainv[i * q + cc] = e[i];
    }
    return 0;
}

// Dominant eigenvalue of the q x q row-major C (q <= 8) by power iteration, with its
// right (v) and left (u) eigenvectors scaled so that u . v = 1.  false unless it is a real
// positive eigenvalue the iteration resolves (residual below 1e-10 of it).
static bool dominant_mode(int q, const double *C, double &lam, double *u, double *v) {
    auto power = [&](bool tr, double *x, double &l) -> bool {
        for (int i = 0; i < q; ++i) x[i] = 1.0 + 0.37 * i;   // generic start
        double nx = 0.0;
        for (int i = 0; i < q; ++i) nx += x[i] * x[i];
        for (int i = 0; i < q; ++i) x[i] /= std::sqrt(nx);
        for (int it = 0; it < 5000; ++it) {
            double y[8], ny = 0.0, xy = 0.0;
            for (int i = 0; i < q; ++i) {
                double t = 0.0;
                for (int j = 0; j < q; ++j) t += (tr ? C[j * q + i] : C[i * q + j]) * x[j];
                y[i] = t;
                ny += t * t;
                xy += x[i] * t;
            }
            if (ny == 0.0) return false;
            l = xy;   // Rayleigh quotient (||x|| = 1)
            double res = 0.0;
            for (int i = 0; i < q; ++i) res += (y[i] - l * x[i]) * (y[i] - l * x[i]);
            ny = std::sqrt(ny);
            for (int i = 0; i < q; ++i) x[i] = y[i] / ny;
            if (std::sqrt(res) <= 1e-10 * std::fabs(l)) return l > 0.0;
        }
        return false;
    };
    double lr = 0.0, ll = 0.0;
    if (!power(false, v, lr) || !power(true, u, ll)) return false;
    if (std::fabs(lr - ll) > 1e-8 * lr) return false;
    double d = 0.0;
    for (int i = 0; i < q; ++i) d += u[i] * v[i];
    if (std::fabs(d) < 1e-8) return false;   // (nearly) defective: no stable split
    for (int i = 0; i < q; ++i) u[i] /= d;
    lam = lr;
    return true;
}

// Temporal split of the preconditioner for the cases of one solve (PrecSplit): case r
// goes to the shifted factor whose shift is nearest lambda_r on a log scale, or to none
// when lambda_r is below a quarter of the smallest shift (the mass term is then small
// against K for every mode the shifts were chosen for).  mi (row-major A_r^-1) becomes
// (I - v_r u_r^T) A_r^-1 for a split case.
static int build_split(PrecSplit &sp, int p, int q, int R, const std::vector<double> &tm,
                       std::vector<double> &mi, int nsh, const double *shifts, hipStream_t s) {
    double smin = shifts[0];
    for (int g = 1; g < nsh; ++g) smin = std::min(smin, shifts[g]);
    std::vector<int> grp(R, -1);
    std::vector<double> uu((size_t)R * q), vv((size_t)R * q);
    for (int r = 0; r < R; ++r) {
        const double *Ai = &mi[(size_t)r * q * q];
        const double *B = &tm[((size_t)r * 2 + 1) * q * q];
        double C[64];
        for (int i = 0; i < q; ++i)
            for (int j = 0; j < q; ++j) {
                double t = 0.0;
                for (int k = 0; k < q; ++k) t += Ai[i * q + k] * B[k * q + j];
                C[i * q + j] = t;
            }
        double lam = 0.0;
        if (!dominant_mode(q, C, lam, &uu[(size_t)r * q], &vv[(size_t)r * q])) continue;
        if (lam < 0.25 * smin) continue;
        int best = 0;
        for (int g = 1; g < nsh; ++g)
            if (std::fabs(std::log(lam / shifts[g])) < std::fabs(std::log(lam / shifts[best])))
                best = g;
        grp[r] = best;
    }
    sp.ns = 0;
    sp.goff.assign(1, 0);
    sp.gfac.clear();
    sp.hcase.clear();
    sp.hbase.clear();
    sp.hstride.clear();
    sp.h64.clear();
    for (int g = 0; g < nsh; ++g) {
        int ng = 0;
        for (int r = 0; r < R; ++r) ng += grp[r] == g;
        if (!ng) continue;
        const int o = sp.goff.back();
        int k = 0;
        for (int r = 0; r < R; ++r) {
            if (grp[r] != g) continue;
            const double *u = &uu[(size_t)r * q], *v = &vv[(size_t)r * q];
            double *Ai = &mi[(size_t)r * q * q];
            double m2[8];
            for (int j = 0; j < q; ++j) {
                double t = 0.0;
                for (int kk = 0; kk < q; ++kk) t += u[kk] * Ai[kk * q + j];
                m2[j] = t;
            }
            for (int i = 0; i < q; ++i)
                for (int j = 0; j < q; ++j) Ai[i * q + j] -= v[i] * m2[j];
            sp.hcase.push_back(r);
            sp.hbase.push_back((long long)o * p + k);
            sp.hstride.push_back(ng);
            sp.h64.insert(sp.h64.end(), m2, m2 + q);
            sp.h64.insert(sp.h64.end(), v, v + q);
            ++k;
        }
        sp.goff.push_back(o + ng);
        sp.gfac.push_back(g);
    }
    sp.ns = sp.goff.back();
    // the group structure decides which shifted factors a split apply sweeps and on which
    // columns: captured step graphs are keyed on it (FNV-1a over the two lists)
    sp.sig = 1469598103934665603ULL;
    for (const std::vector<int> *vec : {&sp.goff, &sp.gfac})
        for (int v : *vec) {
            sp.sig ^= (unsigned long long)(unsigned)v;
            sp.sig *= 1099511628211ULL;
        }
    if (!sp.ns) return XT_OK;
    sp.h32.assign(sp.h64.begin(), sp.h64.end());
    XT_TRY(sp.dcase.upload(sp.hcase.data(), sizeof(int) * sp.ns, s));
    XT_TRY(sp.dbase.upload(sp.hbase.data(), sizeof(long long) * sp.ns, s));
    XT_TRY(sp.dstride.upload(sp.hstride.data(), sizeof(int) * sp.ns, s));
    XT_TRY(sp.m2v[0].upload(sp.h64.data(), sizeof(double) * sp.h64.size(), s));
    XT_TRY(sp.m2v[1].upload(sp.h32.data(), sizeof(float) * sp.h32.size(), s));
    return XT_OK;
}

}  // namespace xt

using namespace xt;

namespace xt {
int small_inverse_rowmajor(int q, const double *A, double *Ainv) { return small_inverse(q, A, Ainv); }
static int op_apply_internal(xt_op op, int R, const void *tmat, double alpha, double beta,
                             const void *x, void *y, int fp32, hipStream_t s) {
    return xt_op_apply(op, R, tmat, alpha, beta, x, y, fp32, (void *)s);
}
}

namespace {

struct Solver {
    xt_op op;
    xt_prec pr;
    int p, q, R, c, im;
    long long n;
    hipStream_t s;
    // buffers live in the operator's workspace (reused across calls)
    DevBuf &tm64, &tm32, &mi64, &mi32;   // temporal matrices and inverses
    DevBuf &V;                           // (im+1) x n FP64 basis
    DevBuf &Z;                           // im x n preconditioned basis (flexible variant)
    DevBuf &w32a, &w32b;                 // FP32 work vectors
    DevBuf &z64, &t64;                   // FP64 work vectors
    DevBuf &partial, &hdev, &hdev2, &ndev, &coef;
    DevBuf &gstate, &geps, &ghist;
    PinnedBuf &hstage;
    SolverWorkspace &ws;
    // Segments: nseg independent systems (meshes) stacked block-diagonally, rows
    // seg_dof[s]..seg_dof[s+1]-1; instance g = s*R + r has its own Krylov space,
    // Hessenberg, rotations and convergence.  I = nseg*R instances.
    int nseg = 1, I = 0;
    std::vector<long long> seg_dof;
    const int *segof = nullptr;   // device: segment of each row (NULL for one segment)
    // Packed cycle state, the same byte offsets in gstate (device) and hstage (pinned
    // host): doubles hh | cs | sn | rs | ro, then ints its | icyc | frozen.  A cycle
    // uploads rs..frozen in one copy and reads the whole state back in one.  hstage
    // continues with staging areas for norms, coefficients, targets and the temporal
    // matrices.
    size_t o_hh = 0, o_cs = 0, o_sn = 0, o_rs = 0, o_ro = 0, o_its = 0, o_icyc = 0, o_fz = 0;
    size_t st_bytes = 0, o_nrm = 0, o_coef = 0, o_eps = 0, o_tm = 0, o_mi = 0, o_tm32 = 0,
           o_mi32 = 0;
    double *ghh = nullptr, *gcs = nullptr, *gsn = nullptr, *grs = nullptr, *gro = nullptr;
    int *gits = nullptr, *gicyc = nullptr, *frozen = nullptr;
    explicit Solver(SolverWorkspace &w)
        : tm64(w.tm64), tm32(w.tm32), mi64(w.mi64), mi32(w.mi32), V(w.V), Z(w.Z), w32a(w.w32a),
          w32b(w.w32b), z64(w.z64), t64(w.t64), partial(w.partial), hdev(w.hdev),
          hdev2(w.hdev2), ndev(w.ndev), coef(w.coef), gstate(w.gstate), geps(w.geps),
          ghist(w.ghist), hstage(w.hstage), ws(w) {}
    template <class T> T *dst(size_t off) {
        return reinterpret_cast<T *>(static_cast<char *>(gstate.ptr) + off);
    }
    // Arnoldi steps between host checks of the frozen flags: a few for small systems,
    // where a step costs less than the host round trip; every step for large ones, where
    // a step past convergence would cost more than the check
    int check_every() const {
        const int k = tune_knob(XT_TUNE_GMRES_CHECK);
        return k ? k : (n <= (1LL << 16) ? 4 : 1);
    }
    int nblk = 0, chunk = 0;
    std::vector<double> *hist = nullptr;  // R == 1: ro/ro0 per iteration (kronpgmres log)
    double hist_ro0 = 1.0;
    // temporal split (shifted factors) of the FP64 and FP32 applies, or none
    const PrecSplit *split64 = nullptr, *split32 = nullptr;

    // Active-column compaction of the banded-factor sweeps: the load cases with an
    // instance still iterating in the current cycle, per run of the factor (instances
    // frozen in the cycle have zero basis columns, so their preconditioned columns are
    // zero and need no sweep).  Updated by set_active at the host's checks of the frozen
    // flags; act_version changes with the lists (it keys the step graphs).
    ActiveCols act;
    bool act_on = false, act_full = true;
    std::vector<long long> run_off;
    std::vector<int> act_host;
    long long act_version = -1;   // ws.act_seq when the lists were made (unique per process)
    int set_active(const std::vector<int> &fz) {
        if (!act_on) return XT_OK;
        const int nrun = (int)run_off.size() - 1;
        std::vector<int> lists((size_t)nrun * R), Ra(nrun, 0);
        bool full = true;
        for (int g = 0; g < nrun; ++g) {
            int *l = &lists[(size_t)g * R];
            std::vector<char> on(R, 0);
            for (int sg = 0; sg < nseg; ++sg) {
                const long long lo = nseg > 1 ? seg_dof[sg] : 0, hi = nseg > 1 ? seg_dof[sg + 1] : p;
                if (lo >= run_off[g + 1] || hi <= run_off[g]) continue;   // not in this run
                for (int r = 0; r < R; ++r) on[r] |= !fz[(size_t)sg * R + r];
            }
            int k = 0;
            for (int r = 0; r < R; ++r) if (on[r]) l[k++] = r;
            Ra[g] = k;
            for (int r = 0; r < R; ++r) if (!on[r]) l[k++] = r;
            full = full && Ra[g] == R;
        }
        act_full = full;
        if (full) return XT_OK;
        if (lists == act_host && Ra == act.Ra) return XT_OK;
        act_host = lists;
        act.Ra = Ra;
        XT_TRY(ws.actcols.reserve(sizeof(int) * (size_t)nrun * R));
        XT_HIP(hipMemcpyAsync(ws.actcols.ptr, act_host.data(), sizeof(int) * act_host.size(),
                              hipMemcpyHostToDevice, s));
        XT_HIP(hipStreamSynchronize(s));   // act_host may change at the next check
        act.dcols = ws.actcols.as<int>();
        act_version = ++ws.act_seq;
        return XT_OK;
    }
    const ActiveCols *cur_act() const { return act_on && !act_full ? &act : nullptr; }

    // y = P^-1 x: the temporal inverses (or split mixes) mi64 / mi32, then the factor(s);
    // compact: only the active cases' columns (zero elsewhere)
    int papply64(const double *x, double *y, bool compact = true) {
        return prec_apply_split<double>(pr, R, q, mi64.as<double>(), split64, x, y, s,
                                        compact ? cur_act() : nullptr);
    }
    int papply32(const float *x, float *y, bool compact = true) {
        return prec_apply_split<float>(pr, R, q, mi32.as<float>(), split32, x, y, s,
                                       compact ? cur_act() : nullptr);
    }

    double *Vj(int j) { return V.as<double>() + (size_t)j * n; }
    double *Zj(int j) { return Z.as<double>() + (size_t)j * n; }

    int init(bool flexible) {
        c = q * R;
        n = (long long)p * c;
        I = nseg * R;
        prec_runs(pr, run_off);   // banded factors: compact the sweeps (XT_TUNE_GMRES_COMPACT)
        act_on = !run_off.empty() && R > 1 && tune_knob(XT_TUNE_GMRES_COMPACT) != 1;
        XT_ARG(c <= 256, "q*R must be <= 256 for the batched solver");
        XT_TRY(V.reserve(sizeof(double) * n * (im + 1)));
        if (flexible) XT_TRY(Z.reserve(sizeof(double) * n * im));
        XT_TRY(z64.reserve(sizeof(double) * n));
        XT_TRY(t64.reserve(sizeof(double) * n));
        XT_TRY(w32a.reserve(sizeof(float) * n));
        XT_TRY(w32b.reserve(sizeof(float) * n));
        // row chunks of the dot products: about 1024 over the whole range, each inside one
        // segment
        const int rpl = 256 / c;
        chunk = std::max(rpl, ((p + 1023) / 1024 + rpl - 1) / rpl * rpl);
        if (seg_dof.empty()) seg_dof = {0, (long long)p};
        ws.hchunks.clear();
        ws.hsegchunk.assign(1, 0);
        for (int sg = 0; sg < nseg; ++sg) {
            for (long long r0 = seg_dof[sg]; r0 < seg_dof[sg + 1]; r0 += chunk) {
                ws.hchunks.push_back((int)r0);
                ws.hchunks.push_back((int)std::min(seg_dof[sg + 1], r0 + chunk));
            }
            ws.hsegchunk.push_back((int)(ws.hchunks.size() / 2));
        }
        nblk = (int)(ws.hchunks.size() / 2);
        XT_TRY(ws.chunks.upload(ws.hchunks.data(), sizeof(int) * ws.hchunks.size(), s));
        XT_TRY(ws.segchunk.upload(ws.hsegchunk.data(), sizeof(int) * ws.hsegchunk.size(), s));
        segof = nullptr;
        if (nseg > 1) {
            ws.hsegof.resize(p);
            for (int sg = 0; sg < nseg; ++sg)
                for (long long r0 = seg_dof[sg]; r0 < seg_dof[sg + 1]; ++r0) ws.hsegof[r0] = sg;
            XT_TRY(ws.segof.upload(ws.hsegof.data(), sizeof(int) * p, s));
            segof = ws.segof.as<int>();
        }
        XT_TRY(partial.reserve(sizeof(double) * (size_t)nblk * (im + 1) * c));
        XT_TRY(hdev.reserve(sizeof(double) * (im + 1) * I));
        XT_TRY(hdev2.reserve(sizeof(double) * (im + 1) * I));
        XT_TRY(ndev.reserve(sizeof(double) * I));
        XT_TRY(ws.reorth.reserve(sizeof(int)));
        XT_TRY(coef.reserve(sizeof(double) * (im + 1) * I));
        XT_TRY(geps.reserve(sizeof(double) * I));
        size_t o = 0;
        auto take = [&o](size_t bytes) {
            const size_t at = o;
            o += (bytes + 255) / 256 * 256;
            return at;
        };
        o_hh = take(sizeof(double) * (size_t)I * (im + 1) * im);
        o_cs = take(sizeof(double) * (size_t)I * im);
        o_sn = take(sizeof(double) * (size_t)I * im);
        o_rs = take(sizeof(double) * (size_t)I * (im + 1));
        o_ro = take(sizeof(double) * I);
        o_its = take(sizeof(int) * I);
        o_icyc = take(sizeof(int) * I);
        o_fz = take(sizeof(int) * I);
        st_bytes = o;
        o_nrm = take(sizeof(double) * I);
        o_coef = take(sizeof(double) * (im + 1) * I);
        o_eps = take(sizeof(double) * I);
        o_tm = take(sizeof(double) * (size_t)R * 2 * q * q);
        o_mi = take(sizeof(double) * (size_t)R * q * q);
        o_tm32 = take(sizeof(float) * (size_t)R * 2 * q * q);
        o_mi32 = take(sizeof(float) * (size_t)R * q * q);
        XT_TRY(gstate.reserve(st_bytes));
        XT_TRY(hstage.reserve(o));
        ghh = dst<double>(o_hh);
        gcs = dst<double>(o_cs);
        gsn = dst<double>(o_sn);
        grs = dst<double>(o_rs);
        gro = dst<double>(o_ro);
        gits = dst<int>(o_its);
        gicyc = dst<int>(o_icyc);
        frozen = dst<int>(o_fz);
        return XT_OK;
    }

    // out[j][g] = <V_j, w>_g for j < k and every instance g
    int dots(int k, const double *Vbase, const double *w, double *out,
             const int *skip = nullptr, const int *frz = nullptr) {
        dim3 grid(nblk, ceil_div(k, JB));
        // the kernel's 256 / c row lanes of c columns, rounded up to whole waves: no wave
        // that only waits at the barriers holds a slot (64 of 256 threads at c = 192)
        const int nthr = ((256 / c) * c + 63) / 64 * 64;
        hipLaunchKernelGGL(multidot_partial, grid, dim3(nthr), 0, s, c, k,
                           ws.chunks.as<const int2>(), Vbase, (size_t)n, w, partial.as<double>(),
                           skip, q, R, segof, frz);
        hipLaunchKernelGGL(reduce_partials, dim3(k, I), dim3(256), 0, s, k, c, q, R, I,
                           ws.segchunk.as<int>(), partial.as<double>(), out, skip);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    int maxpy(int k, double alpha, const double *cf, const double *Vbase, const double *yin,
              double *y, const int *skip = nullptr, const int *frz = nullptr) {
        hipLaunchKernelGGL(multiaxpy, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, c, q, R, I,
                           segof, k, alpha, cf, Vbase, (size_t)n, yin, y, skip, frz);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }

    // w = Op(P^-1 v) with FP64 or FP32 operator/preconditioner
    int op_prec(const double *v, double *w, bool f32) {
        if (!f32) {
            XT_TRY(papply64(v, z64.as<double>()));
            XT_TRY(op_apply_internal(op, R, tm64.ptr, 1.0, 0.0, z64.ptr, w, 0, s));
            return XT_OK;
        }
        hipLaunchKernelGGL(cvt_f64_f32, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, v, w32a.as<float>());
        XT_TRY(papply32(w32a.as<float>(), w32b.as<float>()));
        XT_TRY(op_apply_internal(op, R, tm32.ptr, 1.0, 0.0, w32b.ptr, w32a.ptr, 1, s));
        hipLaunchKernelGGL(cvt_f32_f64, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, w32a.as<float>(), w);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    // Flexible step: Z_i = P^-1 v (FP32 or FP64 preconditioner, stored FP64), w = Op Z_i
    // (FP32 or FP64 operator).
    int op_prec_flex(int i, const double *v, double *w, bool f32prec, bool f32op) {
        double *z = Zj(i);
        if (f32prec) {
            hipLaunchKernelGGL(cvt_f64_f32, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, v, w32a.as<float>());
            XT_TRY(papply32(w32a.as<float>(), w32b.as<float>()));
            hipLaunchKernelGGL(cvt_f32_f64, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, w32b.as<float>(), z);
        } else {
            XT_TRY(papply64(v, z));
        }
        if (f32op && f32prec) {   // w32b holds Z_i in FP32 exactly (Z_i is its widening)
            XT_TRY(op_apply_internal(op, R, tm32.ptr, 1.0, 0.0, w32b.ptr, w32a.ptr, 1, s));
            hipLaunchKernelGGL(cvt_f32_f64, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, w32a.as<float>(), w);
        } else if (f32op) {
            hipLaunchKernelGGL(cvt_f64_f32, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, z, w32a.as<float>());
            XT_TRY(op_apply_internal(op, R, tm32.ptr, 1.0, 0.0, w32a.ptr, w32b.ptr, 1, s));
            hipLaunchKernelGGL(cvt_f32_f64, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, w32b.as<float>(), w);
        } else {
            XT_TRY(op_apply_internal(op, R, tm64.ptr, 1.0, 0.0, z, w, 0, s));
        }
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    // x += P^-1 u
    int prec_add(const double *u, double *x, bool f32) {
        // every case of the cycle: those frozen during it need their correction too
        if (!f32) {
            XT_TRY(papply64(u, z64.as<double>(), false));
            hipLaunchKernelGGL(add_into<double>, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, z64.as<double>(), x);
        } else {
            hipLaunchKernelGGL(cvt_f64_f32, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, u, w32a.as<float>());
            XT_TRY(papply32(w32a.as<float>(), w32b.as<float>(), false));
            hipLaunchKernelGGL(add_into<float>, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, w32b.as<float>(), x);
        }
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    // r = b - A x (FP64 operator)
    int residual(const double *b, const double *x, double *r) {
        XT_TRY(op_apply_internal(op, R, tm64.ptr, 1.0, 0.0, x, r, 0, s));
        hipLaunchKernelGGL(sub_from, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, b, r);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    int norms2(const double *v, std::vector<double> &out) {
        XT_TRY(dots(1, v, v, ndev.as<double>()));
        double *h = hstage.at<double>(o_nrm);
        XT_HIP(hipMemcpyAsync(h, ndev.ptr, sizeof(double) * I, hipMemcpyDeviceToHost, s));
        XT_HIP(hipStreamSynchronize(s));
        out.assign(h, h + I);
        return XT_OK;
    }

    // The launches of Arnoldi step i -> i+1 on stream s: w = Op P^-1 V_i (or the flexible
    // form), CGS2 against V_0..V_i, the new norm, V_{i+1} normalised (zeroed for frozen
    // instances), the Givens update of every instance.
    int step_launches(int i, int K, bool flexible, bool f32, bool f32prec, int hist_cap) {
        if (flexible) XT_TRY(op_prec_flex(i, Vj(i), Vj(i + 1), f32prec, f32));
        else XT_TRY(op_prec(Vj(i), Vj(i + 1), f32));
        // CGS2; with selective reorthogonalisation the first pass also takes <w, w> (V_{i+1}
        // follows V_i in the basis) and the second pass runs only where a check on the
        // device finds cancellation (XT_TUNE_GMRES_REORTH: 1 always two passes)
        const bool sel = tune_knob(XT_TUNE_GMRES_REORTH) != 1;
        // (instances frozen in this cycle have zero columns: the CGS2 kernels skip them;
        // the device flags are current here, the Givens update of the last step set them)
        XT_TRY(dots(sel ? i + 2 : i + 1, Vj(0), Vj(i + 1), hdev.as<double>(), nullptr, frozen));
        XT_TRY(maxpy(i + 1, -1.0, hdev.as<double>(), Vj(0), Vj(i + 1), Vj(i + 1), nullptr,
                     frozen));
        const int *skip = nullptr;
        if (sel) {
            hipLaunchKernelGGL(reorth_check_kernel, dim3(1), dim3(256), 0, s, i + 1, I,
                               hdev.as<double>(), ws.reorth.as<int>());
            skip = ws.reorth.as<int>();
        }
        XT_TRY(dots(i + 1, Vj(0), Vj(i + 1), hdev2.as<double>(), skip, frozen));
        XT_TRY(maxpy(i + 1, -1.0, hdev2.as<double>(), Vj(0), Vj(i + 1), Vj(i + 1), skip, frozen));
        XT_TRY(dots(1, Vj(i + 1), Vj(i + 1), ndev.as<double>(), nullptr, frozen));
        // normalise V_{i+1} for every instance still in its cycle at this step (the
        // restart recombination reads it); zero it for the others
        hipLaunchKernelGGL(scale_by_norm, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, c, q, R,
                           segof, ndev.as<double>(), frozen, Vj(i + 1));
        hipLaunchKernelGGL(givens_kernel, dim3(ceil_div(I, 64)), dim3(64), 0, s, I, i, K,
                           hdev.as<double>(), hdev2.as<double>(), ndev.as<double>(), ghh, gcs,
                           gsn, grs, frozen, geps.as<double>(), gits, gicyc, gro,
                           hist ? ghist.as<double>() : nullptr, hist_ro0, hist_cap);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }

    // One Arnoldi step.  With graphs enabled (and no per-launch timing), the step's
    // launches are captured into a HIP graph on the second use of the same step index,
    // variant and buffers, and replayed from it afterwards: the dozens of launches of a
    // step (operator, preconditioner sweeps, CGS2, Givens) cost one graph launch.  The key
    // holds every pointer and size the launches read, so a grown buffer, another operator,
    // preconditioner or split captures anew, and any device allocation or free since the
    // graphs were made drops them (the key cannot list the buffers inside the operator
    // and preconditioner handles).
    int arnoldi_step(int i, int K, bool flexible, bool f32, bool f32prec, int hist_cap) {
        bool plain = !graphs_on();
        for (int k = 0; k < XT_PROF_NKINDS && !plain; ++k) plain = prof_on(k);
        hipStreamCaptureStatus cs_ = hipStreamCaptureStatusNone;
        if (!plain) {
            XT_HIP(hipStreamIsCapturing(s, &cs_));
            plain = cs_ != hipStreamCaptureStatusNone;
        }
        if (plain) return step_launches(i, K, flexible, f32, f32prec, hist_cap);
        // a device buffer allocated or freed anywhere since the cached graphs were made
        // (a preconditioner workspace grown for FP64 sweeps, a handle destroyed) may be
        // one they name: drop them all
        const unsigned long long epoch = g_dev_epoch.load();
        if (epoch != ws.step_epoch) {
            for (SolverWorkspace::StepGraph &old : ws.steps)
                if (old.exec) (void)hipGraphExecDestroy(old.exec);
            ws.steps.clear();
            ws.step_epoch = epoch;
        }
        const std::vector<const void *> key = {
            (const void *)op, (const void *)pr, V.ptr, Z.ptr, w32a.ptr, w32b.ptr, z64.ptr,
            partial.ptr, hdev.ptr, hdev2.ptr, ndev.ptr, gstate.ptr, geps.ptr, ghist.ptr,
            mi64.ptr, mi32.ptr, tm64.ptr, tm32.ptr, ws.chunks.ptr, ws.segchunk.ptr, segof,
            split64, split32, ws.split.dcase.ptr, ws.split.m2v[0].ptr, ws.split.m2v[1].ptr,
            (const void *)(intptr_t)i, (const void *)(intptr_t)K,
            (const void *)(intptr_t)(flexible * 4 + f32 * 2 + f32prec),
            (const void *)(intptr_t)n, (const void *)(intptr_t)I, (const void *)(intptr_t)nblk,
            (const void *)(intptr_t)(hist != nullptr), (const void *)(intptr_t)hist_cap,
            (const void *)(intptr_t)(ws.split.ns), (const void *)(uintptr_t)ws.split.sig,
            (const void *)(intptr_t)R,
            (const void *)(intptr_t)tune_knob(XT_TUNE_GMRES_REORTH), ws.reorth.ptr,
            (const void *)(intptr_t)tune_epoch(),
            (const void *)(intptr_t)(cur_act() ? act_version : -1)};
        SolverWorkspace::StepGraph *g = nullptr;
        for (SolverWorkspace::StepGraph &cand : ws.steps)
            if (cand.key == key) { g = &cand; break; }
        if (!g) {
            if (ws.steps.size() >= 256) {   // a bounded cache: start over
                for (SolverWorkspace::StepGraph &old : ws.steps)
                    if (old.exec) (void)hipGraphExecDestroy(old.exec);
                ws.steps.clear();
            }
            ws.steps.push_back({key, 0, nullptr});
            g = &ws.steps.back();
        }
        if (!g->exec && ++g->uses < 2) return step_launches(i, K, flexible, f32, f32prec, hist_cap);
        if (!g->exec) {
            if (!ws.cap) XT_HIP(hipStreamCreateWithFlags(&ws.cap, hipStreamNonBlocking));
            XT_HIP(hipStreamBeginCapture(ws.cap, hipStreamCaptureModeThreadLocal));
            const hipStream_t saved = s;
            s = ws.cap;
            const int rc = step_launches(i, K, flexible, f32, f32prec, hist_cap);
            s = saved;
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(ws.cap, &graph);
            if (rc != XT_OK || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                if (rc != XT_OK) return rc;
                XT_HIP(e);
            }
            if (g_dev_epoch.load() != ws.step_epoch) {
                // a launch of the step allocated while it was recorded: the graph may name
                // the memory it replaced; the recorded launches never ran, so run them
                (void)hipGraphDestroy(graph);
                for (SolverWorkspace::StepGraph &old : ws.steps)
                    if (old.exec) (void)hipGraphExecDestroy(old.exec);
                ws.steps.clear();
                ws.step_epoch = g_dev_epoch.load();
                return step_launches(i, K, flexible, f32, f32prec, hist_cap);
            }
            const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) g->exec = nullptr;
            XT_HIP(ei);
        }
        XT_HIP(hipGraphLaunch(g->exec, s));
        return XT_OK;
    }

    // GMRES on the residual system already stored in V_0 (unnormalised, = r0).
    // x is updated in place.  eps1[r]: absolute target on |rs(i+1)|.
    // active[r] in: RHS to solve; out: its_r, conv_r (1 converged, 0 not).
    // max_cycles bounds the restart cycles of this call (GMRES-IR inner solves run one).
    // Precision: f32 = FP32 operator and preconditioner; flexible (FGMRES) stores
    // Z_i = P^-1 V_i and updates x += Z y, with the preconditioner in FP32 when f32prec
    // and the operator in FP32 when f32.  nrm0: the squared column norms of V_0 when the
    // caller has just computed them with norms2 (ndev then still holds them), so the
    // first cycle does not compute them again.
    int gmres(double *x, bool f32, const std::vector<double> &eps1, int maxits,
              std::vector<int> &active, std::vector<int> &its, std::vector<int> &conv,
              std::vector<double> &ro_last, int max_cycles = 1 << 30, bool flexible = false,
              bool f32prec = false, const std::vector<double> *nrm0 = nullptr, int kcyc = 0) {
        const int K = (kcyc > 0 && kcyc < im) ? kcyc : im;   // cycle length
        std::vector<int> ncyc(I, 0);
        std::vector<double> hh((size_t)I * (K + 1) * K), cs((size_t)I * K), sn((size_t)I * K),
            rs((size_t)I * (K + 1));
        std::vector<int> icyc(I, 0), fz(I, 0);
        std::vector<double> nrm;
        auto HH = [&](int r, int a, int b) -> double & { return hh[(size_t)r * (K + 1) * K + a + (size_t)b * (K + 1)]; };
        for (int r = 0; r < I; ++r) conv[r] = 0;
        const int hist_cap = maxits + K + 2;
        if (hist) XT_TRY(ghist.reserve(sizeof(double) * hist_cap));
        std::copy(eps1.begin(), eps1.begin() + I, hstage.at<double>(o_eps));
        XT_HIP(hipMemcpyAsync(geps.ptr, hstage.at<double>(o_eps), sizeof(double) * I,
                              hipMemcpyHostToDevice, s));
        std::vector<int> its_before(I);
        for (;;) {
            // start of a cycle: normalise V_0
            if (nrm0) {
                nrm = *nrm0;
                nrm0 = nullptr;
            } else {
                XT_TRY(norms2(Vj(0), nrm));
            }
            bool any = false;
            for (int r = 0; r < I; ++r) {
                if (!active[r]) { fz[r] = 1; continue; }
                const double ro = std::sqrt(nrm[r]);
                ro_last[r] = ro;
                if (hist && its[r] == 0) hist->push_back(ro / hist_ro0);
                if (ro == 0.0 || ro <= eps1[r]) { conv[r] = 1; active[r] = 0; fz[r] = 1; continue; }
                fz[r] = 0;
                any = true;
                rs[(size_t)r * (K + 1)] = ro;
                icyc[r] = 0;
            }
            if (!any) break;
            XT_TRY(set_active(fz));
            // the cycle's Givens state lives on the device; the host reads the frozen
            // flags every kCheck steps and the whole state at the end of the cycle
            its_before = its;
            // rs | ro | its | icyc | frozen in one copy (the previous copies that read this
            // staging area finished at the last synchronisation)
            std::copy(rs.begin(), rs.end(), hstage.at<double>(o_rs));
            std::copy(ro_last.begin(), ro_last.begin() + I, hstage.at<double>(o_ro));
            std::copy(its.begin(), its.begin() + I, hstage.at<int>(o_its));
            std::copy(icyc.begin(), icyc.end(), hstage.at<int>(o_icyc));
            std::copy(fz.begin(), fz.end(), hstage.at<int>(o_fz));
            XT_HIP(hipMemcpyAsync(grs, hstage.at<char>(o_rs), st_bytes - o_rs,
                                  hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(scale_by_norm, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, c, q,
                               R, segof, ndev.as<double>(), frozen, Vj(0));
            int i = 0;
            for (;;) {
                // Arnoldi step i -> i+1, replayed from a HIP graph when one is captured
                XT_TRY(arnoldi_step(i, K, flexible, f32, f32prec, hist_cap));
                if ((i + 1) % check_every() == 0 || i + 1 == K) {
                    // the whole state in one copy: the cycle ends at a check where every
                    // case is frozen (all are at i + 1 == K), so this is then its final state
                    XT_HIP(hipMemcpyAsync(hstage.ptr, gstate.ptr, st_bytes, hipMemcpyDeviceToHost,
                                          s));
                    XT_HIP(hipStreamSynchronize(s));
                    std::copy(hstage.at<int>(o_fz), hstage.at<int>(o_fz) + I, fz.begin());
                    bool cont = false;
                    for (int r = 0; r < I; ++r) cont = cont || !fz[r];
                    if (!cont) break;
                    XT_TRY(set_active(fz));
                }
                ++i;
            }
            {   // the state read at the last check
                const double *h = hstage.at<double>(o_hh);
                std::copy(h, h + hh.size(), hh.begin());
                h = hstage.at<double>(o_cs);
                std::copy(h, h + cs.size(), cs.begin());
                h = hstage.at<double>(o_sn);
                std::copy(h, h + sn.size(), sn.begin());
                h = hstage.at<double>(o_rs);
                std::copy(h, h + rs.size(), rs.begin());
                h = hstage.at<double>(o_ro);
                std::copy(h, h + I, ro_last.begin());
                const int *hi = hstage.at<int>(o_icyc);
                std::copy(hi, hi + I, icyc.begin());
                hi = hstage.at<int>(o_its);
                std::copy(hi, hi + I, its.begin());
            }
            if (hist && its[0] > its_before[0]) {
                std::vector<double> hv(its[0] - its_before[0]);
                XT_HIP(hipMemcpy(hv.data(), ghist.as<double>() + its_before[0] + 1,
                                 sizeof(double) * hv.size(), hipMemcpyDeviceToHost));
                hist->insert(hist->end(), hv.begin(), hv.end());
            }
            // back substitution and solution update, per RHS with its own cycle length
            const int kmax = i + 1;
            std::vector<double> y((size_t)kmax * I, 0.0), rr((size_t)(kmax + 1) * I, 0.0);
            for (int r = 0; r < I; ++r) {
                if (!active[r] || icyc[r] == 0) continue;
                const int ir = icyc[r];
                double *rs_ = &rs[(size_t)r * (K + 1)];
                std::vector<double> yy(rs_, rs_ + ir);
                yy[ir - 1] = yy[ir - 1] / HH(r, ir - 1, ir - 1);
                for (int kk = ir - 2; kk >= 0; --kk) {
                    double t = yy[kk];
                    for (int j = kk + 1; j < ir; ++j) t -= HH(r, kk, j) * yy[j];
                    yy[kk] = t / HH(r, kk, kk);
                }
                for (int j = 0; j < ir; ++j) y[(size_t)j * I + r] = yy[j];
            }
            std::copy(y.begin(), y.end(), hstage.at<double>(o_coef));
            XT_HIP(hipMemcpyAsync(coef.ptr, hstage.at<double>(o_coef), sizeof(double) * kmax * I,
                                  hipMemcpyHostToDevice, s));
            if (flexible) {
                XT_TRY(maxpy(kmax, 1.0, coef.as<double>(), Zj(0), x, x));
            } else {
                XT_TRY(maxpy(kmax, 1.0, coef.as<double>(), Vj(0), nullptr, t64.as<double>()));
                XT_TRY(prec_add(t64.as<double>(), x, f32));
            }
            // decide per RHS; build restart residual V_0 = sum_j rs'_j V_j
            bool again = false;
            for (int r = 0; r < I; ++r) {
                if (!active[r]) continue;
                const int ir = icyc[r];
                if (ro_last[r] <= eps1[r]) { conv[r] = 1; active[r] = 0; continue; }
                if (its[r] >= maxits || ++ncyc[r] >= max_cycles) { active[r] = 0; continue; }
                double *rs_ = &rs[(size_t)r * (K + 1)], *c_ = &cs[(size_t)r * K], *s_ = &sn[(size_t)r * K];
                for (int j = 1; j <= ir; ++j) {
                    const int jj = ir + 1 - j + 1;  // 1-based as the reference
                    rs_[jj - 2] = -s_[jj - 2] * rs_[jj - 1];
                    rs_[jj - 1] = c_[jj - 2] * rs_[jj - 1];
                }
                for (int j = 0; j <= ir; ++j) rr[(size_t)j * I + r] = rs_[j];
                again = true;
            }
            if (!again) break;
            XT_HIP(hipStreamSynchronize(s));   // y's copy read the staging area rr reuses
            std::copy(rr.begin(), rr.end(), hstage.at<double>(o_coef));
            XT_HIP(hipMemcpyAsync(coef.ptr, hstage.at<double>(o_coef),
                                  sizeof(double) * (kmax + 1) * I, hipMemcpyHostToDevice, s));
            XT_TRY(maxpy(kmax + 1, 1.0, coef.as<double>(), Vj(0), nullptr, t64.as<double>()));
            XT_HIP(hipMemcpyAsync(Vj(0), t64.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            XT_HIP(hipStreamSynchronize(s));
        }
        return XT_OK;
    }
};

}  // namespace

extern "C" int xt_slab_solve(xt_op op, xt_prec pr, int R, const double *ab, const double *b,
                             double *x, int im, double eps, int maxits, int mode, int *its,
                             int *ierr, double *relres, void *stream) {
    return xt::slab_solve_impl(op, pr, R, ab, b, x, im, eps, maxits, mode, its, ierr, relres,
                               (hipStream_t)stream, nullptr);
}

extern "C" int xt_slab_solve_tol(xt_op op, xt_prec pr, int R, const double *ab, const double *b,
                                 double *x, int im, double eps, double xtol, int maxits,
                                 int mode, int *its, int *ierr, double *relres, void *stream) {
    if (xtol < 0.0) {
        set_error("argument error: xtol must be >= 0");
        return XT_E_ARG;
    }
    return xt::slab_solve_impl(op, pr, R, ab, b, x, im, eps, maxits, mode, its, ierr, relres,
                               (hipStream_t)stream, nullptr, xtol);
}

extern "C" int xt_slab_stats(xt_op op, long long *it32, long long *it64) {
    XT_ARG(op, "slab_stats: op");
    SolverWorkspace *ws = op_workspace(op);
    if (it32) *it32 = ws->last_it32;
    if (it64) *it64 = ws->last_it64;
    return XT_OK;
}

extern "C" int xt_slab_solve_seg(xt_op op, xt_prec pr, int nseg, const long long *seg_dof,
                                 int R, const double *ab, const double *b, double *x, int im,
                                 double eps, double xtol, int maxits, int mode, int *its,
                                 int *ierr, double *relres, void *stream) {
    if (xtol < 0.0 || nseg < 1 || !seg_dof) {
        set_error("argument error: xtol >= 0, nseg >= 1 and seg_dof are required");
        return XT_E_ARG;
    }
    return xt::slab_solve_impl(op, pr, R, ab, b, x, im, eps, maxits, mode, its, ierr, relres,
                               (hipStream_t)stream, nullptr, xtol, nseg, seg_dof);
}

int xt::slab_solve_impl(xt_op op, xt_prec pr, int R, const double *ab, const double *b,
                        double *x, int im, double eps, int maxits, int mode, int *its,
                        int *ierr, double *relres, hipStream_t stream,
                        std::vector<double> *hist, double xtol, int nseg,
                        const long long *seg_dof) {
    XT_ARG(op && pr && R > 0 && ab && b && x && its && ierr && relres, "slab_solve arguments");
    XT_ARG(im >= 1 && im <= 256 && maxits >= 1 && mode >= 0 && mode <= 2, "slab_solve parameters");
    SolverWorkspace *ws = op_workspace(op);
    if (!ws) return XT_E_NOMEM;
    Solver S(*ws);
    S.op = op;
    S.pr = pr;
    S.R = R;
    S.im = im;
    S.s = stream;
    S.hist = (R == 1 && nseg == 1 && mode == 0) ? hist : nullptr;
    {
        int p = 0, q = 0, nt = 0;
        XT_TRY(op_dims(op, &p, &q, &nt));
        XT_ARG(nt == 2, "slab_solve needs a two-term (K, M) operator");
        XT_ARG(prec_rows(pr) == p, "the preconditioner and the operator differ in size");
        S.p = p;
        S.q = q;
    }
    if (nseg > 1 || seg_dof) {
        XT_ARG(nseg >= 1 && seg_dof && seg_dof[0] == 0 && seg_dof[nseg] == S.p,
               "segments must cover the rows 0..p");
        for (int sg = 0; sg < nseg; ++sg)
            XT_ARG(seg_dof[sg + 1] > seg_dof[sg], "segments must be non-empty and increasing");
        S.nseg = nseg;
        S.seg_dof.assign(seg_dof, seg_dof + nseg + 1);
    }
    XT_TRY(S.init(mode != 0));
    const int q = S.q;
    // temporal matrices: ab[r] = (A_r, B_r) column-major -> row-major T; inverse of A_r
    std::vector<double> tm((size_t)R * 2 * q * q), mi((size_t)R * q * q);
    for (int r = 0; r < R; ++r) {
        for (int t = 0; t < 2; ++t)
            for (int i = 0; i < q; ++i)
                for (int j = 0; j < q; ++j)
                    tm[(((size_t)r * 2 + t) * q + i) * q + j] = ab[((size_t)r * 2 + t) * q * q + i + j * q];
        if (small_inverse(q, &tm[(size_t)r * 2 * q * q], &mi[(size_t)r * q * q])) {
            set_error("temporal matrix A of right-hand side %d is numerically singular", r);
            return XT_E_SINGULAR;
        }
    }
    // shifted factors: each case's dominant temporal mode goes to the nearest shift; the
    // split mixes replace A_r^-1 for the applies whose precision the factors hold
    std::vector<double> mis(mi);
    {
        const double *shifts = nullptr;
        bool fp64 = false;
        const int nsh = prec_shifts(pr, &shifts, &fp64);
        if (nsh > 0) {
            XT_HIP(hipStreamSynchronize(S.s));   // the previous split's uploads have finished
            XT_TRY(build_split(ws->split, S.p, q, R, tm, mis, nsh, shifts, S.s));
            if (ws->split.ns) {
                S.split32 = &ws->split;
                if (fp64) {
                    S.split64 = &ws->split;
                    mi = mis;
                }
            }
        }
    }
    std::vector<float> tm32(tm.begin(), tm.end()), mi32(mis.begin(), mis.end());
    std::copy(tm.begin(), tm.end(), S.hstage.at<double>(S.o_tm));
    std::copy(mi.begin(), mi.end(), S.hstage.at<double>(S.o_mi));
    std::copy(tm32.begin(), tm32.end(), S.hstage.at<float>(S.o_tm32));
    std::copy(mi32.begin(), mi32.end(), S.hstage.at<float>(S.o_mi32));
    XT_TRY(S.tm64.upload(S.hstage.at<double>(S.o_tm), sizeof(double) * tm.size(), S.s));
    XT_TRY(S.mi64.upload(S.hstage.at<double>(S.o_mi), sizeof(double) * mi.size(), S.s));
    XT_TRY(S.tm32.upload(S.hstage.at<float>(S.o_tm32), sizeof(float) * tm32.size(), S.s));
    XT_TRY(S.mi32.upload(S.hstage.at<float>(S.o_mi32), sizeof(float) * mi32.size(), S.s));

    // FP64 preconditioner sweeps available (a factor kept in FP32 only has none: no
    // FP64 fallback then)
    const bool fp64_prec = prec_has_fp64(pr);
    std::vector<int> active(S.I, 1), itv(S.I, 0), conv(S.I, 0);
    std::vector<double> ro_last(S.I, 0.0), eps1(S.I, 0.0), r0(S.I), bn(S.I);
    // V_0 = b - A x0; when every x0 is zero that is b itself (A 0 = 0 exactly), so the
    // operator apply and a second norm pass are skipped (ndev then holds b's norms,
    // which are V_0's)
    std::vector<double> xn;
    XT_TRY(S.norms2(x, xn));
    bool x0zero = true;
    for (int r = 0; r < S.I; ++r) x0zero = x0zero && xn[r] == 0.0;
    XT_TRY(S.norms2(b, bn));
    S.hist_ro0 = std::sqrt(bn[0]);
    if (x0zero) {
        XT_HIP(hipMemcpyAsync(S.Vj(0), b, sizeof(double) * S.n, hipMemcpyDeviceToDevice, S.s));
        r0 = bn;
    } else {
        XT_TRY(S.residual(b, x, S.Vj(0)));
        XT_TRY(S.norms2(S.Vj(0), r0));
    }
    const std::vector<double> r0sq = r0;   // squared norms of V_0 (ndev still holds them)
    for (int r = 0; r < S.I; ++r) {
        r0[r] = std::sqrt(r0[r]);
        eps1[r] = eps * r0[r];
        ierr[r] = 0;
        if (r0[r] == 0.0) { ierr[r] = -1; active[r] = 0; }
    }
    ws->last_it32 = ws->last_it64 = 0;
    auto itsum = [&]() {
        long long t = 0;
        for (int r = 0; r < S.I; ++r) t += itv[r];
        return t;
    };
    if (mode == 0 || mode == 2) {
        // mode 2: FP64 flexible GMRES, FP32 preconditioner
        XT_TRY(S.gmres(x, false, eps1, maxits, active, itv, conv, ro_last, 1 << 30, mode == 2,
                       mode == 2));
        (mode == 2 ? ws->last_it32 : ws->last_it64) = itsum();
        for (int r = 0; r < S.I; ++r) if (ierr[r] == 0 && !conv[r]) ierr[r] = 1;
    } else {
        // GMRES-IR: one flexible GMRES cycle per refinement step on the current residual
        // with the FP32 operator and FP32 preconditioner (FP64 basis, stored
        // preconditioned vectors), FP64 outer residual.  The inner target starts at 1e-3
        // of the residual (one inner iteration a step with an exact factor) and is
        // relaxed to the reduction the last step achieved when FP32 rounding is the floor
        // (ill-conditioned temporal matrices).  A right-hand side continues with inner
        // cycles in FP64 (operator and preconditioner), where FP32 does not pay, when its
        // step reduced the true residual by less than 16x (the XT_TUNE_IR_MIN_GAIN knob:
        // an FP64 cycle with the exact factor gains orders of magnitude per iteration, an
        // FP32 iteration costs about half of one), or when its FP32 cycle did not meet
        // the inner target within kFp32Cycle iterations (2, or the XT_TUNE_IR_FP32_CYCLE
        // knob; with an exact factor of a well-conditioned K one iteration meets it; an
        // FP32 factor of an
        // ill-conditioned K preconditions poorly: FP64 cycles then need a few iterations
        // where FP32 ones need tens).
        // With xtol > 0 a case also needs a forward-error test (the residual eps does not
        // bound the error of an ill-conditioned slab): see etight below.
        std::vector<double> rn(S.I), irel(S.I, 1.0e-3), prev(S.I, 0.0), dn(S.I, 0.0), xn(S.I, 0.0);
        std::vector<int> fb(S.I, 0);
        // Forward-error target xtol > 0 with FP64 sweeps at hand: a case converges when its
        // residual reaches etight = 1e-5 xtol of the initial residual (no tighter than
        // 1e-14): FP32 refinement steps keep going while each gains at least gmin32 (the
        // FP64 residual makes their progress real), FP64 cycles aim there in one long
        // Krylov cycle.  The
        // last correction does not bound the error here: FP32 cycles of an ill-conditioned
        // slab leave the modes the FP32 factor resolves poorly almost untouched, so their
        // corrections are small while the error is not (the thin panel: corrections below
        // 1e-7 of the solution, error 1.5e-5); its slab operator turns a relative residual
        // into up to ~2.5e4 times that error, which 1e-5 xtol covers.  With an FP32-only
        // factor the last-correction test is the one available.
        std::vector<double> etight(S.I);
        for (int r = 0; r < S.I; ++r) etight[r] = std::max(1.0e-5 * xtol, 1.0e-14) * r0[r];
        // Whether FP32 cycles pay is measured, per operator and preconditioner: the first
        // solve runs FP32 cycles first (probe A), the next one FP64 cycles only (probe B),
        // and the following solves take whichever was faster by wall time, until the
        // next probe pair after XT_TUNE_IR_SKIP32 = n solves in all (built-in 16; 1: FP32
        // first always).  An FP32 factor of an ill-conditioned K (the thin panel) can
        // leave FP32 cycles a net loss even when each one converges.
        const int nskip = tune_knob(XT_TUNE_IR_SKIP32) ? tune_knob(XT_TUNE_IR_SKIP32) : 16;
        const bool probing = fp64_prec && nskip > 1;
        if (probing && ws->ir_prec != prec_id(pr)) {
            ws->ir_prec = prec_id(pr);
            ws->ir_phase = 0;
        }
        const bool fp64only = probing && (ws->ir_phase == 1 || (ws->ir_phase == 2 && ws->ir_fp64));
        if (fp64only) std::fill(fb.begin(), fb.end(), 1);
        const auto tstart = std::chrono::steady_clock::now();
        const double gmin32 = tune_knob(XT_TUNE_IR_MIN_GAIN) ? tune_knob(XT_TUNE_IR_MIN_GAIN) : 16;
        if (xtol > 0.0 && !fp64_prec) XT_TRY(ws->xprev.reserve(sizeof(double) * S.n));
        for (int outer = 0;; ++outer) {
            const bool dtest = xtol > 0.0 && !fp64_prec;   // the last-correction test
            if (dtest && outer > 0) {   // correction of the last step: x - xprev
                double *d = ws->xprev.as<double>();
                hipLaunchKernelGGL(sub_from, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.s, S.n, x, d);
                XT_HIP(hipGetLastError());
                XT_TRY(S.norms2(d, dn));
                XT_TRY(S.norms2(x, xn));
            }
            // V_0 holds b - A x for the current x; its norms give the true residual (the
            // first pass has them from the setup).  They are computed last: the first
            // inner cycle normalises V_0 by the norms the device still holds.
            if (outer == 0) rn = r0sq;
            else XT_TRY(S.norms2(S.Vj(0), rn));
            if (outer == 64) break;
            bool any32 = false, any64 = false;
            std::vector<double> e1(S.I, 0.0);
            std::vector<int> act32(S.I, 0), act64(S.I, 0);
            for (int r = 0; r < S.I; ++r) {
                if (ierr[r] == -1) continue;
                const double nr = std::sqrt(rn[r]);
                const bool fwd = xtol == 0.0 ||
                                 (fp64_prec ? nr <= etight[r]
                                            : outer > 0 && dn[r] <= xtol * xtol * xn[r]);
                if (nr <= eps1[r] && fwd) { conv[r] = 1; continue; }
                conv[r] = 0;
                if (itv[r] >= maxits || nr == 0.0) continue;
                if (prev[r] > 0.0) {
                    const double rho = nr / prev[r];
                    if (rho * gmin32 > 1.0 && fp64_prec && !fb[r]) {
                        fb[r] = 1;
                        irel[r] = 1.0e-3;
                    } else if (!fb[r] && rho > 10.0 * irel[r]) {
                        irel[r] = std::min(0.1, rho);   // the FP32 floor
                    }
                }
                prev[r] = nr;
                // FP64 cycles after a fallback aim at the final target at once (one long
                // Krylov cycle converges faster than several restarted ones)
                e1[r] = fb[r] && xtol > 0.0 && fp64_prec ? etight[r]
                        : nr <= eps1[r]         ? irel[r] * nr
                        : fb[r]                 ? 0.5 * eps1[r]
                                                : std::max(irel[r] * nr, 0.5 * eps1[r]);
                if (fb[r]) { act64[r] = 1; any64 = true; }
                else { act32[r] = 1; any32 = true; }
            }
            if (!any32 && !any64) break;
            if (dtest)
                XT_HIP(hipMemcpyAsync(ws->xprev.ptr, x, sizeof(double) * S.n,
                                      hipMemcpyDeviceToDevice, S.s));
            // inner solves for the correction on V_0 = the residual (already there): each
            // flexible cycle adds its correction Z y to x itself
            std::vector<int> cin(S.I, 0);
            // the first inner call starts from the norms just computed (V_0 unchanged)
            if (any32) {
                const int kFp32Cycle = tune_knob(XT_TUNE_IR_FP32_CYCLE) ? tune_knob(XT_TUNE_IR_FP32_CYCLE) : 2;
                const std::vector<int> was32 = act32;   // the call clears the active flags
                const long long t0 = itsum();
                XT_TRY(S.gmres(x, true, e1, maxits, act32, itv, cin, ro_last, 1, true, true, &rn,
                               kFp32Cycle));
                ws->last_it32 += itsum() - t0;
                for (int r = 0; r < S.I; ++r)
                    if (was32[r] && !cin[r] && fp64_prec && !fb[r]) {
                        fb[r] = 1;
                        irel[r] = 1.0e-3;   // FP64 cycles reach the full reduction
                    }
            }
            if (any64) {
                const long long t0 = itsum();
                XT_TRY(S.gmres(x, false, e1, maxits, act64, itv, cin, ro_last, 1, true, false,
                               any32 ? nullptr : &rn));
                ws->last_it64 += itsum() - t0;
            }
            XT_TRY(S.residual(b, x, S.Vj(0)));
        }
        if (probing) {
            const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                           tstart).count();
            if (ws->ir_phase == 0) {
                ws->ir_cost_a = t;
                ws->ir_phase = 1;
            } else if (ws->ir_phase == 1) {
                ws->ir_fp64 = t < ws->ir_cost_a;
                ws->ir_phase = 2;
                ws->ir_left = nskip - 2;
            } else if (--ws->ir_left <= 0) {
                ws->ir_phase = 0;
            }
        }
        for (int r = 0; r < S.I; ++r) if (ierr[r] == 0 && !conv[r]) ierr[r] = 1;
        for (int r = 0; r < S.I; ++r) {   // true residual: the loop's last b - A x
            its[r] = itv[r];
            relres[r] = bn[r] > 0 ? std::sqrt(rn[r]) / std::sqrt(bn[r]) : 0.0;
        }
        return XT_OK;
    }
    // true residual
    std::vector<double> rn;
    XT_TRY(S.residual(b, x, S.t64.as<double>()));
    XT_TRY(S.norms2(S.t64.as<double>(), rn));
    for (int r = 0; r < S.I; ++r) {
        its[r] = itv[r];
        relres[r] = bn[r] > 0 ? std::sqrt(rn[r]) / std::sqrt(bn[r]) : 0.0;
    }
    return XT_OK;
}
