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
#include <atomic>
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

// Both sweeps in one workgroup without a barrier per level: each wave takes rows in
// level order (tasks w, w + nwaves, ... of the level-ordered rows of L, then of U) and
// starts a row as soon as the rows it reads are finished, told by per-row flags in LDS
// (1: forward result, 2: backward result).  Every row depends only on rows earlier in
// that order and each wave walks its rows in that order, so the earliest unfinished
// row can always proceed.  Each row's leading entries sit in a fixed-size slot in
// level order, so a wave loads its next two rows' entries with no dependent address
// while it waits for and computes the current one; the solution, the inverse diagonal
// and the flags live in LDS.  Lanes: S slices of the row's entries times 64/S column
// groups; the slices' partial sums are reduced by DPP lane moves in a fixed order.  Waits are bounded: a wave that waits past kSpinMax polls raises an abort
// flag that releases every wave, and the output is then NaN.
constexpr int kSlot = 32;           // leading entries of a row kept in its fixed-size slot
constexpr int kSfMaxRows = 4096;    // slots are built for factors up to this order
constexpr int kSpinMax = 1 << 22;

// LDS: x [(n + 1) c] (row n stays zero), inverse diagonal [n], flags [n + 1] (flag[n] = 3,
// always satisfied), row pointers of L and U [n + 1] each, abort flag.
template <typename T>
static size_t sf_lds_bytes(int n, int c) {
    return sizeof(T) * ((size_t)(n + 1) * c + n) + sizeof(int) * (3 * (size_t)n + 4);
}

// Sum over aligned groups of S lanes (S = 4 or 8) with DPP lane moves: swap
// neighbours, swap pairs within a quad, then mirror within the half-row of 8.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int S, typename T>
__device__ __forceinline__ T group_sum(T s) {
    if (S >= 2) s += dpp_move<0xB1>(s);    // quad_perm [1,0,3,2]
    if (S >= 4) s += dpp_move<0x4E>(s);    // quad_perm [2,3,0,1]
    if (S >= 8) s += dpp_move<0x141>(s);   // row_half_mirror
    return s;
}

__device__ __forceinline__ int lds_flag(const int *f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One triangle as the single-workgroup sweep reads it (device pointers): rows in level
// order (bit 30 set when the row has more than kSlot entries), the first kSlot entries
// of each of those rows in its slot (columns past the row are n, values 0), and the
// CSR for the entries past the slot.
struct SfTri {
    const int *trow, *sci, *rp, *ci;
    const void *sv, *v;
};

template <typename T, int S>
__global__ void __launch_bounds__(1024)
trsv_sf_kernel(int n, int c, SfTri L, SfTri U, const T *__restrict__ dinv,
               const T *__restrict__ b, T *__restrict__ xg) {
    constexpr int E = kSlot / S;    // slot entries per lane
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *x = reinterpret_cast<T *>(smem_raw);
    T *sdinv = x + (size_t)(n + 1) * c;
    int *flag = reinterpret_cast<int *>(sdinv + n);
    int *spL = flag + n + 1, *spU = spL + n + 1;
    int *abort_flag = spU + n + 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n * c; i += nt) x[i] = b[i];
    for (int i = tid; i < c; i += nt) x[(size_t)n * c + i] = T(0);
    for (int i = tid; i < n; i += nt) {
        sdinv[i] = dinv[i];
        flag[i] = 0;
    }
    for (int i = tid; i <= n; i += nt) {
        spL[i] = L.rp[i];
        spU[i] = U.rp[i];
    }
    if (tid == 0) {
        flag[n] = 3;
        *abort_flag = 0;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6, ntask = 2 * n;
    const int slice = lane % S, cg = lane / S, ncg = 64 / S;
    // task t < n: forward row of level-ordered position t; t >= n: backward row at t - n.
    // Branch-free loads (t clamped to the last task) with no use here, so they stay in
    // flight while the wave works on earlier tasks.
    auto task_load = [&](int t, int &row, T *pv, int *pc) {
        t = t < ntask ? t : ntask - 1;
        const bool up = t >= n;
        const int k = up ? t - n : t;
        const int *trow = up ? U.trow : L.trow, *sci = up ? U.sci : L.sci;
        const T *sv = static_cast<const T *>(up ? U.sv : L.sv);
        row = trow[k];
        const size_t base = (size_t)k * kSlot + slice;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            pv[j] = sv[base + j * S];
            pc[j] = sci[base + j * S];
        }
    };
    auto process = [&](int t, int rowf, const T *pv, const int *pc) {
        const bool up = t >= n;
        const int row = rowf & 0x3fffffff;
        const bool longrow = (rowf >> 30) != 0;
        const int need = up ? 2 : 1;
        const int *rp = up ? spU : spL;
        const int *ci = up ? U.ci : L.ci;
        const T *v = static_cast<const T *>(up ? U.v : L.v);
        // wait for the rows this one reads (all flag reads issued together)
        for (int spins = 0;; ++spins) {
            int f[E];
#pragma unroll
            for (int j = 0; j < E; ++j) f[j] = lds_flag(flag + pc[j]);
            const int fo = up ? lds_flag(flag + row) : 1;
            const int ab = lds_flag(abort_flag);
            bool ok = fo >= 1;
#pragma unroll
            for (int j = 0; j < E; ++j) ok = ok && f[j] >= need;
            if (longrow)
                for (int e = rp[row] + kSlot + slice; e < rp[row + 1]; e += S)
                    ok = ok && lds_flag(flag + ci[e]) >= need;
            if (__all(ok) || ab) break;
            if (spins >= kSpinMax) {
                __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                break;
            }
        }
        // a wave's LDS operations complete in order: a compiler fence keeps the x reads
        // after the flag reads (a workgroup fence would also wait for the prefetches)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        for (int col = cg; col < c; col += ncg) {
            const size_t o = (size_t)row * c + col;
            const T xo = x[o];
            T s = 0;
#pragma unroll
            for (int j = 0; j < E; ++j) s += pv[j] * x[(size_t)pc[j] * c + col];
            if (longrow)
                for (int e = rp[row] + kSlot + slice; e < rp[row + 1]; e += S)
                    s += v[e] * x[(size_t)ci[e] * c + col];
            s = group_sum<S>(s);
            if (slice == 0) x[o] = up ? sdinv[row] * (xo - s) : xo - s;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);   // x stores before the flag store
        if (lane == 0)
            __hip_atomic_store(flag + row, need, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    // three register sets in turn: the loads of tasks t + nw and t + 2 nw fly while task
    // t is processed (compiler fences keep them issued before the current task's work)
    T av[E], bv[E], cv[E];
    int ac[E], bc[E], cc[E], ar, br, cr;
    task_load(wave, ar, av, ac);
    task_load(wave + nw, br, bv, bc);
    for (int t = wave; t < ntask; t += 3 * nw) {
        task_load(t + 2 * nw, cr, cv, cc);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        process(t, ar, av, ac);
        if (t + nw >= ntask) break;
        task_load(t + 3 * nw, ar, av, ac);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        process(t + nw, br, bv, bc);
        if (t + 2 * nw >= ntask) break;
        task_load(t + 4 * nw, br, bv, bc);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        process(t + 2 * nw, cr, cv, cc);
    }
    __syncthreads();
    const bool bad = lds_flag(abort_flag) != 0;
    for (int i = tid; i < n * c; i += nt) xg[i] = bad ? T(__builtin_nan("")) : x[i];
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
    // single-workgroup sweeps (small factors only): level-ordered rows with the long-row
    // bit, fixed-size entry slots
    DevBuf task_row, slot_ci, slot_v64, slot_v32;
};

}  // namespace xt

using namespace xt;

namespace xt {
std::atomic<unsigned long long> g_prec_ids{0};
}
struct xt_prec_s {
    unsigned long long id = ++xt::g_prec_ids;   // never reused, unlike the handle's address
    int p = 0;
    Tri L, U;                  // ILUT factors (level-scheduled solves)
    DevBuf dinv64, dinv32;
    DevBuf work;
    BandPrec *band = nullptr;  // or: exact banded factor (block sweeps)
    // shifted factors (K + s_g M)^-1 of the temporal split, and their gathered columns
    std::vector<double> shifts;
    std::vector<BandPrec *> shifted;
    DevBuf gin[2], gout[2];    // [0] FP64, [1] FP32
    // or: banded factors of runs of independent systems, run g on DOFs goff[g]..goff[g+1]-1
    std::vector<BandPrec *> groups;
    std::vector<int> goff;
    ~xt_prec_s() {
        if (band) band_delete(band);
        for (BandPrec *b : shifted) band_delete(b);
        for (BandPrec *b : groups) band_delete(b);
    }
};

namespace xt {

int prec_rows(xt_prec pr) { return pr ? pr->p : -1; }

static int upload_tri(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                      const std::vector<double> &v, bool forward, Tri &T) {
    T.m.n = n;
    T.m.nnz = rp[n];
    XT_TRY(T.m.rowptr.upload(rp.data(), sizeof(int) * (n + 1)));
    // one zero entry past the end: the single-workgroup sweeps read it, branch-free, in
    // place of a missing entry
    std::vector<int> cp(ci);
    std::vector<double> vp(v);
    cp.push_back(0);
    vp.push_back(0.0);
    XT_TRY(T.m.col.upload(cp.data(), sizeof(int) * cp.size()));
    XT_TRY(T.m.val64.upload(vp.data(), sizeof(double) * vp.size()));
    std::vector<float> v32(vp.begin(), vp.end());
    XT_TRY(T.m.val32.upload(v32.data(), sizeof(float) * v32.size()));
    std::vector<int> lrows;
    build_levels(n, rp, ci, forward, T.lev_ptr, lrows);
    XT_TRY(T.lev_ptr_dev.upload(T.lev_ptr.data(), sizeof(int) * T.lev_ptr.size()));
    T.lev_nnz.assign(T.lev_ptr.size() - 1, 0);
    for (size_t l = 0; l + 1 < T.lev_ptr.size(); ++l)
        for (int k = T.lev_ptr[l]; k < T.lev_ptr[l + 1]; ++k)
            T.lev_nnz[l] += rp[lrows[k] + 1] - rp[lrows[k]];
    XT_TRY(T.lev_rows.upload(lrows.data(), sizeof(int) * n));
    if (n <= kSfMaxRows) {
        std::vector<int> trow(n), sci((size_t)n * kSlot, n);
        std::vector<double> sv((size_t)n * kSlot, 0.0);
        for (int k = 0; k < n; ++k) {
            const int r = lrows[k], len = rp[r + 1] - rp[r];
            trow[k] = r | (len > kSlot ? (1 << 30) : 0);
            for (int q = 0; q < std::min(len, kSlot); ++q) {
                sci[(size_t)k * kSlot + q] = ci[rp[r] + q];
                sv[(size_t)k * kSlot + q] = v[rp[r] + q];
            }
        }
        std::vector<float> sv32(sv.begin(), sv.end());
        XT_TRY(T.task_row.upload(trow.data(), sizeof(int) * n));
        XT_TRY(T.slot_ci.upload(sci.data(), sizeof(int) * sci.size()));
        XT_TRY(T.slot_v64.upload(sv.data(), sizeof(double) * sv.size()));
        XT_TRY(T.slot_v32.upload(sv32.data(), sizeof(float) * sv32.size()));
    }
    XT_HIP(hipDeviceSynchronize());
    return XT_OK;
}

template <typename T>
static int run_tri(const Tri &tri, const T *vals, const T *dinv, int c, const T *b, T *x,
                   bool timed, hipStream_t s) {
    const int nlev = (int)tri.lev_ptr.size() - 1;
    const int *rows = tri.lev_rows.as<int>();
    for (int l = 0; l < nlev; ++l) {
        const int r0 = tri.lev_ptr[l], nr = tri.lev_ptr[l + 1] - r0;
        if (!nr) continue;
        long long n = (long long)nr * c;
        const bool prof = timed && prof_on(XT_PROF_TRSV);
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
// LDS of one workgroup (gfx950: 160 KB); the staged single-workgroup sweeps are used
// when the solution and the level schedule fit.
constexpr size_t kWgLdsMax = 160 * 1024;

// ILUT sweeps Y = (LU)^-1 X Minv_r^T.  stage: p*q*R elements of device memory for the
// temporal mix (used only when minv is given); timed: per-launch HIP-event timing when
// enabled (it creates events, so the allocation-free entry point passes false).
template <typename T>
static int ilut_apply(xt_prec pr, int R, int q, const T *minv, const T *x, T *y, bool f32,
                      T *stage, bool timed, hipStream_t s) {
    const int c = q * R;
    const T *b = x;
    if (minv) {
        XT_TRY(temporal_mix<T, T>(pr->p, q, R, minv, x, stage, s));
        b = stage;
    }
    const T *lv = f32 ? (const T *)pr->L.m.val32.ptr : (const T *)pr->L.m.val64.ptr;
    const T *uv = f32 ? (const T *)pr->U.m.val32.ptr : (const T *)pr->U.m.val64.ptr;
    const T *dv = f32 ? (const T *)pr->dinv32.ptr : (const T *)pr->dinv64.ptr;
    // Small factors: both sweeps in one workgroup (row flags in LDS when the solution
    // fits there, else a barrier per level) instead of a launch per level; large ones:
    // one launch per level over the whole device.
    const long long work = (pr->L.m.nnz + pr->U.m.nnz + 2LL * pr->p) * c;
    if (work <= kOneWgWork) {
        const bool prof = timed && prof_on(XT_PROF_TRSV);
        if (prof) prof_begin(XT_PROF_TRSV, s);
        const size_t xl = sizeof(T) * (size_t)pr->p * c;
        const size_t sl = sf_lds_bytes<T>(pr->p, c);
        const int S = c <= 8 ? 8 : c <= 16 ? 4 : 0;   // entry slices per column
        if (S && sl <= kWgLdsMax && pr->L.slot_ci.ptr && pr->U.slot_ci.ptr) {
            auto tri = [&](const Tri &t, const T *vals) {
                SfTri w;
                w.trow = t.task_row.as<int>();
                w.sci = t.slot_ci.as<int>();
                w.sv = f32 ? t.slot_v32.ptr : t.slot_v64.ptr;
                w.rp = t.m.rowptr.as<int>();
                w.ci = t.m.col.as<int>();
                w.v = vals;
                return w;
            };
#define XT_SF(SS)                                                                           \
    hipLaunchKernelGGL((trsv_sf_kernel<T, SS>), dim3(1), dim3(1024), sl, s, pr->p, c,       \
                       tri(pr->L, lv), tri(pr->U, uv), dv, b, y)
            if (S == 8) XT_SF(8);
            else XT_SF(4);
#undef XT_SF
        } else {
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
        }
        if (prof)
            prof_end(XT_PROF_TRSV, s,
                     (double)(pr->L.m.nnz + pr->U.m.nnz) * (sizeof(T) + 4 + c * sizeof(T)) +
                         (double)pr->p * c * sizeof(T) * 3,
                     2.0 * (pr->L.m.nnz + pr->U.m.nnz) * c);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    XT_TRY(run_tri<T>(pr->L, lv, nullptr, c, b, y, timed, s));
    XT_TRY(run_tri<T>(pr->U, uv, dv, c, nullptr, y, timed, s));
    return XT_OK;
}

// Gathered column of split case t: g[base_t + dof * stride_t] = sum_j m2_t[j] X[dof][r_t q + j]
template <typename T>
__global__ void __launch_bounds__(256)
split_gather_kernel(int p, int q, int c, int ns, const int *__restrict__ cs,
                    const long long *__restrict__ base, const int *__restrict__ stride,
                    const T *__restrict__ m2v, const T *__restrict__ x, T *__restrict__ g) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)p * ns) return;
    const long long dof = idx / ns;
    const int t = (int)(idx - dof * ns);
    const T *m2 = m2v + (size_t)t * 2 * q;
    const T *xr = x + dof * c + (long long)cs[t] * q;
    T acc = 0;
    for (int j = 0; j < q; ++j) acc += m2[j] * xr[j];
    g[base[t] + dof * stride[t]] = acc;
}

// Y[dof][r_t q + j] += v_t[j] h[base_t + dof * stride_t] (each case is one t at most)
template <typename T>
__global__ void __launch_bounds__(256)
split_scatter_kernel(int p, int q, int c, int ns, const int *__restrict__ cs,
                     const long long *__restrict__ base, const int *__restrict__ stride,
                     const T *__restrict__ m2v, const T *__restrict__ h, T *__restrict__ y) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)p * ns) return;
    const long long dof = idx / ns;
    const int t = (int)(idx - dof * ns);
    const T *v = m2v + (size_t)t * 2 * q + q;
    const T hv = h[base[t] + dof * stride[t]];
    T *yr = y + dof * c + (long long)cs[t] * q;
    for (int j = 0; j < q; ++j) yr[j] += v[j] * hv;
}

template <typename T>
int prec_apply_t(xt_prec pr, int R, int q, const T *minv, const T *x, T *y, bool f32,
                 hipStream_t s, const ActiveCols *ac) {
    if (pr->band)
        return band_apply<T>(pr->band, R, q, minv, x, y, s, ac ? ac->Ra[0] : -1,
                             ac ? ac->dcols : nullptr);
    if (!pr->groups.empty()) {   // rows of run g: goff[g] .. goff[g+1]-1, all R q columns
        const long long c = (long long)q * R;
        for (size_t g = 0; g < pr->groups.size(); ++g)
            XT_TRY(band_apply<T>(pr->groups[g], R, q, minv, x + pr->goff[g] * c,
                                 y + pr->goff[g] * c, s, ac ? ac->Ra[g] : -1,
                                 ac ? ac->dcols + g * R : nullptr));
        return XT_OK;
    }
    T *stage = nullptr;
    if (minv) {   // sized for FP64, so switching precision does not reallocate
        XT_TRY(pr->work.reserve(sizeof(double) * (size_t)pr->p * q * R));
        stage = pr->work.as<T>();
    }
    return ilut_apply<T>(pr, R, q, minv, x, y, f32, stage, true, s);
}
unsigned long long prec_id(xt_prec pr) { return pr->id; }
void prec_runs(xt_prec pr, std::vector<long long> &off) {
    off.clear();
    if (pr->band) {
        off = {0, (long long)pr->p};
    } else {
        for (int o : pr->goff) off.push_back(o);
    }
}
bool prec_has_fp64(xt_prec pr) {
    for (const BandPrec *b : pr->groups)
        if (!band_has_fp64(b)) return false;
    return !pr->band || band_has_fp64(pr->band);
}

int prec_shifts(xt_prec pr, const double **shifts, bool *fp64) {
    if (shifts) *shifts = pr->shifts.empty() ? nullptr : pr->shifts.data();
    if (fp64) *fp64 = !pr->shifted.empty() && band_has_fp64(pr->shifted[0]);
    return (int)pr->shifts.size();
}

template <typename T>
int prec_apply_split(xt_prec pr, int R, int q, const T *minv, const PrecSplit *sp, const T *x,
                     T *y, hipStream_t s, const ActiveCols *ac) {
    const bool f32 = sizeof(T) == 4;
    if (!sp || sp->ns == 0 || !pr->band) return prec_apply_t<T>(pr, R, q, minv, x, y, f32, s, ac);
    const int pi = f32 ? 1 : 0, p = pr->p, ns = sp->ns, c = q * R;
    XT_TRY(pr->gin[pi].reserve(sizeof(T) * (size_t)p * ns));
    XT_TRY(pr->gout[pi].reserve(sizeof(T) * (size_t)p * ns));
    T *gi = pr->gin[pi].as<T>(), *go = pr->gout[pi].as<T>();
    const T *m2v = sp->m2v[pi].as<T>();
    const int nblk = ceil_div((long long)p * ns, 256);
    // the gather reads x before the K^-1 sweeps, which may write y over nothing it needs
    hipLaunchKernelGGL(split_gather_kernel<T>, dim3(nblk), dim3(256), 0, s, p, q, c, ns,
                       sp->dcase.as<int>(), sp->dbase.as<long long>(), sp->dstride.as<int>(), m2v,
                       x, gi);
    XT_HIP(hipGetLastError());
    XT_TRY(band_apply<T>(pr->band, R, q, minv, x, y, s, ac ? ac->Ra[0] : -1,
                         ac ? ac->dcols : nullptr));
    for (size_t g = 0; g + 1 < sp->goff.size(); ++g) {
        const int o = sp->goff[g], ng = sp->goff[g + 1] - o;
        XT_ARG(sp->gfac[g] >= 0 && sp->gfac[g] < (int)pr->shifted.size(), "split factor index");
        XT_TRY(band_apply<T>(pr->shifted[sp->gfac[g]], ng, 1, nullptr, gi + (size_t)o * p,
                             go + (size_t)o * p, s));
    }
    hipLaunchKernelGGL(split_scatter_kernel<T>, dim3(nblk), dim3(256), 0, s, p, q, c, ns,
                       sp->dcase.as<int>(), sp->dbase.as<long long>(), sp->dstride.as<int>(), m2v,
                       go, y);
    XT_HIP(hipGetLastError());
    return XT_OK;
}
template int prec_apply_split<double>(xt_prec, int, int, const double *, const PrecSplit *,
                                      const double *, double *, hipStream_t, const ActiveCols *);
template int prec_apply_split<float>(xt_prec, int, int, const float *, const PrecSplit *,
                                     const float *, float *, hipStream_t, const ActiveCols *);

template int prec_apply_t<double>(xt_prec, int, int, const double *, const double *, double *,
                                  bool, hipStream_t, const ActiveCols *);
template int prec_apply_t<float>(xt_prec, int, int, const float *, const float *, float *, bool,
                                 hipStream_t, const ActiveCols *);

// Staging of the allocation-free apply (xt_prec_apply_noalloc): one array per device,
// held until the process ends, shared by every call on that device.  Host code never
// passes the array's name to a kernel or a copy (on the host it is a stand-in, not a
// device address): it resolves the device address once per device with
// hipGetSymbolAddress and uses that.
static __device__ __align__(256) unsigned char scratch[XT_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::atomic<void *> g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (resolved on the first call there).
static int scratch_ptr(void **out) {
    int dev = 0;
    XT_HIP(hipGetDevice(&dev));
    XT_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the scratch table");
    void *ptr = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!ptr) {
        XT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        XT_ARG(ptr != nullptr, "hipGetSymbolAddress returned no address for scratch");
        g_scratch_ptr[dev].store(ptr, std::memory_order_release);
    }
    *out = ptr;
    return XT_OK;
}

}  // namespace xt

extern "C" {

int xt_prec_create(int p, const double *alu, const int *jlu, const int *ju, xt_prec *out) {
    XT_ARG(out && p > 0 && alu && jlu && ju, "prec arguments");
    XT_HIP(hipFuncSetAttribute((const void *)trsv_onewg_kernel<double, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOneWgLds));
    XT_HIP(hipFuncSetAttribute((const void *)trsv_onewg_kernel<float, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOneWgLds));
#define XT_SF_ATTR(T, SS)                                                                   \
    XT_HIP(hipFuncSetAttribute((const void *)trsv_sf_kernel<T, SS>,                          \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kWgLdsMax))
    XT_SF_ATTR(double, 8); XT_SF_ATTR(double, 4);
    XT_SF_ATTR(float, 8); XT_SF_ATTR(float, 4);
#undef XT_SF_ATTR
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

int xt_prec_create_band_shifted(int p, const int *kia, const int *kja, const double *ka,
                                const int *mia, const int *mja, const double *ma, int nshift,
                                const double *shifts, int block, int flags, xt_prec *out) {
    XT_ARG(out && p > 0 && kia && kja && ka && nshift >= 0 && (nshift == 0 || (shifts && mia &&
           mja && ma)), "prec_create_band_shifted arguments");
    XT_ARG(kia[0] == 1 && (nshift == 0 || mia[0] == 1), "CSR must be 1-based");
    for (int g = 0; g < nshift; ++g) XT_ARG(shifts[g] > 0.0, "shifts must be positive");
    xt_prec pr = new (std::nothrow) xt_prec_s;
    if (!pr) return XT_E_NOMEM;
    pr->p = p;
    int rc = band_build(p, kia, kja, ka, block, &pr->band, !(flags & XT_BAND_FP32_ONLY));
    int b = 0;
    if (rc == XT_OK) band_dims(pr->band, &b, nullptr, nullptr);
    // K + s M row by row (both rows sorted by column: merged; K's pattern holds M's for a
    // mesh whose mass couples only DOFs that the stiffness couples, but any union works)
    auto sorted = [p](const int *ia_, const int *ja_) {
        for (int i = 0; i < p; ++i)
            for (int k = ia_[i]; k + 1 < ia_[i + 1]; ++k)
                if (ja_[k] <= ja_[k - 1]) return false;
        return true;
    };
    if (rc == XT_OK && nshift > 0 && (!sorted(kia, kja) || !sorted(mia, mja))) {
        set_error("prec_create_band_shifted: CSR rows must be sorted by column");
        rc = XT_E_ARG;
    }
    std::vector<int> ia(p + 1, 1), ja;
    std::vector<double> a;
    for (int g = 0; g < nshift && rc == XT_OK; ++g) {
        ja.clear();
        a.clear();
        for (int i = 0; i < p; ++i) {
            int k = kia[i] - 1, ke = kia[i + 1] - 1, m = mia[i] - 1, me = mia[i + 1] - 1;
            while (k < ke || m < me) {
                const int jk = k < ke ? kja[k] : 1 << 30, jm = m < me ? mja[m] : 1 << 30;
                if (jk == jm) { ja.push_back(jk); a.push_back(ka[k++] + shifts[g] * ma[m++]); }
                else if (jk < jm) { ja.push_back(jk); a.push_back(ka[k++]); }
                else { ja.push_back(jm); a.push_back(shifts[g] * ma[m++]); }
            }
            ia[i + 1] = (int)ja.size() + 1;
        }
        BandPrec *bp = nullptr;
        if (rc == XT_OK)
            rc = band_build(p, ia.data(), ja.data(), a.data(), b, &bp,
                            !(flags & XT_SHIFT_FP32_ONLY));
        if (rc == XT_OK) {
            pr->shifted.push_back(bp);
            pr->shifts.push_back(shifts[g]);
        }
    }
    if (rc != XT_OK) { delete pr; return rc; }
    *out = pr;
    return XT_OK;
}

// Runs of consecutive systems for xt_prec_create_band_seg: minimise over the partitions
// of the systems into runs  sum_runs (dofs of the run x its block size) + kRunCost,
// the factor entries an apply streams plus a launch cost per run in the same units
// (about 8 us of sweeps at the HBM rate, FP32: 8e-6 s x 5e12 B/s / 16 B per entry),
// with runs of several systems capped in size.
static void band_runs(const std::vector<long long> &n, const std::vector<int> &b,
                      std::vector<int> &cut) {
    constexpr double kRunCost = 2.5e6;
    // a run of several systems holds at most this many factor entries per block set (the
    // FP64 factorisation of a run needs about six such sets at once)
    constexpr double kRunMaxEntries = 1.0e9;
    const int S = (int)n.size();
    std::vector<double> best(S + 1, 0.0);
    std::vector<int> from(S + 1, 0);
    for (int j = 1; j <= S; ++j) {
        best[j] = 1e300;
        long long dofs = 0;
        int bmax = 0;
        for (int i = j - 1; i >= 0; --i) {   // run i..j-1
            dofs += n[i];
            bmax = std::max(bmax, b[i]);
            if (i < j - 1 && (double)dofs * bmax > kRunMaxEntries) break;
            const double c = best[i] + (double)dofs * bmax + kRunCost;
            if (c < best[j]) {
                best[j] = c;
                from[j] = i;
            }
        }
    }
    cut.clear();
    for (int j = S; j > 0; j = from[j]) cut.push_back(j);
    cut.push_back(0);
    std::reverse(cut.begin(), cut.end());
}

int xt_prec_create_band_seg(int p, const int *ia, const int *ja, const double *a, int nseg,
                            const int *seg_dof, int block, int flags, xt_prec *out) {
    XT_ARG(out && p > 0 && ia && ja && a && nseg > 0 && seg_dof && block >= 0,
           "prec_create_band_seg arguments");
    XT_ARG(ia[0] == 1, "CSR must be 1-based");
    XT_ARG(seg_dof[0] == 0 && seg_dof[nseg] == p, "seg_dof must run from 0 to p");
    for (int t = 0; t < nseg; ++t) XT_ARG(seg_dof[t + 1] > seg_dof[t], "empty or unordered segment");
    std::vector<long long> n(nseg);
    std::vector<int> b(nseg);
    for (int t = 0; t < nseg; ++t) {
        const int lo = seg_dof[t], hi = seg_dof[t + 1];
        int bw = 0;
        for (int i = lo; i < hi; ++i) {
            XT_ARG(ia[i + 1] >= ia[i] && ia[i + 1] <= ia[p], "CSR row pointers must not decrease");
            for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
                const int j = ja[k] - 1;
                if (j < lo || j >= hi) {
                    set_error("prec_create_band_seg: row %d (segment %d) couples column %d "
                              "outside its segment", i + 1, t, j + 1);
                    return XT_E_ARG;
                }
                if (j <= i) bw = std::max(bw, i - j);
            }
        }
        n[t] = hi - lo;
        b[t] = (std::max(std::max(block, bw), 64) + 63) / 64 * 64;
    }
    std::vector<int> cut;
    band_runs(n, b, cut);
    xt_prec pr = new (std::nothrow) xt_prec_s;
    if (!pr) return XT_E_NOMEM;
    pr->p = p;
    int rc = XT_OK;
    std::vector<int> sia, sja;
    std::vector<double> sa;
    for (size_t g = 0; g + 1 < cut.size() && rc == XT_OK; ++g) {
        const int lo = seg_dof[cut[g]], hi = seg_dof[cut[g + 1]], pg = hi - lo;
        int bg = 0;
        for (int t = cut[g]; t < cut[g + 1]; ++t) bg = std::max(bg, b[t]);
        // the run's rows, columns shifted to the run (1-based)
        const long long k0 = ia[lo] - 1, k1 = ia[hi] - 1;
        sia.resize(pg + 1);
        for (int i = 0; i <= pg; ++i) sia[i] = (int)(ia[lo + i] - k0);
        sja.resize(k1 - k0);
        for (long long k = k0; k < k1; ++k) sja[k - k0] = ja[k] - lo;
        sa.assign(a + k0, a + k1);
        BandPrec *bp = nullptr;
        rc = band_build(pg, sia.data(), sja.data(), sa.data(), bg, &bp,
                        !(flags & XT_BAND_FP32_ONLY));
        if (rc == XT_OK) {
            pr->groups.push_back(bp);
            pr->goff.push_back(lo);
        }
    }
    if (rc != XT_OK) { delete pr; return rc; }
    pr->goff.push_back(p);
    *out = pr;
    return XT_OK;
}

int xt_prec_band_groups(xt_prec pr, int *ngroup, int *goff, int *gblock) {
    XT_ARG(pr && ngroup, "prec_band_groups arguments");
    if (!pr->groups.empty()) {
        *ngroup = (int)pr->groups.size();
        for (size_t g = 0; g < pr->groups.size(); ++g) {
            if (goff) goff[g] = pr->goff[g];
            if (gblock) band_dims(pr->groups[g], &gblock[g], nullptr, nullptr);
        }
        if (goff) goff[pr->groups.size()] = pr->p;
        return XT_OK;
    }
    *ngroup = pr->band ? 1 : 0;
    if (pr->band) {
        if (goff) { goff[0] = 0; goff[1] = pr->p; }
        if (gblock) band_dims(pr->band, &gblock[0], nullptr, nullptr);
    }
    return XT_OK;
}

int xt_prec_shifts(xt_prec pr, int *nshift, double *shifts) {
    XT_ARG(pr, "prec");
    if (nshift) *nshift = (int)pr->shifts.size();
    if (shifts) std::copy(pr->shifts.begin(), pr->shifts.end(), shifts);
    return XT_OK;
}

int xt_prec_refactor_band(xt_prec pr, const int *ia, const int *ja, const double *a,
                          void *stream) {
    XT_ARG(pr && (pr->band || !pr->groups.empty()) && ia && ja && a,
           "prec_refactor_band arguments");
    if (!pr->groups.empty()) {
        set_error("xt_prec_refactor_band: a grouped factor is rebuilt with "
                  "xt_prec_create_band_seg");
        return XT_E_UNSUPPORTED;
    }
    if (!pr->shifted.empty()) {
        set_error("xt_prec_refactor_band: a preconditioner with shifted factors is rebuilt with "
                  "xt_prec_create_band_shifted");
        return XT_E_UNSUPPORTED;
    }
    return band_refactor(pr->band, ia, ja, a, (hipStream_t)stream);
}

int xt_prec_refresh_band(xt_prec pr, const int *ia, const int *ja, const double *a,
                         int first_row, void *stream) {
    XT_ARG(pr && pr->band && pr->groups.empty() && pr->shifted.empty() && ia && ja && a,
           "prec_refresh_band arguments (a plain banded factor)");
    XT_ARG(first_row >= 0 && first_row < pr->p, "prec_refresh_band: first_row out of range");
    return band_refactor(pr->band, ia, ja, a, (hipStream_t)stream, first_row);
}

int xt_prec_band_info(xt_prec pr, int *block, int *nblocks, int *bandwidth) {
    XT_ARG(pr && (pr->band || !pr->groups.empty()), "not a banded-factor preconditioner");
    if (pr->band) {
        band_dims(pr->band, block, nblocks, bandwidth);
        return XT_OK;
    }
    int bmax = 0, nbs = 0, bwmax = 0;
    for (const BandPrec *g : pr->groups) {
        int b = 0, nb = 0, bw = 0;
        band_dims(g, &b, &nb, &bw);
        bmax = std::max(bmax, b);
        nbs += nb;
        bwmax = std::max(bwmax, bw);
    }
    if (block) *block = bmax;
    if (nblocks) *nblocks = nbs;
    if (bandwidth) *bandwidth = bwmax;
    return XT_OK;
}

int xt_prec_levels(xt_prec pr, int *nlev_lower, int *nlev_upper) {
    XT_ARG(pr, "prec");
    if (pr->band || !pr->groups.empty()) {   // sequential block steps of each sweep
        int nb = 0;
        if (pr->band) band_dims(pr->band, nullptr, &nb, nullptr);
        for (size_t g = 0; g < pr->groups.size(); ++g) {   // the runs follow one another
            int nbg = 0;
            band_dims(pr->groups[g], nullptr, &nbg, nullptr);
            nb += nbg;
        }
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

int xt_prec_apply_noalloc(xt_prec pr, int R, int q, const void *minv, const void *x, void *y,
                          int fp32, void *stream) {
    XT_ARG(pr && R > 0 && q > 0 && x && y && x != y, "prec_apply_noalloc arguments");
    if (pr->band || !pr->groups.empty()) {
        set_error("xt_prec_apply_noalloc: ILUT factors only (the banded factor keeps its own "
                  "workspace)");
        return XT_E_UNSUPPORTED;
    }
    const size_t need = (size_t)pr->p * q * R * (fp32 ? sizeof(float) : sizeof(double));
    if (minv && need > XT_NOALLOC_SCRATCH_BYTES) {
        set_error("xt_prec_apply_noalloc: p*q*R = %lld values need %zu bytes of staging, more "
                  "than the %zu-byte scratch; use xt_prec_apply",
                  (long long)pr->p * q * R, need, (size_t)XT_NOALLOC_SCRATCH_BYTES);
        return XT_E_ARG;
    }
    void *stage = nullptr;
    if (minv) XT_TRY(scratch_ptr(&stage));
    hipStream_t s = (hipStream_t)stream;
    if (fp32)
        return ilut_apply<float>(pr, R, q, (const float *)minv, (const float *)x, (float *)y,
                                 true, (float *)stage, false, s);
    return ilut_apply<double>(pr, R, q, (const double *)minv, (const double *)x, (double *)y,
                              false, (double *)stage, false, s);
}

}  // extern "C"
