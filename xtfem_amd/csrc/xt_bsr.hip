// Node-block Kronecker operator for 3-DOF meshes: Y = alpha (K X A_r^T + M X B_r^T) + beta Y.
//
// Same operator as kronspmv/kronspgemv (kronecker.f90:19-127) for the mechanical
// system, whose spatial matrices are made of 3x3 node-pair blocks (a C3D8 element
// couples all 3 displacement components of its 8 nodes; consistent mass couples
// only equal components, elesubs.f90:51-53, so each M block is diagonal).
//
// Storage ("node slices"): node rows, in their natural order, are packed into
// slices of 64 consecutive nodes.  Inside a slice the entries are aligned by
// neighbour offset: the slice keeps the sorted set of offsets d = m - n that
// occur in any of its rows, and entry e of every slot refers to node n + d_e
// (a zero block where that neighbour is absent).  For a structured mesh the set
// is the 27-point stencil, so at every step the 64 lanes of a wave read 64
// consecutive blocks and gather 64 consecutive rows of X: all wave-wide loads
// are contiguous.  A slice whose offset set would be much longer than its
// longest row (irregular meshes) stores its rows by position instead.  Block e
// of slot k of slice s lives at index base[s] + e*64 + k and holds the 3x3 K
// values (row-major) and the 3 diagonal M values (12 numbers), plus one int32
// column-node index.
//
// Work mapping: one lane per (node row, load case r).  The lane accumulates the
// 3 rows x 6 temporal columns of K X and M X for its node in registers (72 FMA
// per block), then applies the 6x6 temporal matrices of its load case in
// registers and stores its 18 outputs once.  A wave covers 64/LPN nodes of one
// slice (LPN = load cases rounded up to a power of two), so loop bounds are
// wave-uniform.  The path is for q = 6 temporal functions (three quadratic
// temporal nodes, each with a sine enrichment: 2*nnte, main3d.f90:172).
#include <algorithm>
#include <vector>

#include "xt_common.h"

namespace xt {

struct BsrPlan {
    int nn = 0;             // nodes (block rows)
    int nslices = 0;
    long long nblocks = 0;  // stored blocks including padding
    long long nreal = 0;    // blocks without padding
    DevBuf slot_node;       // [nslices*64] node of each slot, -1 for empty slots
    DevBuf slice_base;      // [nslices] first block of each slice
    DevBuf slice_len;       // [nslices] blocks per slot
    DevBuf bcol;            // [nblocks] column node
    DevBuf bval32, bval64;  // [nblocks][12]
    DevBuf kmap, mmap;      // value position in bval of every CSR entry of K and of M
    long long nk = 0, nm = 0;
    // Staged form: the distinct column nodes of each slice (its neighbour set, sorted),
    // unode[ubase[s] .. ubase[s] + ulen[s]), and for every stored block the position of
    // its column node in that list (uint16), so a workgroup can stage the slice's rows
    // of X in LDS once and read them from there.
    DevBuf unode, ubase, ulen, lidx;
    int umax = 0;           // longest neighbour set (0: no staged form)
};

static constexpr int kSlice = 64;

// Build the node-slice form; returns XT_E_UNSUPPORTED if K/M are not 3x3-blocked
// with diagonal M blocks.
int bsr_build(int p, const int *krp, const int *kci, const double *kv, const int *mrp,
              const int *mci, const double *mv, BsrPlan &P) {
    if (p % 3) return XT_E_UNSUPPORTED;
    const int nn = p / 3;
    // M must be block-diagonal per node pair (equal components only)
    for (int i = 0; i < p; ++i)
        for (int e = mrp[i]; e < mrp[i + 1]; ++e)
            if (mci[e] % 3 != i % 3) return XT_E_UNSUPPORTED;
    // block pattern of each node row: union of the column nodes of K and M
    std::vector<std::vector<int>> cols(nn);
    std::vector<int> mark(nn, -1);
    for (int a = 0; a < nn; ++a) {
        std::vector<int> &c = cols[a];
        for (int d = 0; d < 3; ++d) {
            const int i = 3 * a + d;
            for (int e = krp[i]; e < krp[i + 1]; ++e) {
                const int b = kci[e] / 3;
                if (mark[b] != a) { mark[b] = a; c.push_back(b); }
            }
            for (int e = mrp[i]; e < mrp[i + 1]; ++e) {
                const int b = mci[e] / 3;
                if (mark[b] != a) { mark[b] = a; c.push_back(b); }
            }
        }
        std::sort(c.begin(), c.end());
        if (c.empty()) c.push_back(a);  // keep a (zero) block so the row is well formed
    }
    const int nslices = ceil_div(nn, kSlice);
    std::vector<int> slot_node((size_t)nslices * kSlice, -1), base(nslices), len(nslices);
    std::vector<std::vector<int>> offs(nslices);  // aligned offsets (empty: by position)
    long long nb = 0;
    for (int s = 0; s < nslices; ++s) {
        int maxlen = 0;
        std::vector<int> u;
        for (int k = 0; k < kSlice; ++k) {
            const int a = s * kSlice + k;
            if (a >= nn) break;
            slot_node[a] = a;
            maxlen = std::max(maxlen, (int)cols[a].size());
            for (int b : cols[a]) u.push_back(b - a);
        }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if ((int)u.size() * 4 <= maxlen * 5) offs[s] = u;
        len[s] = offs[s].empty() ? maxlen : (int)u.size();
        base[s] = (int)nb;
        nb += (long long)len[s] * kSlice;
    }
    if (nb * 12 > 0x7fffffffLL) return XT_E_UNSUPPORTED;
    std::vector<int> bcol(nb);
    std::vector<double> bval((size_t)nb * 12, 0.0);
    std::vector<int> pos(nn, -1);  // position of column node b in the current row
    std::vector<int> kmap(krp[p]), mmap(mrp[p]);
    long long nreal = 0;
    for (int s = 0; s < nslices; ++s)
        for (int k = 0; k < kSlice; ++k) {
            const int a = s * kSlice + k;
            const int L = len[s];
            if (a >= nn) {
                for (int e = 0; e < L; ++e) bcol[(long long)base[s] + (long long)e * kSlice + k] = 0;
                continue;
            }
            const std::vector<int> &c = cols[a];
            nreal += (long long)c.size();
            for (int e = 0; e < L; ++e) {
                const long long idx = (long long)base[s] + (long long)e * kSlice + k;
                if (!offs[s].empty()) {
                    // aligned slot: neighbour a + d_e, clamped into range when absent
                    const int m = a + offs[s][e];
                    bcol[idx] = std::min(std::max(m, 0), nn - 1);
                } else {
                    bcol[idx] = e < (int)c.size() ? c[e] : a;
                }
            }
            // entry index of every real neighbour
            if (!offs[s].empty()) {
                for (int b : c) {
                    const int e = (int)(std::lower_bound(offs[s].begin(), offs[s].end(), b - a) -
                                        offs[s].begin());
                    pos[b] = e;
                }
            } else {
                for (int e = 0; e < (int)c.size(); ++e) pos[c[e]] = e;
            }
            for (int d = 0; d < 3; ++d) {
                const int i = 3 * a + d;
                for (int e = krp[i]; e < krp[i + 1]; ++e) {
                    const int b = kci[e] / 3, cd = kci[e] % 3;
                    const long long idx = (long long)base[s] + (long long)pos[b] * kSlice + k;
                    bval[idx * 12 + d * 3 + cd] = kv[e];
                    kmap[e] = (int)(idx * 12 + d * 3 + cd);
                }
                for (int e = mrp[i]; e < mrp[i + 1]; ++e) {
                    const int b = mci[e] / 3;
                    const long long idx = (long long)base[s] + (long long)pos[b] * kSlice + k;
                    bval[idx * 12 + 9 + d] = mv[e];
                    mmap[e] = (int)(idx * 12 + 9 + d);
                }
            }
            for (int b : c) pos[b] = -1;
        }
    // staged form: per slice the sorted distinct column nodes and each block's position
    {
        std::vector<int> unode, ub(nslices), ul(nslices);
        std::vector<unsigned short> li(nb);
        int umax = 0;
        bool ok = true;
        std::vector<int> tmp;
        for (int sl = 0; sl < nslices && ok; ++sl) {
            const long long b0 = base[sl], b1 = b0 + (long long)len[sl] * kSlice;
            tmp.assign(bcol.begin() + b0, bcol.begin() + b1);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            if (tmp.size() > 65535) { ok = false; break; }
            ub[sl] = (int)unode.size();
            ul[sl] = (int)tmp.size();
            umax = std::max(umax, (int)tmp.size());
            for (long long i = b0; i < b1; ++i)
                li[i] = (unsigned short)(std::lower_bound(tmp.begin(), tmp.end(), bcol[i]) -
                                         tmp.begin());
            unode.insert(unode.end(), tmp.begin(), tmp.end());
        }
        if (ok && !unode.empty()) {
            XT_TRY(P.unode.upload(unode.data(), sizeof(int) * unode.size()));
            XT_TRY(P.ubase.upload(ub.data(), sizeof(int) * nslices));
            XT_TRY(P.ulen.upload(ul.data(), sizeof(int) * nslices));
            XT_TRY(P.lidx.upload(li.data(), sizeof(unsigned short) * nb));
            P.umax = umax;
        }
    }
    P.nn = nn;
    P.nslices = nslices;
    P.nblocks = nb;
    P.nreal = nreal;
    XT_TRY(P.slot_node.upload(slot_node.data(), sizeof(int) * slot_node.size()));
    XT_TRY(P.slice_base.upload(base.data(), sizeof(int) * nslices));
    XT_TRY(P.slice_len.upload(len.data(), sizeof(int) * nslices));
    XT_TRY(P.bcol.upload(bcol.data(), sizeof(int) * nb));
    P.nk = krp[p];
    P.nm = mrp[p];
    if (P.nk) XT_TRY(P.kmap.upload(kmap.data(), sizeof(int) * P.nk));
    if (P.nm) XT_TRY(P.mmap.upload(mmap.data(), sizeof(int) * P.nm));
    XT_TRY(P.bval64.upload(bval.data(), sizeof(double) * bval.size()));
    std::vector<float> b32(bval.begin(), bval.end());
    XT_TRY(P.bval32.upload(b32.data(), sizeof(float) * b32.size()));
    XT_HIP(hipDeviceSynchronize());
    return XT_OK;
}

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// 12 block values (3x3 K + diagonal M) with 16-byte loads.  (Non-temporal loads
// were measured 40% slower here on gfx950 and are not used.)
typedef float xt_f4 __attribute__((ext_vector_type(4)));
typedef double xt_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void load_block(const float *bv, float *kv) {
    const xt_f4 *p = reinterpret_cast<const xt_f4 *>(bv);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const xt_f4 v = p[t];
        kv[4 * t + 0] = v.x;
        kv[4 * t + 1] = v.y;
        kv[4 * t + 2] = v.z;
        kv[4 * t + 3] = v.w;
    }
}
__device__ __forceinline__ void load_block(const double *bv, double *kv) {
    const xt_d2 *p = reinterpret_cast<const xt_d2 *>(bv);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const xt_d2 v = p[t];
        kv[2 * t + 0] = v.x;
        kv[2 * t + 1] = v.y;
    }
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8);
// renumber them so each XCD walks one contiguous eighth of the node range and its
// L2 holds that range's neighbour band of X.  The grid is a multiple of 8, which
// makes the renumbering a permutation of [0, grid).
__device__ __forceinline__ long long xcd_block(long long b, long long grid) {
    const long long per = grid / 8;
    return (b % 8) * per + b / 8;
}

// LPN: lanes per node (power of two >= load cases in this launch, <= 64); UNR: entries
// per pass (loads in flight)
template <typename T, int LPN, int UNR>
__global__ void __launch_bounds__(256)
kron_bsr3_kernel(int nslices, const int *__restrict__ slot_node,
                 const int *__restrict__ slice_base, const int *__restrict__ slice_len,
                 const int *__restrict__ bcol, const T *__restrict__ bval, int R, int r0,
                 int c, const T *__restrict__ tmat, T alpha, T beta, const T *__restrict__ x,
                 T *__restrict__ y) {
    using V2 = typename Vec2<T>::type;
    constexpr int NPW = 64 / LPN;
    const int lane = threadIdx.x & 63;
    const long long blk = xcd_block(blockIdx.x, gridDim.x);
    const long long wave = blk * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long long slot = wave * NPW + lane / LPN;
    const int s = (int)(slot / kSlice);
    if (s >= nslices) return;
    const int k = (int)(slot - (long long)s * kSlice);
    const int r = r0 + (lane % LPN);
    const int node = slot_node[slot];
    const bool act = node >= 0 && r < R;
    const int L = slice_len[s];
    const long long b0 = slice_base[s];
    T aK[3][6], aM[3][6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) { aK[i][j] = T(0); aM[i][j] = T(0); }
    const int rc = act ? r * 6 : 0;
    // UNR entries per pass: their column indices, blocks and X rows are all loaded
    // before the products, so one pass costs two dependent memory latencies (column
    // index, then X) instead of two per entry.
    auto accumulate = [&](const T (&kv)[12], const T (&xv)[3][6]) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                aK[i][j] += kv[3 * i + 0] * xv[0][j] + kv[3 * i + 1] * xv[1][j] + kv[3 * i + 2] * xv[2][j];
                aM[i][j] += kv[9 + i] * xv[i][j];
            }
    };
    auto gather = [&](int m, T (&xv)[3][6]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T *xr = x + (size_t)(3 * m + d) * c + rc;
#pragma unroll
            for (int j = 0; j < 6; j += 2) {
                const V2 v = *reinterpret_cast<const V2 *>(xr + j);
                xv[d][j] = v.x;
                xv[d][j + 1] = v.y;
            }
        }
    };
    int e = 0;
    for (; e + UNR <= L; e += UNR) {
        int m[UNR];
        T kv[UNR][12], xv[UNR][3][6];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const long long idx = b0 + (long long)(e + u) * kSlice + k;
            m[u] = bcol[idx];
            load_block(bval + idx * 12, kv[u]);
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) gather(m[u], xv[u]);
#pragma unroll
        for (int u = 0; u < UNR; ++u) accumulate(kv[u], xv[u]);
    }
    for (; e < L; ++e) {
        const long long idx = b0 + (long long)e * kSlice + k;
        T kv[12], xv[3][6];
        load_block(bval + idx * 12, kv);
        gather(bcol[idx], xv);
        accumulate(kv, xv);
    }
    if (!act) return;
    const T *TA = tmat + (size_t)(r * 2 + 0) * 36;
    const T *TB = tmat + (size_t)(r * 2 + 1) * 36;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T out[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) {
            T sacc = T(0);
#pragma unroll
            for (int j = 0; j < 6; ++j) sacc += TA[o * 6 + j] * aK[i][j] + TB[o * 6 + j] * aM[i][j];
            out[o] = sacc;
        }
        T *yr = y + (size_t)(3 * node + i) * c + rc;
#pragma unroll
        for (int o = 0; o < 6; o += 2) {
            V2 v;
            if (beta == T(0)) {
                v.x = alpha * out[o];
                v.y = alpha * out[o + 1];
            } else {
                const V2 old = *reinterpret_cast<const V2 *>(yr + o);
                v.x = alpha * out[o] + beta * old.x;
                v.y = alpha * out[o + 1] + beta * old.y;
            }
            *reinterpret_cast<V2 *>(yr + o) = v;
        }
    }
}

// Staged variant for many load cases: one workgroup per (slice, chunk of CC columns =
// CC/6 load cases).  The workgroup first copies the rows of X of the slice's neighbour
// nodes, restricted to its columns, into LDS (16-byte loads; each row of X is read once
// per slice instead of once per neighbouring node row), then thread (slot k, case rl)
// walks the slot's entries reading the block values from memory (consecutive slots
// consecutive) and the neighbour's X rows from LDS through the block's position in the
// neighbour list.  Same products and summation order as kron_bsr3_kernel.
template <typename T, int CC>
__global__ void __launch_bounds__(64 * (CC / 6))
kron_bsr3_lds_kernel(int nslices, int nch, const int *__restrict__ slot_node,
                     const int *__restrict__ slice_base, const int *__restrict__ slice_len,
                     const int *__restrict__ unode, const int *__restrict__ ubase,
                     const int *__restrict__ ulen, const unsigned short *__restrict__ lidx,
                     const T *__restrict__ bval, int c, const T *__restrict__ tmat, T alpha,
                     T beta, const T *__restrict__ x, T *__restrict__ y) {
    using V2 = typename Vec2<T>::type;
    constexpr int VEC = 16 / sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *xs = reinterpret_cast<T *>(smem_raw);          // [u][3][CC]
    const long long blk = xcd_block(blockIdx.x, gridDim.x);
    const int s = (int)(blk / nch), ch = (int)(blk - (long long)s * nch);
    if (s >= nslices) return;                           // (grid padded to a multiple of 8)
    const int col0 = ch * CC;
    const int U = ulen[s];
    const int *un = unode + ubase[s];
    constexpr int RPN = 3 * CC / VEC;                   // vector loads per neighbour node
    for (int t = threadIdx.x; t < U * RPN; t += blockDim.x) {
        const int u = t / RPN, v = t - u * RPN;
        const int d = v / (CC / VEC), cv = v - d * (CC / VEC);
        *reinterpret_cast<vec_t *>(xs + ((size_t)u * 3 + d) * CC + cv * VEC) =
            *reinterpret_cast<const vec_t *>(x + ((size_t)un[u] * 3 + d) * c + col0 + cv * VEC);
    }
    __syncthreads();
    const int k = threadIdx.x & 63, rl = threadIdx.x >> 6;
    const int r = ch * (CC / 6) + rl;
    const int node = slot_node[(size_t)s * kSlice + k];
    const int L = slice_len[s];
    const long long b0 = slice_base[s];
    T aK[3][6], aM[3][6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) { aK[i][j] = T(0); aM[i][j] = T(0); }
    constexpr int UNR = 4;
    int e = 0;
    auto step = [&](const T (&kv)[12], int li) {
        T xv[3][6];
        const T *xr = xs + (size_t)li * 3 * CC + rl * 6;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int j = 0; j < 6; j += 2) {
                const V2 v = *reinterpret_cast<const V2 *>(xr + d * CC + j);
                xv[d][j] = v.x;
                xv[d][j + 1] = v.y;
            }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                aK[i][j] += kv[3 * i + 0] * xv[0][j] + kv[3 * i + 1] * xv[1][j] + kv[3 * i + 2] * xv[2][j];
                aM[i][j] += kv[9 + i] * xv[i][j];
            }
    };
    for (; e + UNR <= L; e += UNR) {
        T kv[UNR][12];
        int li[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const long long idx = b0 + (long long)(e + u) * kSlice + k;
            li[u] = lidx[idx];
            load_block(bval + idx * 12, kv[u]);
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) step(kv[u], li[u]);
    }
    for (; e < L; ++e) {
        const long long idx = b0 + (long long)e * kSlice + k;
        T kv[12];
        load_block(bval + idx * 12, kv);
        step(kv, lidx[idx]);
    }
    if (node < 0) return;
    const T *TA = tmat + (size_t)(r * 2 + 0) * 36;
    const T *TB = tmat + (size_t)(r * 2 + 1) * 36;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T out[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) {
            T sacc = T(0);
#pragma unroll
            for (int j = 0; j < 6; ++j) sacc += TA[o * 6 + j] * aK[i][j] + TB[o * 6 + j] * aM[i][j];
            out[o] = sacc;
        }
        T *yr = y + (size_t)(3 * node + i) * c + r * 6;
#pragma unroll
        for (int o = 0; o < 6; o += 2) {
            V2 v;
            if (beta == T(0)) {
                v.x = alpha * out[o];
                v.y = alpha * out[o + 1];
            } else {
                const V2 old = *reinterpret_cast<const V2 *>(yr + o);
                v.x = alpha * out[o] + beta * old.x;
                v.y = alpha * out[o + 1] + beta * old.y;
            }
            *reinterpret_cast<V2 *>(yr + o) = v;
        }
    }
}

// Sum over aligned groups of 4 or 8 lanes with DPP moves (swap neighbours, swap pairs
// in a quad, mirror in the half-row): every lane of the group ends with the group's sum.
template <int CTRL>
__device__ __forceinline__ float quad_move(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double quad_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int SPL, typename T>
__device__ __forceinline__ T group_sum(T v) {
    v += quad_move<0xB1>(v);                     // quad_perm [1,0,3,2]
    v += quad_move<0x4E>(v);                     // quad_perm [2,3,0,1]
    if (SPL == 8) v += quad_move<0x141>(v);      // row_half_mirror
    return v;
}

// One load case (R = 1): SPL lanes per node row, lane `sub` taking the slot's entries
// sub, sub + SPL, ...; the partial K X and M X sums are added across the SPL lanes, and
// lane i < 3 applies the temporal matrices to row i of the node.  SPL times the waves
// of the lane-per-node kernel, and 1/SPL of its dependent load chain per lane, for the
// small launches where that chain is the bound.
template <typename T, int SPL, int UNR>
__global__ void __launch_bounds__(256)
kron_bsr3_r1_kernel(int nslices, const int *__restrict__ slot_node,
                    const int *__restrict__ slice_base, const int *__restrict__ slice_len,
                    const int *__restrict__ bcol, const T *__restrict__ bval,
                    const T *__restrict__ tmat, T alpha, T beta, const T *__restrict__ x,
                    T *__restrict__ y) {
    using V2 = typename Vec2<T>::type;
    constexpr int NPW = 64 / SPL, c = 6;
    const int lane = threadIdx.x & 63, sub = lane % SPL;
    const long long blk = xcd_block(blockIdx.x, gridDim.x);
    const long long wave = blk * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long long slot = wave * NPW + lane / SPL;
    const int s = (int)(slot / kSlice);
    if (s >= nslices) return;   // whole quads leave together
    const int k = (int)(slot - (long long)s * kSlice);
    const int node = slot_node[slot];
    const int L = slice_len[s];
    const long long b0 = slice_base[s];
    T aK[3][6], aM[3][6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) { aK[i][j] = T(0); aM[i][j] = T(0); }
    auto accumulate = [&](const T (&kv)[12], const T (&xv)[3][6]) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                aK[i][j] += kv[3 * i + 0] * xv[0][j] + kv[3 * i + 1] * xv[1][j] + kv[3 * i + 2] * xv[2][j];
                aM[i][j] += kv[9 + i] * xv[i][j];
            }
    };
    auto gather = [&](int m, T (&xv)[3][6]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T *xr = x + (size_t)(3 * m + d) * c;
#pragma unroll
            for (int j = 0; j < 6; j += 2) {
                const V2 v = *reinterpret_cast<const V2 *>(xr + j);
                xv[d][j] = v.x;
                xv[d][j + 1] = v.y;
            }
        }
    };
    int e = sub;
    for (; e + (UNR - 1) * SPL < L; e += UNR * SPL) {
        int m[UNR];
        T kv[UNR][12], xv[UNR][3][6];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const long long idx = b0 + (long long)(e + u * SPL) * kSlice + k;
            m[u] = bcol[idx];
            load_block(bval + idx * 12, kv[u]);
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) gather(m[u], xv[u]);
#pragma unroll
        for (int u = 0; u < UNR; ++u) accumulate(kv[u], xv[u]);
    }
    for (; e < L; e += SPL) {
        const long long idx = b0 + (long long)e * kSlice + k;
        T kv[12], xv[3][6];
        load_block(bval + idx * 12, kv);
        gather(bcol[idx], xv);
        accumulate(kv, xv);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            aK[i][j] = group_sum<SPL>(aK[i][j]);
            aM[i][j] = group_sum<SPL>(aM[i][j]);
        }
    if (node < 0 || sub >= 3) return;
    const T *TA = tmat, *TB = tmat + 36;
    T out[6];
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        T sacc = T(0);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            // row sub of the node: select without dynamic register indexing
            const T kk = sub == 0 ? aK[0][j] : sub == 1 ? aK[1][j] : aK[2][j];
            const T mm = sub == 0 ? aM[0][j] : sub == 1 ? aM[1][j] : aM[2][j];
            sacc += TA[o * 6 + j] * kk + TB[o * 6 + j] * mm;
        }
        out[o] = sacc;
    }
    T *yr = y + (size_t)(3 * node + sub) * c;
#pragma unroll
    for (int o = 0; o < 6; o += 2) {
        V2 v;
        if (beta == T(0)) {
            v.x = alpha * out[o];
            v.y = alpha * out[o + 1];
        } else {
            const V2 old = *reinterpret_cast<const V2 *>(yr + o);
            v.x = alpha * out[o] + beta * old.x;
            v.y = alpha * out[o + 1] + beta * old.y;
        }
        *reinterpret_cast<V2 *>(yr + o) = v;
    }
}

// New K and M values (CSR entries in the order the plan was built from) into the node
// blocks, FP64 and FP32.
__global__ void __launch_bounds__(256)
bsr_scatter_kernel(long long n, const int *__restrict__ map, const double *__restrict__ v,
                   double *__restrict__ b64, float *__restrict__ b32) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = map[i];
    b64[t] = v[i];
    b32[t] = (float)v[i];
}

int bsr_set_values(BsrPlan &P, const double *kv, const double *mv, hipStream_t s) {
    if (P.nk)
        hipLaunchKernelGGL(bsr_scatter_kernel, dim3(ceil_div(P.nk, 256)), dim3(256), 0, s, P.nk,
                           P.kmap.as<int>(), kv, P.bval64.as<double>(), P.bval32.as<float>());
    if (P.nm)
        hipLaunchKernelGGL(bsr_scatter_kernel, dim3(ceil_div(P.nm, 256)), dim3(256), 0, s, P.nm,
                           P.mmap.as<int>(), mv, P.bval64.as<double>(), P.bval32.as<float>());
    XT_HIP(hipGetLastError());
    return XT_OK;
}

static constexpr int kR1Split = 4;   // lanes per node row for one load case (8: slower)

// Kernel variant of a launch: the tuning knobs, else the built-in choices (4 entries in
// flight for one load case in FP32, else 2: with several cases in FP32 two entries took
// 849 us against 909 at 32 cases and 233 against 271 at 8 on a 435,969-DOF operator; one
// wave per workgroup when four-wave workgroups would leave CUs idle; kR1Split lanes per
// node for one load case).
template <typename T>
static int spmm_unr(int R) {
    const int k = tune_knob(XT_TUNE_SPMM_UNR);
    return k ? k : (sizeof(T) == 4 && R == 1 ? 4 : 2);
}

// LDS budget of one staged workgroup (two fit a CU's 160 KB).
constexpr size_t kStageLds = 64 * 1024;

// Columns per staged workgroup: the largest of 48, 24, 12 (whole load cases, 16-byte
// rows) that divides c and whose neighbour rows fit kStageLds; 0 when none does.
template <typename T>
static int stage_cols(const BsrPlan &P, int c) {
    if (!P.umax) return 0;
    for (int cc : {48, 24, 12})
        if (c % cc == 0 && (size_t)P.umax * 3 * cc * sizeof(T) <= kStageLds) return cc;
    return 0;
}

template <typename T>
int bsr_apply(const BsrPlan &P, int R, const T *tmat, double alpha, double beta, const T *x,
              T *y, bool f32, hipStream_t s) {
    const T *bv = f32 ? (const T *)P.bval32.ptr : (const T *)P.bval64.ptr;
    const int c = 6 * R;
    // staged kernel: only when the tuning knob asks for it (2; measured slower than the
    // lane-per-case kernel on the structured meshes here at 32 load cases)
    const int lk = tune_knob(XT_TUNE_SPMM_LDS);
    const int cc = (lk != 2 || R < 2) ? 0 : stage_cols<T>(P, c);
    if (cc) {
        const int nch = c / cc;
        const int grid = (int)(((long long)P.nslices * nch + 7) / 8 * 8);
        const size_t lds = (size_t)P.umax * 3 * cc * sizeof(T);
        if (prof_on(XT_PROF_SPMM)) prof_begin(XT_PROF_SPMM, s);
#define XT_BSR_LDS(CCV)                                                                        \
    hipLaunchKernelGGL((kron_bsr3_lds_kernel<T, CCV>), dim3(grid), dim3(64 * (CCV / 6)), lds, s, \
                       P.nslices, nch, P.slot_node.as<int>(), P.slice_base.as<int>(),          \
                       P.slice_len.as<int>(), P.unode.as<int>(), P.ubase.as<int>(),            \
                       P.ulen.as<int>(), P.lidx.as<unsigned short>(), bv, c, tmat, (T)alpha,   \
                       (T)beta, x, y)
        if (cc == 48) XT_BSR_LDS(48);
        else if (cc == 24) XT_BSR_LDS(24);
        else XT_BSR_LDS(12);
#undef XT_BSR_LDS
        if (prof_on(XT_PROF_SPMM)) {
            // as the lane-per-case kernel's: the stored format read once, X read once,
            // Y written (and read when beta != 0); the neighbour lists (2 B per block
            // in place of the 4-byte column index, 4 B per listed node) are counted
            const double sv = sizeof(T);
            const double vec = (double)P.nn * 3 * c * sv;
            const double bytes = (double)P.nblocks * (12 * sv + 2) +
                                 (double)P.nslices * (kSlice * 4 + 16) +
                                 (double)(P.unode.bytes) + vec * (beta != 0 ? 3 : 2);
            const double flops = 2.0 * 12 * P.nreal * c + 2.0 * 2 * 6 * (3.0 * P.nn) * c;
            prof_end(XT_PROF_SPMM, s, bytes, flops);
        }
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    for (int r0 = 0; r0 < R; r0 += 64) {
        const int Rc = std::min(64, R - r0);
        int lpn = 1;
        while (lpn < Rc) lpn <<= 1;
        const long long slots = (long long)P.nslices * kSlice;
        const long long waves = slots * lpn / 64;
        const int wk = tune_knob(XT_TUNE_SPMM_WPB);
        const int wpb = wk ? wk : (waves >= 4 * 1024 ? 4 : 1);
        const int grid = (int)(((waves + wpb - 1) / wpb + 7) / 8 * 8);
        const int unr = spmm_unr<T>(R);
#define XT_BSR_LAUNCH(N, U)                                                                   \
    hipLaunchKernelGGL((kron_bsr3_kernel<T, N, U>), dim3(grid), dim3(64 * wpb), 0, s,         \
                       P.nslices, P.slot_node.as<int>(), P.slice_base.as<int>(),              \
                       P.slice_len.as<int>(), P.bcol.as<int>(), bv, R, r0, c, tmat, (T)alpha, \
                       (T)beta, x, y)
#define XT_BSR_UNR(N)                                                                         \
    if (unr == 2) XT_BSR_LAUNCH(N, 2);                                                        \
    else if (unr == 8) XT_BSR_LAUNCH(N, 8);                                                   \
    else XT_BSR_LAUNCH(N, 4)
#define XT_BSR_R1(S, U)                                                                       \
    hipLaunchKernelGGL((kron_bsr3_r1_kernel<T, S, U>), dim3(grid1), dim3(64), 0, s, P.nslices, \
                       P.slot_node.as<int>(), P.slice_base.as<int>(), P.slice_len.as<int>(),  \
                       P.bcol.as<int>(), bv, tmat, (T)alpha, (T)beta, x, y)
        if (prof_on(XT_PROF_SPMM)) prof_begin(XT_PROF_SPMM, s);
        if (R == 1) {   // spl lanes per node
            const int sk = tune_knob(XT_TUNE_SPMM_R1SPLIT);
            const int spl = sk ? sk : kR1Split;
            const int grid1 = (int)(((slots * spl / 64) + 7) / 8 * 8);
            if (spl == 8) {
                if (unr == 2) XT_BSR_R1(8, 2);
                else if (unr == 8) XT_BSR_R1(8, 8);
                else XT_BSR_R1(8, 4);
            } else {
                if (unr == 2) XT_BSR_R1(4, 2);
                else if (unr == 8) XT_BSR_R1(4, 8);
                else XT_BSR_R1(4, 4);
            }
        } else switch (lpn) {
            case 1: XT_BSR_UNR(1); break;
            case 2: XT_BSR_UNR(2); break;
            case 4: XT_BSR_UNR(4); break;
            case 8: XT_BSR_UNR(8); break;
            case 16: XT_BSR_UNR(16); break;
            case 32: XT_BSR_UNR(32); break;
            default: XT_BSR_UNR(64); break;
        }
#undef XT_BSR_R1
#undef XT_BSR_UNR
#undef XT_BSR_LAUNCH
        if (prof_on(XT_PROF_SPMM)) {
            // 12 values (3x3 K block + diagonal of the M block) + 1 column index per
            // stored block, slice tables, X read, Y written (and read when beta != 0)
            const double sv = sizeof(T);
            const double vec = (double)P.nn * 3 * 6 * Rc * sv;
            const double bytes = (double)P.nblocks * (12 * sv + 4) +
                                 (double)P.nslices * (kSlice * 4 + 8) + vec * (beta != 0 ? 3 : 2);
            const double flops = 2.0 * 12 * P.nreal * 6 * Rc + 2.0 * 2 * 6 * (3.0 * P.nn) * 6 * Rc;
            prof_end(XT_PROF_SPMM, s, bytes, flops);
        }
        XT_HIP(hipGetLastError());
    }
    return XT_OK;
}
template int bsr_apply<float>(const BsrPlan &, int, const float *, double, double, const float *,
                              float *, bool, hipStream_t);
template int bsr_apply<double>(const BsrPlan &, int, const double *, double, double,
                               const double *, double *, bool, hipStream_t);

}  // namespace xt

namespace xt {
BsrPlan *bsr_new() { return new BsrPlan; }
void bsr_delete(BsrPlan *p) { delete p; }
long long bsr_blocks(const BsrPlan *p, int real) { return real ? p->nreal : p->nblocks; }
}  // namespace xt
