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
// The D x and D^T y products of all blocks run in one batched launch.  A recurrence
// w_v -= M_v w_{v-1} (v = 1..m, m = nb - 1; M_v = G_{v-1} forward, and H in reverse
// block order backward) advances 2^L blocks per dependent step: pairing two steps,
//   w_v -= M_v w_{v-1}  for even v            (one batched launch: the new right-hand sides)
//   w_v -= M'_v w_{v-2} for v = 2, 4, .. in order,   M'_v = -M_v M_{v-1}
//   w_v -= M_v w_{v-1}  for odd v             (one batched launch: the blocks between)
// and the middle line is again such a recurrence over every second block, paired in
// turn for L levels.  The level matrices are formed once per factorization; the
// dependent chain shrinks from m launches to m / 2^L plus 2 L batched launches.
// All matrix sets are stored row-major (the row of the operator a tile needs is
// contiguous) in FP64 and FP32.
//
// The factorization runs on the device once per K, and again at every refactorization
// (xt_prec_refactor_band: per slab when K follows the temperature; from the first
// changed block on after an element deletion): this file's panel kernels for the block
// Cholesky, rocSOLVER trtri for D_i and rocBLAS gemm for G_i, H_i and the level matrices
// (b^3 products near the FP64 matrix-core peak).  None of it is on the per-iteration path.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "xt_common.h"

namespace xt {

// The sweeps of one apply are a fixed sequence of 2 + 2 (nb - 1) launches on fixed
// buffers for a given column count and precision: captured once into a HIP graph and
// replayed, they skip the per-launch host overhead.  Keyed on every pointer the
// launches use, so a refactorization or a grown workspace captures anew.
constexpr int kLevels = 4;     // pairing levels of the block recurrences, at most
constexpr int kMinChain = 6;   // a level is added while its chain keeps this many steps
constexpr int kKey = 6 + 2 * kLevels;
struct SweepGraph {
    int c = 0, pi = -1, tile = 0, smallc = 0, nl = 0;   // columns, precision, knobs,
    long long epoch = -1;                                // levels used, knob epoch
    const void *key[kKey] = {};
    hipGraphExec_t exec = nullptr;
};

struct BandPrec {
    int p = 0, b = 0, nb = 0, bw = 0;
    // [.][0] FP64, [.][1] FP32.  Dl: D_i row-major (lower), Du: D_i^T row-major (upper),
    // G, H: recurrence matrices (nb - 1 of each); Lv[dir][L - 1]: the level-L matrices of
    // the forward (0) and backward (1) recurrence, (nb - 1) >> L of them in chain order.
    DevBuf Dl[2], Du[2], G[2], H[2];
    DevBuf Lv[2][kLevels][2];
    int nlev = 0;
    bool keep64 = true;          // false: the FP64 sets are released after the FP32 copies
    DevBuf w1, w2;
    DevBuf part;                 // partial products of the split-k chain steps
    SweepGraph graph[4];
    int next_graph = 0;
    hipStream_t cap = nullptr;   // capture stream (non-blocking, default priority)
    BandPrec() = default;
    BandPrec(const BandPrec &) = delete;
    BandPrec &operator=(const BandPrec &) = delete;
    ~BandPrec() {
        for (SweepGraph &g : graph)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

// ---------------------------------------------------------------------------
// Block products of the sweeps:
//   out_blk[r][col] = (add ? add_blk[r][col] - s : s),  s = sum_k A_blk[r][k] V_blk[k][col]
// A_blk: b x b row-major; V, add, out: b x c row-major (c columns of one row contiguous).
// tri = 1: A lower triangular (k <= r), tri = 2: upper (k >= r); the stored zeros of
// the other triangle are skipped where a whole tile lies there.  Grid z runs over the
// blocks of a batch (a_stride, v_stride, ... between consecutive blocks).
// ---------------------------------------------------------------------------

// Few columns (c <= CM <= 16, V block in LDS): a workgroup owns RW rows.  It first
// copies the rows k0..k1-1 of V (one contiguous span) into LDS with 16-byte loads;
// its 256 threads then stride over k, loading the RW A rows (coalesced, every load
// of a group of GK k issued before the products, indices clamped and out-of-range
// terms multiplied by zero) and reading V from LDS.  The RW * CM <= 16 partial sums
// of the 256 threads are summed through LDS: 16 threads per output add 16 partials
// each, then a 16-lane butterfly (one value per lane).  Grid x = b / RW.
constexpr size_t kRowsLdsMax = 96 * 1024;
constexpr int kRowsMaxB = 1024;   // block rows from which few columns go to matrix cores
constexpr int kRowsMaxB64 = 512;  // the same in FP64 (768-row blocks, 6-16 columns: the
                                  // matrix-core sweeps took a plate step from 49.7 to 38 ms)
constexpr bool kChainSplitDefault = true;    // chain steps as split-k tile products
constexpr bool kChainSplitFewDefault = false;   // the same with at most 16 columns
// Columns up to which the FP64 sweeps use the factor's pairing levels (the 1856-row panel
// factor: pairing 11.6 vs 14.5 ms per apply at 48 columns, 20.7 vs 18.4 at 72, 31.0 vs
// 27.6 at 192); FP32 sweeps use them at every width (16.1 vs 17.5 ms at 192).
constexpr int kPairMaxC64 = 64;
// Batched passes with at most 16 columns on 16 x 16 tiles instead of 64 x 64 ones (the
// panel factor at one load case: FP64 apply 11.3 -> 8.0 ms, FP32 6.0 -> 4.6 ms).
constexpr bool kSmallcMfmaBatched = true;

template <typename T, int CM, int RW>
__global__ void __launch_bounds__(256)
band_rows_kernel(int b, int c, int tri, const T *__restrict__ A, long long a_stride,
                 const T *__restrict__ V, long long v_stride, const T *add, long long add_stride,
                 T *out, long long out_stride) {
    constexpr int VEC = 16 / sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    static_assert(RW * CM <= 16, "at most 16 outputs per workgroup");
    T *Vsh = reinterpret_cast<T *>(smem_raw);   // [k1 - k0][c]
    __shared__ T red[16][256];                   // [output][thread]
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * RW;
    const long long z = blockIdx.z;
    A += z * a_stride;
    V += z * v_stride;
    out += z * out_stride;
    if (add) add += z * add_stride;
    int k0 = 0, k1 = b;
    if (tri == 1) k1 = r0 + RW;
    if (tri == 2) k0 = r0;
    {
        const int n = (k1 - k0) * c;
        const T *src = V + (size_t)k0 * c;
        if ((((size_t)k0 * c) % VEC) == 0 && (n % VEC) == 0) {
            for (int e = tid; e < n / VEC; e += 256)
                *reinterpret_cast<vec_t *>(Vsh + (size_t)e * VEC) =
                    *reinterpret_cast<const vec_t *>(src + (size_t)e * VEC);
        } else {
            for (int e = tid; e < n; e += 256) Vsh[e] = src[e];
        }
    }
    __syncthreads();
    T acc[RW][CM];
#pragma unroll
    for (int rr = 0; rr < RW; ++rr)
#pragma unroll
        for (int j = 0; j < CM; ++j) acc[rr][j] = T(0);
    const T *a0 = A + (size_t)r0 * b;
    constexpr int GK = RW >= 8 ? 2 : 4;
    for (int kb = k0; kb < k1; kb += GK * 256) {
        T av[GK][RW];
#pragma unroll
        for (int it = 0; it < GK; ++it) {
            const int kk = min(kb + tid + 256 * it, k1 - 1);
#pragma unroll
            for (int rr = 0; rr < RW; ++rr) av[it][rr] = a0[(size_t)rr * b + kk];
        }
#pragma unroll
        for (int it = 0; it < GK; ++it) {
            const int k = kb + tid + 256 * it;
            const int kk = min(k, k1 - 1);
            const T m = (k < k1) ? T(1) : T(0);
            T vv[CM];
#pragma unroll
            for (int j = 0; j < CM; ++j) vv[j] = Vsh[(kk - k0) * c + min(j, c - 1)];
#pragma unroll
            for (int rr = 0; rr < RW; ++rr) {
                const T a = av[it][rr] * m;
#pragma unroll
                for (int j = 0; j < CM; ++j) acc[rr][j] = fma(a, vv[j], acc[rr][j]);
            }
        }
    }
#pragma unroll
    for (int rr = 0; rr < RW; ++rr)
#pragma unroll
        for (int j = 0; j < CM; ++j) red[rr * CM + j][tid] = acc[rr][j];
    __syncthreads();
    {
        const int o = tid >> 4, part = tid & 15;   // output o, partials part + 16 i
        T sum = T(0);
        if (o < RW * CM) {
#pragma unroll
            for (int i = 0; i < 16; ++i) sum += red[o][part + 16 * i];
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 16);
        const int rr = o / CM, j = o - rr * CM;
        if (part == 0 && o < RW * CM && j < c) {
            const size_t idx = (size_t)(r0 + rr) * c + j;
            out[idx] = add ? add[idx] - sum : sum;
        }
    }
}

// Many columns (c > 16) on the matrix cores: a workgroup computes one 16 x 16 output
// tile; its 4 waves split the k range into quarters (whole chunks) and each runs
// v_mfma_*_16x16x4 steps over its quarter; the 4 partial tiles are summed through
// LDS in a fixed order (deterministic, one launch).  Lane l = 16 g + t feeds row t of
// A and column t of V; the k index of a step is assigned per lane group g (group g
// holds KG consecutive k of each chunk of 4 KG, loaded as one vector), which any
// order of the k sum allows, so A rows are read with vector loads.  Operands come
// straight from global memory (L2), UNR chunks of loads in flight per pass.  Output
// layout of the 4 accumulators per lane: f32 row 4 g + i, f64 row g + 4 i, column t.
// Grid x = b / 16, y = ceil(c / 16), z = blocks of the batch.
typedef float xt_f4m __attribute__((ext_vector_type(4)));
typedef double xt_d4m __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    typedef xt_f4m acc_t;
    static constexpr int KG = 4, UNR = 4;
    __device__ static acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int g, int i) { return 4 * g + i; }
};
template <>
struct Mfma<double> {
    typedef xt_d4m acc_t;
    static constexpr int KG = 2, UNR = 8;
    __device__ static acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int g, int i) { return g + 4 * i; }
};

template <typename T>
__global__ void __launch_bounds__(256)
band_mfma_kernel(int b, int c, int tri, const T *__restrict__ A, long long a_stride,
                 const T *__restrict__ V, long long v_stride, const T *add, long long add_stride,
                 T *out, long long out_stride) {
    using M = Mfma<T>;
    constexpr int KG = M::KG, CH = 4 * KG, UNR = M::UNR;
    typedef T vec_t __attribute__((ext_vector_type(KG)));
    __shared__ T red[3][64][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, t = lane & 15;
    const int r0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const long long z = blockIdx.z;
    A += z * a_stride;
    V += z * v_stride;
    out += z * out_stride;
    if (add) add += z * add_stride;
    int k0 = 0, k1 = b;
    if (tri == 1) k1 = r0 + 16;
    if (tri == 2) k0 = r0;
    // this wave's quarter of [k0, k1), in whole chunks
    const int per = ((k1 - k0) / CH + 3) / 4 * CH;
    const int kw0 = k0 + wave * per, kw1 = min(k1, kw0 + per);
    const T *arow = A + (size_t)(r0 + t) * b;
    const int cc = min(n0 + t, c - 1);
    typename M::acc_t acc = {T(0), T(0), T(0), T(0)};
    for (int kc = kw0; kc < kw1; kc += CH * UNR) {
        vec_t ra[UNR];
        T rb[UNR][KG];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int kk = kc + u * CH + KG * g;
            const bool ok = kk < kw1;
            const int kl = ok ? kk : kw0;
            ra[u] = *reinterpret_cast<const vec_t *>(arow + kl);
#pragma unroll
            for (int s2 = 0; s2 < KG; ++s2) rb[u][s2] = V[(size_t)(kl + s2) * c + cc];
            if (!ok) {
#pragma unroll
                for (int s2 = 0; s2 < KG; ++s2) ra[u][s2] = T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int s2 = 0; s2 < KG; ++s2) acc = M::step(ra[u][s2], rb[u][s2], acc);
    }
    if (wave > 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) red[wave - 1][lane][i] = acc[i];
    }
    __syncthreads();
    if (wave == 0 && n0 + t < c) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const T sum = ((acc[i] + red[0][lane][i]) + red[1][lane][i]) + red[2][lane][i];
            const size_t idx = (size_t)(r0 + M::row(g, i)) * c + n0 + t;
            out[idx] = add ? add[idx] - sum : sum;
        }
    }
}

// Batched products of many blocks (the block-diagonal passes and the pairing levels'
// passes, where parallelism is plentiful and L2 bandwidth is the limit of the tile
// kernel above): a workgroup computes a TM x 64 output tile of one block.  TM = 64: each
// of its 4 waves a 32 x 32 quadrant as 2 x 2 matrix-core tiles; TM = 128 (large
// batches): each wave a 32 x 64 strip as 2 x 4 tiles, fewer operand reads per product.
// k advances in chunks of 16 staged through LDS, double-buffered (the next chunk's
// global loads are in flight while the current chunk is multiplied), so every operand
// element fetched feeds TM or 64 outputs instead of 16.  In MFMA step s of a chunk,
// lane group g takes k = 4 g + s for both operands.  Columns past c read a clamped
// column and are not stored.  One-dimensional grid, a multiple of 8 covering the
// nblk (b / TM) ceil(c / 64) tiles: workgroups are dealt to the 8 XCDs round-robin
// (workgroup w runs on XCD w % 8), so tile t = (w % 8) (grid / 8) + w / 8 gives each XCD
// a contiguous run of whole blocks and one L2 serves all tiles of a block (the A row
// tiles and V column tiles each tile shares with its neighbours); surplus workgroups
// return at once.  ksplit > 0 (one product split over k, tri = 0, A and V strides 0):
// "block" z covers k in [z ksplit, (z+1) ksplit) and writes its partial product to
// out + z out_stride, summed afterwards by chain_reduce_kernel.
template <typename T, int TM>
__global__ void __launch_bounds__(256)
band_tile_kernel(int b, int c, int tri, int nblk, const T *__restrict__ A, long long a_stride,
                 const T *__restrict__ V, long long v_stride, const T *add,
                 long long add_stride, T *out, long long out_stride, int ksplit) {
    using M = Mfma<T>;
    constexpr int KC = 16, AP = KC + 4, VP = 64 + 4;
    constexpr int AV = TM / 64;                    // A vec4 loads per thread per chunk
    constexpr int RT = 2, CT = TM == 64 ? 2 : 4;   // MFMA tiles per wave (rows, columns)
    typedef T vec4_t __attribute__((ext_vector_type(4)));
    __shared__ T As[2][TM][AP];
    __shared__ T Vs[2][KC][VP];
    const int ntx = b / TM;
    const long long per_blk = (long long)ntx * ((c + 63) / 64);
    const long long tile = (long long)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    if (tile >= per_blk * nblk) return;
    const long long z = tile / per_blk;
    const int rem = (int)(tile - z * per_blk);
    const int R0 = (rem % ntx) * TM, N0 = (rem / ntx) * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, t = lane & 15;
    A += z * a_stride;
    V += z * v_stride;
    out += z * out_stride;
    if (add) add += z * add_stride;
    int k0 = 0, k1 = b;
    if (tri == 1) k1 = R0 + TM;
    if (tri == 2) k0 = R0;
    if (ksplit > 0) {   // split-k product: "block" z is the k range [z ksplit, (z+1) ksplit)
        k0 = (int)z * ksplit;
        k1 = min(b, k0 + ksplit);
    }
    const int ar = tid >> 2, ac = (tid & 3) * 4;   // A chunk: rows ar + 64 u, k ac..ac+3
    const int vr = tid >> 4, vc = (tid & 15) * 4;  // V chunk: k row vr, columns vc..vc+3
    const T *ap = A + (size_t)(R0 + ar) * b + ac;  // 4-element aligned: b % 64 == 0
    int vcol[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) vcol[e] = min(N0 + vc + e, c - 1);
    const int wr = TM == 64 ? (wave >> 1) * 32 : wave * 32;
    const int wc = TM == 64 ? (wave & 1) * 32 : 0;
    typename M::acc_t acc[RT][CT];
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int j = 0; j < CT; ++j) acc[i][j] = typename M::acc_t{T(0), T(0), T(0), T(0)};
    vec4_t ga[AV];
#pragma unroll
    for (int u = 0; u < AV; ++u)
        ga[u] = *reinterpret_cast<const vec4_t *>(ap + (size_t)64 * u * b + k0);
    T gv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) gv[e] = V[(size_t)(k0 + vr) * c + vcol[e]];
    int buf = 0;
    for (int kc = k0; kc < k1; kc += KC) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
            for (int u = 0; u < AV; ++u) As[buf][ar + 64 * u][ac + e] = ga[u][e];
            Vs[buf][vr][vc + e] = gv[e];
        }
        __syncthreads();
        if (kc + KC < k1) {
#pragma unroll
            for (int u = 0; u < AV; ++u)
                ga[u] = *reinterpret_cast<const vec4_t *>(ap + (size_t)64 * u * b + kc + KC);
#pragma unroll
            for (int e = 0; e < 4; ++e) gv[e] = V[(size_t)(kc + KC + vr) * c + vcol[e]];
        }
        T a[RT][4], bv[CT][4];
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) {
#pragma unroll
            for (int i = 0; i < RT; ++i) a[i][s2] = As[buf][wr + 16 * i + t][4 * g + s2];
#pragma unroll
            for (int j = 0; j < CT; ++j) bv[j][s2] = Vs[buf][4 * g + s2][wc + 16 * j + t];
        }
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
            for (int i = 0; i < RT; ++i)
#pragma unroll
                for (int j = 0; j < CT; ++j) acc[i][j] = M::step(a[i][s2], bv[j][s2], acc[i][j]);
        buf ^= 1;
    }
#pragma unroll
    for (int j = 0; j < CT; ++j) {
        const int col = N0 + wc + 16 * j + t;
        if (col >= c) continue;
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const size_t idx = (size_t)(R0 + wr + 16 * i + M::row(g, e)) * c + col;
                out[idx] = add ? add[idx] - acc[i][j][e] : acc[i][j][e];
            }
    }
}

// out = add - (part_0 + part_1 + ... ) (or the sum when add is null), the partials of a
// split-k product summed in a fixed order (the result does not depend on scheduling).
template <typename T>
__global__ void __launch_bounds__(256)
chain_reduce_kernel(long long n, int ns, const T *__restrict__ part, const T *add, T *out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    T sum = part[i];
    for (int z = 1; z < ns; ++z) sum += part[(long long)z * n + i];
    out[i] = add ? add[i] - sum : sum;
}

// Split of one chain-step product b x b times b x c over k so that the 64 x 64 tiles of
// all splits give about four workgroups per CU: ns splits of ks rows (a multiple of the
// 16-row k chunk); XT_TUNE_BAND_CHAIN 2 / 3 aim at 512 / 256 workgroups instead (1024
// measured fastest on a 1856-row, 192-column chain, 512 within 2%).
static void chain_split(int b, int c, int *ks, int *ns) {
    const int kn = tune_knob(XT_TUNE_BAND_CHAIN);
    const long long target = kn == 3 ? 256 : kn == 2 ? 512 : 1024;   // workgroups
    const long long tiles = (long long)(b / 64) * ceil_div(c, 64);
    int want = (int)std::min<long long>(b / 64, std::max<long long>(1, ceil_div(target, tiles)));
    int k = (int)ceil_div(ceil_div(b, want), 16) * 16;
    *ks = k;
    *ns = (int)ceil_div(b, k);
}
// The chain steps use the split-k tile products (else one workgroup per 16 x 16 output
// tile with its 4 waves splitting k): XT_TUNE_BAND_CHAIN 2-4 / 1, built-in: split.
static bool chain_split_on(int c) {
    const int k = tune_knob(XT_TUNE_BAND_CHAIN);
    if (c <= 16) {   // few columns (XT_TUNE_BAND_CHAIN_FEW): split-k tiles 2, one 16 x 16
        const int kf = tune_knob(XT_TUNE_BAND_CHAIN_FEW);   // tile per workgroup 1
        return kf == 2 || (kf == 0 && kChainSplitFewDefault);
    }
    return k >= 2 || (k == 0 && kChainSplitDefault);
}

// Algorithmic bytes and flops of one sweep kernel launch: the matrix elements of the
// stored operator (triangular: the lower or upper half and the diagonal), V read and
// out_elems elements written or read besides, per block of the batch.
// Profile kind of the sweeps in precision T.
template <typename T>
constexpr int band_kind() { return sizeof(T) == 8 ? XT_PROF_BAND64 : XT_PROF_BAND; }

template <typename T>
static void band_prof_end(int b, int c, int tri, int nblk, double out_elems, hipStream_t s) {
    const double mat = tri ? (double)b * (b + 1) / 2 : (double)b * b;
    const double bytes = sizeof(T) * nblk * (mat + (double)b * c + out_elems);
    prof_end(band_kind<T>(), s, bytes, 2.0 * nblk * mat * c, false);
}

// Matrix-core launch for many columns.
template <typename T>
static int launch_tile(const BandPrec &bp, int c, int tri, const T *A, long long as,
                       const T *V, long long vs, const T *add, long long ads, T *out,
                       long long os, int nblk, hipStream_t s) {
    const int b = bp.b;
    const bool prof = prof_on(band_kind<T>());
    if (prof) prof_begin(band_kind<T>(), s);
    const int skn = tune_knob(XT_TUNE_BAND_SMALLC);
    if (nblk > 1 && c <= 16 && (skn == 3 || (skn == 0 && kSmallcMfmaBatched))) {
        // few columns, many blocks: 16 x 16 output tiles (k split over the 4 waves), a
        // quarter of the 64-column tiles' padded products
        hipLaunchKernelGGL(band_mfma_kernel<T>, dim3(b / 16, ceil_div(c, 16), nblk), dim3(256),
                           0, s, b, c, tri, A, as, V, vs, add, ads, out, os);
    } else if (nblk > 1) {   // batched pass: 128-row tiles for large FP32 batches, else 64 (the
                      // FP64 128-row tile holds 2 waves a SIMD and measured slower)
        const int tk = tune_knob(XT_TUNE_BAND_TILE);
        const bool can128 = sizeof(T) == 4 && b % 128 == 0;
        const int tm = (can128 && (tk == 128 || (tk == 0 && nblk >= 23))) ? 128 : 64;
        const long long tiles = (long long)(b / tm) * ceil_div(c, 64) * nblk;
        const dim3 grid((unsigned)(ceil_div(tiles, 8) * 8));
        if (tm == 128)
            hipLaunchKernelGGL((band_tile_kernel<T, 128>), grid, dim3(256), 0, s, b, c, tri, nblk,
                               A, as, V, vs, add, ads, out, os, 0);
        else
            hipLaunchKernelGGL((band_tile_kernel<T, 64>), grid, dim3(256), 0, s, b, c, tri, nblk,
                               A, as, V, vs, add, ads, out, os, 0);
    } else if (tri == 0 && chain_split_on(c) && bp.part.ptr &&
               [&] { int ks = 0, ns = 1; chain_split(b, c, &ks, &ns);
                     return sizeof(T) * (size_t)ns * b * c <= bp.part.bytes; }()) {
        // one product split over k: the partials of all splits, then their fixed-order sum
        int ks = 0, ns = 1;
        chain_split(b, c, &ks, &ns);
        const long long tiles = (long long)(b / 64) * ceil_div(c, 64) * ns;
        const long long n = (long long)b * c;
        T *part = bp.part.as<T>();
        hipLaunchKernelGGL((band_tile_kernel<T, 64>), dim3((unsigned)(ceil_div(tiles, 8) * 8)),
                           dim3(256), 0, s, b, c, 0, ns, A, 0LL, V, 0LL, (const T *)nullptr, 0LL,
                           part, n, ks);
        hipLaunchKernelGGL(chain_reduce_kernel<T>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                           s, n, ns, (const T *)part, add, out);
    } else
        hipLaunchKernelGGL(band_mfma_kernel<T>, dim3(b / 16, ceil_div(c, 16), nblk), dim3(256), 0,
                           s, b, c, tri, A, as, V, vs, add, ads, out, os);
    if (prof) band_prof_end<T>(b, c, tri, nblk, (double)b * c * (add ? 2 : 1), s);
    return XT_OK;
}

template <typename T>
static int launch_gemm(const BandPrec &bp, int c, int tri, const T *A, long long as,
                       const T *V, long long vs, const T *add, long long ads, T *out,
                       long long os, int nblk, hipStream_t s) {
    const int b = bp.b;
    const bool prof = prof_on(band_kind<T>());
    const size_t vlds = sizeof(T) * (size_t)b * c;
    // few columns: the VALU rows kernel stages V once per RW rows, which for large blocks
    // re-reads V from L2 b / RW times per block; the matrix-core kernels read it once per
    // tile and win there (the knob overrides)
    const int sk = tune_knob(XT_TUNE_BAND_SMALLC);
#define XT_ROWS(CM, RW)                                                                        \
    do {                                                                                       \
        if (prof) prof_begin(band_kind<T>(), s);                                               \
        hipLaunchKernelGGL((band_rows_kernel<T, CM, RW>), dim3(b / RW, 1, nblk), dim3(256),    \
                           vlds, s, b, c, tri, A, as, V, vs, add, ads, out, os);               \
        if (prof) band_prof_end<T>(b, c, tri, nblk, (double)b * c * (add ? 2 : 1), s);         \
    } while (0)
    if (vlds > kRowsLdsMax) {
        XT_TRY(launch_tile<T>(bp, c, tri, A, as, V, vs, add, ads, out, os, nblk, s));
    } else if (c <= 16 && (sk == 1 || (sk == 0 && b < (sizeof(T) == 8 ? kRowsMaxB64 : kRowsMaxB)))) {
        if (c <= 8) {
            XT_ROWS(8, 2);
        } else {
            XT_ROWS(16, 1);
        }
    } else {
        XT_TRY(launch_tile<T>(bp, c, tri, A, as, V, vs, add, ads, out, os, nblk, s));
    }
#undef XT_ROWS
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

// ---------------------------------------------------------------------------
// Blocked Cholesky of the band, 64-column panels (replaces potrf + trsm + syrk per block)
//
// Block i's column of the band is the tall panel [L_i; E_i] (diagonal block K_ii, then
// the sub-diagonal block K_{i+1,i}), both column-major b x b with leading dimension b.
// Panel k (columns 64k..64k+63) takes two launches:
//   chol_panel_kernel: every workgroup factors the 64 x 64 diagonal tile (identical
//     arithmetic in each workgroup, so identical results) and workgroup t > 0 solves
//     row tile k + t of the tall panel, P = A_t L_kk^-T, alongside.  Results go to the
//     scratch W; the matrix is only read.
//   chol_update_kernel: copies W back into the panel's tiles and applies the rank-64
//     update to every tile to the right: the lower tiles of L_i, all tiles of E_i, and
//     the lower tiles of the next diagonal block (K_{i+1,i+1} -= E_i E_i^T, a piece per
//     panel).
// ---------------------------------------------------------------------------
constexpr int kCT = 64;   // panel width and tile size

__device__ inline const double *tall_tile(const double *L, const double *E, int b, int nt,
                                          int rt, int ct) {
    const double *base = rt < nt ? L + (size_t)rt * kCT : E + (size_t)(rt - nt) * kCT;
    return base + (size_t)ct * kCT * b;
}

// value of v in lane `src` of the wave (every lane gets it)
__device__ inline double lane_bcast(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

typedef double xt_d4c __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
chol_panel_kernel(int b, int nt, int k, const double *__restrict__ L,
                  const double *__restrict__ E, double *__restrict__ W, int *__restrict__ info) {
    // The 64 x 64 diagonal tile (sA) and this workgroup's row tile (sX) in LDS, factored
    // and solved in four 16-column steps: the 16 x 16 diagonal block by one wave in
    // registers, the rows below it (and the row tile's) by forward substitution, one
    // row per thread, then the rank-16 update of the columns to the right on the matrix
    // cores (v_mfma_f64_16x16x4f64, one 16 x 16 block per wave at a time).
    __shared__ double sA[kCT][kCT + 1], sX[kCT][kCT + 1];
    __shared__ double rdiag[16], colb[16];
    __shared__ int sbad;
    const int tid = threadIdx.x, t = blockIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const double *D = tall_tile(L, E, b, nt, k, k);
    for (int e = tid; e < kCT * kCT; e += 256) {
        const int c = e >> 6, r = e & 63;
        sA[r][c] = D[(size_t)c * b + r];
    }
    if (t > 0) {
        const double *At = tall_tile(L, E, b, nt, k + t, k);
        for (int e = tid; e < kCT * kCT; e += 256) {
            const int c = e >> 6, r = e & 63;
            sX[r][c] = At[(size_t)c * b + r];
        }
    }
    if (tid == 0) sbad = 0;
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
        const int c0 = 16 * s;
        if (tid < 16) {   // (1) 16 x 16 Cholesky, lane r holds row r
            const int r = tid;
            double a[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) a[c] = sA[c0 + r][c0 + c];
            int bad = 0;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const double pj = lane_bcast(a[j], j);
                bad = (!(pj > 0.0) && !bad) ? j + 1 : bad;
                // 1 / sqrt(pj): the hardware estimate refined by two Newton steps
                // (y <- y + y (1 - pj y^2) / 2, full FP64 precision), then sqrt(pj) = pj y
                double rd = __builtin_amdgcn_rsq(pj);
#pragma unroll
                for (int it = 0; it < 2; ++it) rd = fma(0.5 * rd, fma(-pj * rd, rd, 1.0), rd);
                const double d = pj * rd;
                const double col = r > j ? a[j] * rd : (r == j ? d : 0.0);
                a[j] = r >= j ? col : a[j];
                if (r == j) rdiag[j] = rd;
                // the column to every lane through LDS (one write, broadcast reads: the
                // lanes are one wave, whose LDS accesses complete in order)
                colb[r] = col;
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int l = j + 1; l < 16; ++l) {
                    const double cl = colb[l];
                    a[l] = l <= r ? a[l] - col * cl : a[l];
                }
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) sA[c0 + r][c0 + c] = a[c];
            if (r == 0 && bad && !sbad) sbad = c0 + bad;
        }
        __syncthreads();
        // (2) rows below the block in the tile, then the row tile's rows: x L_s^T = a
        const int nd = kCT - c0 - 16;
        double *row = nullptr;
        if (tid < nd) row = &sA[c0 + 16 + tid][0];
        else if (t > 0 && tid < nd + kCT) row = &sX[tid - nd][0];
        if (row) {
            double x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = row[c0 + c];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                double v = x[j];
#pragma unroll
                for (int m = 0; m < j; ++m) v = v - x[m] * sA[c0 + j][c0 + m];
                x[j] = v * rdiag[j];
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) row[c0 + c] = x[c];
        }
        __syncthreads();
        // (3) C -= P Q^T over the 16-column step: lower blocks (I, J), s < J <= I, of the
        // tile, and blocks (R, J), J > s, of the row tile; P from the block row, Q from
        // block row J of the tile, both at columns c0..c0+15
        const int nrem = 3 - s;
        const int n1 = nrem * (nrem + 1) / 2, n2 = t > 0 ? 4 * nrem : 0;
        const int g = lane >> 4, q = lane & 15;
        for (int blk = wave; blk < n1 + n2; blk += 4) {
            int I, J;
            double (*Cm)[kCT + 1];
            if (blk < n1) {
                int jj = 0, rem = blk;
                while (rem >= nrem - jj) { rem -= nrem - jj; ++jj; }
                J = s + 1 + jj;
                I = J + rem;
                Cm = sA;
            } else {
                const int e = blk - n1;
                I = e / nrem;
                J = s + 1 + (e - I * nrem);
                Cm = sX;
            }
            xt_d4c acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) {
                const double pa = Cm[16 * I + q][c0 + 4 * k4 + g];
                const double qb = sA[16 * J + q][c0 + 4 * k4 + g];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pa, qb, acc, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double &cv = Cm[16 * I + g + 4 * i][16 * J + q];
                cv = cv - acc[i];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        if (tid == 0 && sbad) info[0] = k * kCT + sbad;
        for (int e = tid; e < kCT * kCT; e += 256) {
            const int c = e >> 6, r = e & 63;
            W[e] = c <= r ? sA[r][c] : 0.0;
        }
        return;
    }
    double *wt = W + (size_t)t * kCT * kCT;
    for (int e = tid; e < kCT * kCT; e += 256) {
        const int c = e >> 6, r = e & 63;
        wt[e] = sX[r][c];
    }
}

// C (64 x 64 at C, leading dimension b) -= P Q^T with P, Q 64 x 64 column-major in W, on
// the matrix cores: wave w takes the 16-row band 16w of C, four 16 x 16 blocks, k = 64.
__device__ inline void tile_update(double *C, int b, const double *P, const double *Q,
                                   double (*sP)[kCT], double (*sQ)[kCT]) {
    const int tid = threadIdx.x;
    for (int e = tid; e < kCT * kCT; e += 256) {
        (&sP[0][0])[e] = P[e];   // sP[m][r] = P(r, m)
        (&sQ[0][0])[e] = Q[e];
    }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6, g = lane >> 4, q = lane & 15;
    xt_d4c acc[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[n] = xt_d4c{0.0, 0.0, 0.0, 0.0};
    for (int k4 = 0; k4 < kCT / 4; ++k4) {
        const int m = 4 * k4 + g;
        const double pa = sP[m][16 * w + q];
#pragma unroll
        for (int n = 0; n < 4; ++n)
            acc[n] = __builtin_amdgcn_mfma_f64_16x16x4f64(pa, sQ[m][16 * n + q], acc[n], 0, 0, 0);
    }
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double &c = C[(size_t)(16 * n + q) * b + 16 * w + g + 4 * i];
            c = c - acc[n][i];
        }
}

// Work items: [0, ncp) copy W tile t to tall-panel tile (k + t, k); then the lower tiles
// (i, j), k < j <= i < nt, of L_i; then (if E) the nt x (nt - 1 - k) tiles right of the
// panel in E_i; then the nt (nt + 1) / 2 lower tiles of Knext.
__global__ void __launch_bounds__(256)
chol_update_kernel(int b, int nt, int k, double *L, double *E, double *Knext,
                   const double *__restrict__ W) {
    __shared__ double sP[kCT][kCT], sQ[kCT][kCT];
    const int ntot = E ? 2 * nt : nt;
    const int ncp = ntot - k, T = nt - 1 - k;
    int id = blockIdx.x;
    if (id < ncp) {
        const int rt = k + id;
        double *dst = (double *)tall_tile(L, E, b, nt, rt, k);
        const double *src = W + (size_t)id * kCT * kCT;
        for (int e = threadIdx.x; e < kCT * kCT; e += 256) {
            const int c = e >> 6, r = e & 63;
            dst[(size_t)c * b + r] = src[e];
        }
        return;
    }
    id -= ncp;
    const int nl = T * (T + 1) / 2;
    if (id < nl) {   // lower tile (i, j) of L_i right of the panel: j = k+1+jj, i = j + ii
        int jj = 0, rem = id;
        while (rem >= T - jj) { rem -= T - jj; ++jj; }
        const int j = k + 1 + jj, i = j + rem;
        tile_update((double *)tall_tile(L, E, b, nt, i, j), b, W + (size_t)(i - k) * kCT * kCT,
                    W + (size_t)(j - k) * kCT * kCT, sP, sQ);
        return;
    }
    id -= nl;
    if (!E) return;
    if (id < nt * T) {   // E_i tile (i, j), j > k
        const int i = id / T, j = k + 1 + (id - (id / T) * T);
        tile_update((double *)tall_tile(L, E, b, nt, nt + i, j), b,
                    W + (size_t)(nt + i - k) * kCT * kCT, W + (size_t)(j - k) * kCT * kCT, sP, sQ);
        return;
    }
    id -= nt * T;
    // Knext lower tile (i, j), j <= i: -= E_i(row tile i, panel) E_i(row tile j, panel)^T
    int j = 0, rem = id;
    while (rem >= nt - j) { rem -= nt - j; ++j; }
    const int i = j + rem;
    tile_update(Knext + (size_t)j * kCT * b + (size_t)i * kCT, b,
                W + (size_t)(nt + i - k) * kCT * kCT, W + (size_t)(nt + j - k) * kCT * kCT, sP, sQ);
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

// Lower triangle of the 1-based CSR K, rows of blocks fb.. on, into the column-major
// b x b diagonal blocks kd (block fb first) and sub-diagonal blocks ko (row block i+1,
// column block i; the first is column block fb - 1 when fb > 0, else 0), one lane per
// row in entry order (a repeated entry: the last one stays).  The rows of the last block
// beyond p get a unit diagonal.  The blocks are zero on entry.
__global__ void __launch_bounds__(256)
band_scatter_kernel(int p, int b, int nb, int fb, const int *__restrict__ ia,
                    const int *__restrict__ ja, const double *__restrict__ a, double *kd,
                    double *ko, int *outside) {
    const int i = fb * b + blockIdx.x * 256 + threadIdx.x;
    const long long bb = (long long)b * b;
    if (i >= nb * b) return;
    const int bi = i / b, li = i - bi * b, ko0 = fb > 0 ? fb - 1 : 0;
    if (i >= p) {
        kd[(bi - fb) * bb + (long long)li * b + li] = 1.0;
        return;
    }
    for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
        const int j = ja[k] - 1;
        if (j > i) continue;
        if (j < 0 || i - j > b) {   // not in the band this factor was sized for
            outside[0] = 1;
            continue;
        }
        const int bj = j / b, lj = j - bj * b;
        // bi == bj + 1 otherwise
        double *dst = bi == bj ? kd + (bi - fb) * bb : ko + (bj - ko0) * bb;
        dst[(long long)lj * b + li] = a[k];
    }
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

static int zero_fill(double *d, long long n, hipStream_t s) {
    if (n <= 0) return XT_OK;
    hipLaunchKernelGGL(fill_zero_kernel<double>, dim3(8192), dim3(256), 0, s, n, d);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

// Level-L matrices t0 <= t < t1 of direction dir: M^L_t = -(M^{L-1}_{2t+1} M^{L-1}_{2t}),
// M^0_t = G_t forward and H_{m-1-t} backward.  Row-major X is column-major X^T, so the
// row-major product is the column-major product of the stored blocks in the order
// M_{2t}, M_{2t+1}.
static int level_products(rocblas_handle h, int b, int m, int dir, int L, int t0, int t1,
                          const double *G, const double *H, double *const lv[2][kLevels]) {
    if (t1 <= t0) return XT_OK;
    const long long bb = (long long)b * b;
    const double minus = -1.0, zero = 0.0;
    double *out = lv[dir][L - 1];
    if (L == 1 && dir == 1) {   // level 0 backward runs through H in reverse
        for (int t = t0; t < t1; ++t) {
            const double *A = H + (long long)(m - 1 - 2 * t) * bb;
            XT_RB(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, b, b, b,
                                &minus, A, b, A - bb, b, &zero, out + (long long)t * bb, b));
        }
        return XT_OK;
    }
    // consecutive pairs of a contiguous set: one batched call
    const double *A = (L == 1 ? G : lv[dir][L - 2]) + (long long)(2 * t0) * bb;
    XT_RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, b, b,
                                        b, &minus, A, b, 2 * bb, A + bb, b, 2 * bb, &zero,
                                        out + (long long)t0 * bb, b, bb, t1 - t0));
    return XT_OK;
}

// G_i = (D_{i+1} E_i)^T column-major, i.e. row-major D_{i+1} E_i, for n consecutive i
// (E_i, D_{i+1} column-major with stride bb); H_i = E_i D_i column-major, row-major
// (E_i D_i)^T.
static int recurrence_products(rocblas_handle h, int b, const double *E, const double *Dnext,
                               double *G, int ng, const double *Dsame, const double *Eh,
                               double *Hout, int nh) {
    const long long bb = (long long)b * b;
    const double one = 1.0, zero = 0.0;
    if (ng > 0)
        XT_RB(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose,
                                            rocblas_operation_transpose, b, b, b, &one, E, b, bb,
                                            Dnext, b, bb, &zero, G, b, bb, ng));
    if (nh > 0)
        XT_RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, b,
                                            b, b, &one, Eh, b, bb, Dsame, b, bb, &zero, Hout, b,
                                            bb, nh));
    return XT_OK;
}

// FP32 copy of elements [off, off + n) of an FP64 set.
static int to_f32_range(DevBuf *set, long long off, long long n, hipStream_t s) {
    if (n <= 0 || !set[0].ptr) return XT_OK;
    hipLaunchKernelGGL(to_f32_kernel, dim3(4096), dim3(256), 0, s, n,
                       set[0].as<double>() + off, set[1].as<float>() + off);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

// Block Cholesky of K given as a 1-based CSR in device memory (lower triangle read).
// fb > 0 (a factor with its FP64 sets, K changed only in rows and columns of blocks fb..):
// blocks before fb are kept and the factor is recomputed from block fb on, starting from
// E_{fb-1} = K_{fb,fb-1} D_{fb-1}^T and the Schur complement K_{fb,fb} - E E^T; the
// pairing levels are recomputed where they depend on a changed block.  The result equals
// a full factorization up to rounding.
static int factor_dev(BandPrec &bp, const int *dia, const int *dja, const double *da,
                      hipStream_t s, int fb = 0) {
    const int b = bp.b, nb = bp.nb, p = bp.p;
    const long long bb = (long long)b * b;
    if (fb > 0 && (fb >= nb || !band_has_fp64(&bp) || !bp.Dl[0].ptr)) fb = 0;
    const int nr = nb - fb;                       // blocks factored here
    DevBuf L, E, Dl, G, H, info, E0;
    XT_TRY(L.reserve(sizeof(double) * bb * nr));
    XT_TRY(zero_fill(L.as<double>(), bb * nr, s));
    const int ne = fb > 0 ? nr : nb - 1;          // sub-diagonal blocks held
    if (ne > 0) {
        XT_TRY(E.reserve(sizeof(double) * bb * ne));
        XT_TRY(zero_fill(E.as<double>(), bb * ne, s));
    }
    XT_TRY(info.reserve(sizeof(int) * (nb + 1)));
    XT_HIP(hipMemsetAsync(info.ptr, 0, sizeof(int) * (nb + 1), s));
    hipLaunchKernelGGL(band_scatter_kernel, dim3(ceil_div((long long)nr * b, 256)), dim3(256), 0,
                       s, p, b, nb, fb, dia, dja, da, L.as<double>(),
                       ne > 0 ? E.as<double>() : nullptr, info.as<int>() + nb);
    XT_HIP(hipGetLastError());
    {
        int outside = 0;
        XT_HIP(hipMemcpyAsync(&outside, info.as<int>() + nb, sizeof(int), hipMemcpyDeviceToHost, s));
        XT_HIP(hipStreamSynchronize(s));
        if (outside) {
            set_error("K has entries outside the band of width %d the factor was built for", b);
            return XT_E_ARG;
        }
    }
    XT_HIP(hipMemsetAsync(info.ptr, 0, sizeof(int) * (nb + 1), s));
    double *Lp = L.as<double>(), *Ep = E.as<double>();
    RbHandle rb;
    XT_RB(rocblas_create_handle(&rb.h));
    XT_RB(rocblas_set_stream(rb.h, s));
    XT_RB(rocblas_set_pointer_mode(rb.h, rocblas_pointer_mode_host));
    // Ep + e(i) bb holds E_i (first K_{i+1,i}): e(i) = i for the full factorization, and
    // i - fb + 1 (E_{fb-1} at 0) for the partial one
    const int eo = fb > 0 ? 1 - fb : 0;
    if (fb > 0) {
        // E_{fb-1} = K_{fb,fb-1} L_{fb-1}^-T = K_{fb,fb-1} D_{fb-1}^T (column-major, D kept
        // column-major in the backward set), in place of the K block; then the Schur
        // complement K_{fb,fb} -= E E^T (lower triangle)
        const double one = 1.0, zero = 0.0, minus = -1.0;
        XT_TRY(E0.reserve(sizeof(double) * bb));
        XT_RB(rocblas_dgemm(rb.h, rocblas_operation_none, rocblas_operation_transpose, b, b, b,
                            &one, Ep, b, bp.Du[0].as<double>() + (long long)(fb - 1) * bb, b,
                            &zero, E0.as<double>(), b));
        XT_HIP(hipMemcpyAsync(Ep, E0.ptr, sizeof(double) * bb, hipMemcpyDeviceToDevice, s));
        XT_RB(rocblas_dsyrk(rb.h, rocblas_fill_lower, rocblas_operation_none, b, b, &minus, Ep,
                            b, &one, Lp, b));
    }
    int *ip = info.as<int>();
    // L_ii, E_i = K_{i+1,i} L_ii^-T and K_{i+1,i+1} -= E_i E_i^T, panel by panel
    const int nt = b / kCT;
    DevBuf Wb;
    XT_TRY(Wb.reserve(sizeof(double) * 2 * nt * kCT * kCT));
    for (int i = fb; i < nb; ++i) {
        double *Li = Lp + (long long)(i - fb) * bb;
        double *Ei = i + 1 < nb ? Ep + (long long)(i + eo) * bb : nullptr;
        double *Kn = i + 1 < nb ? Li + bb : nullptr;
        const int ntot = Ei ? 2 * nt : nt;
        for (int k = 0; k < nt; ++k) {
            const int T = nt - 1 - k;
            hipLaunchKernelGGL(chol_panel_kernel, dim3(ntot - k), dim3(256), 0, s, b, nt, k, Li,
                               Ei, Wb.as<double>(), ip + i);
            const int nwork = (ntot - k) + T * (T + 1) / 2 + (Ei ? nt * T + nt * (nt + 1) / 2 : 0);
            hipLaunchKernelGGL(chol_update_kernel, dim3(nwork), dim3(256), 0, s, b, nt, k, Li, Ei,
                               Kn, Wb.as<const double>());
        }
    }
    XT_HIP(hipGetLastError());
    std::vector<int> hinfo(nb + 1);
    XT_HIP(hipMemcpyAsync(hinfo.data(), ip, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    XT_HIP(hipStreamSynchronize(s));
    for (int i = fb; i < nb; ++i)
        if (hinfo[i] != 0) {
            set_error("K is not positive definite (block %d, minor %d)", i, hinfo[i]);
            return XT_E_SINGULAR;
        }
    hipLaunchKernelGGL(zero_upper_kernel, dim3(ceil_div(bb, 256 * 8), nr), dim3(256), 0, s, b, Lp);
    XT_HIP(hipGetLastError());
    // D_i = L_ii^-1 in place (column-major lower = D_i^T row-major: the backward set)
    XT_RB(rocsolver_dtrtri_strided_batched(rb.h, rocblas_fill_lower, rocblas_diagonal_non_unit,
                                           b, Lp, b, bb, ip + fb, nr));
    XT_HIP(hipMemcpyAsync(hinfo.data(), ip, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    XT_HIP(hipStreamSynchronize(s));
    for (int i = fb; i < nb; ++i)
        if (hinfo[i] != 0) {
            set_error("singular diagonal factor block %d", i);
            return XT_E_SINGULAR;
        }
    const int m = nb - 1;
    if (fb > 0) {
        // nothing can fail numerically from here on: the new blocks go into the kept sets
        double *Du = bp.Du[0].as<double>(), *Dlp = bp.Dl[0].as<double>();
        double *Gp = bp.G[0].as<double>(), *Hp = bp.H[0].as<double>();
        XT_HIP(hipMemcpyAsync(Du + (long long)fb * bb, Lp, sizeof(double) * bb * nr,
                              hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(transpose_kernel, dim3(b / 32, b / 32, nr), dim3(256), 0, s, b,
                           (const double *)Lp, Dlp + (long long)fb * bb);
        XT_HIP(hipGetLastError());
        // G_i for i = fb-1..nb-2 (E_{fb-1} with the new D_fb on), H_i for i = fb..nb-2
        XT_TRY(recurrence_products(rb.h, b, Ep, Lp, Gp + (long long)(fb - 1) * bb, nr, Lp,
                                   Ep + bb, Hp + (long long)fb * bb, nr - 1));
        // levels: forward entries whose G range reaches index fb - 1, backward entries
        // whose H range (H_{m-1-t}) reaches index fb
        double *lvp[2][kLevels] = {};
        for (int dir = 0; dir < 2; ++dir)
            for (int L = 0; L < bp.nlev; ++L) lvp[dir][L] = bp.Lv[dir][L][0].as<double>();
        for (int L = 1; L <= bp.nlev; ++L) {
            const int cnt = m >> L;
            XT_TRY(level_products(rb.h, b, m, 0, L, (fb - 1) >> L, cnt, Gp, Hp, lvp));
            if (m - 1 - fb >= 0)
                XT_TRY(level_products(rb.h, b, m, 1, L, 0,
                                      std::min(cnt, ((m - 1 - fb) >> L) + 1), Gp, Hp, lvp));
        }
        XT_TRY(to_f32_range(bp.Du, (long long)fb * bb, bb * nr, s));
        XT_TRY(to_f32_range(bp.Dl, (long long)fb * bb, bb * nr, s));
        XT_TRY(to_f32_range(bp.G, (long long)(fb - 1) * bb, bb * nr, s));
        XT_TRY(to_f32_range(bp.H, (long long)fb * bb, bb * (nr - 1), s));
        for (int L = 1; L <= bp.nlev; ++L) {
            const int cnt = m >> L, f0 = (fb - 1) >> L;
            XT_TRY(to_f32_range(bp.Lv[0][L - 1], (long long)f0 * bb, bb * (cnt - f0), s));
            if (m - 1 - fb >= 0)
                XT_TRY(to_f32_range(bp.Lv[1][L - 1], 0,
                                    bb * std::min(cnt, ((m - 1 - fb) >> L) + 1), s));
        }
        XT_HIP(hipStreamSynchronize(s));
        return XT_OK;
    }
    XT_TRY(Dl.reserve(sizeof(double) * bb * nb));
    hipLaunchKernelGGL(transpose_kernel, dim3(b / 32, b / 32, nb), dim3(256), 0, s, b, Lp,
                       Dl.as<double>());
    XT_HIP(hipGetLastError());
    if (nb > 1) {
        XT_TRY(G.reserve(sizeof(double) * bb * (nb - 1)));
        XT_TRY(H.reserve(sizeof(double) * bb * (nb - 1)));
        XT_TRY(recurrence_products(rb.h, b, Ep, Lp + bb, G.as<double>(), nb - 1, Lp, Ep,
                                   H.as<double>(), nb - 1));
    }
    // Level matrices: M^L_t = -(M^{L-1}_{2t+1} M^{L-1}_{2t}), t < m >> L, where
    // M^0_t = G_t forward and H_{m-1-t} backward.
    int nlev = 0;
    // XT_TUNE_BAND_LEVELS v caps the pairing levels at v - 1 (set before the factor is built)
    const int maxlev = tune_knob(XT_TUNE_BAND_LEVELS) ? tune_knob(XT_TUNE_BAND_LEVELS) - 1 : kLevels;
    while (nlev < maxlev && (m >> (nlev + 1)) >= kMinChain) ++nlev;
    DevBuf lv[2][kLevels];
    double *lvp[2][kLevels] = {};
    for (int dir = 0; dir < 2; ++dir)
        for (int L = 1; L <= nlev; ++L) {
            XT_TRY(lv[dir][L - 1].reserve(sizeof(double) * bb * (m >> L)));
            lvp[dir][L - 1] = lv[dir][L - 1].as<double>();
        }
    for (int dir = 0; dir < 2; ++dir)
        for (int L = 1; L <= nlev; ++L)
            XT_TRY(level_products(rb.h, b, m, dir, L, 0, m >> L, G.as<double>(), H.as<double>(),
                                  lvp));
    bp.Du[0] = std::move(L);
    bp.Dl[0] = std::move(Dl);
    bp.G[0] = std::move(G);
    bp.H[0] = std::move(H);
    for (int dir = 0; dir < 2; ++dir)
        for (int L = 0; L < kLevels; ++L) bp.Lv[dir][L][0] = std::move(lv[dir][L]);
    bp.nlev = nlev;
    for (int k = 0; k < 4 + 2 * kLevels; ++k) {
        DevBuf *src = k == 0 ? bp.Dl : k == 1 ? bp.Du : k == 2 ? bp.G : k == 3 ? bp.H
                    : bp.Lv[(k - 4) / kLevels][(k - 4) % kLevels];
        if (!src[0].ptr) continue;
        const long long n = (long long)(src[0].bytes / sizeof(double));
        XT_TRY(src[1].reserve(sizeof(float) * n));
        XT_TRY(to_f32_range(src, 0, n, s));
    }
    XT_HIP(hipStreamSynchronize(s));
    if (!bp.keep64)
        for (int k = 0; k < 4 + 2 * kLevels; ++k) {
            DevBuf *src = k == 0 ? bp.Dl : k == 1 ? bp.Du : k == 2 ? bp.G : k == 3 ? bp.H
                        : bp.Lv[(k - 4) / kLevels][(k - 4) % kLevels];
            src[0].release();
        }
    return XT_OK;
}

// Host CSR: uploaded, then assembled into the blocks on the device (a few % of their bytes).
static int factor(BandPrec &bp, const int *ia, const int *ja, const double *a, hipStream_t s) {
    const long long nnz = ia[bp.p] - 1;
    DevBuf dia, dja, da;
    XT_TRY(dia.upload(ia, sizeof(int) * (bp.p + 1), s));
    XT_TRY(dja.upload(ja, sizeof(int) * nnz, s));
    XT_TRY(da.upload(a, sizeof(double) * nnz, s));
    return factor_dev(bp, dia.as<int>(), dja.as<int>(), da.as<double>(), s);
}

// Refactor with new values of a K whose lower-triangle pattern fits the existing band
// (the same mesh, e.g. new material or boundary conditions), CSR in device memory.
// first_row > 0: K is unchanged in rows and columns before first_row (0-based), and the
// blocks before first_row / b are kept (a factor without FP64 sets is refactored whole).
int band_refactor(BandPrec *bp, const int *dia, const int *dja, const double *da,
                  hipStream_t s, int first_row) {
    const int fb = first_row > 0 ? std::min(first_row / bp->b, bp->nb) : 0;
    return factor_dev(*bp, dia, dja, da, s, fb);
}

int band_build(int p, const int *ia, const int *ja, const double *a, int block, BandPrec **out,
               bool keep64) {
    XT_ARG(p > 0 && ia && ja && a && out, "band prec arguments");
#define XT_ROWS_ATTR(T, CM, RW)                                                                \
    XT_HIP(hipFuncSetAttribute((const void *)band_rows_kernel<T, CM, RW>,                      \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRowsLdsMax))
    XT_ROWS_ATTR(double, 8, 2); XT_ROWS_ATTR(double, 16, 1);
    XT_ROWS_ATTR(float, 8, 2); XT_ROWS_ATTR(float, 16, 1);
#undef XT_ROWS_ATTR
    XT_ARG(ia[0] == 1, "CSR must be 1-based");
    int bw = 0;
    for (int i = 0; i < p; ++i) XT_ARG(ia[i + 1] >= ia[i], "CSR row pointers must not decrease");
    for (int i = 0; i < p; ++i)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
            const int j = ja[k] - 1;
            XT_ARG(j >= 0 && j < p, "column index out of range");
            if (j <= i) bw = std::max(bw, i - j);
        }
    int b = std::max(std::max(block, bw), 64);
    b = (b + 63) / 64 * 64;
    if (b > 4096) {
        set_error("bandwidth %d exceeds the supported block size 4096 (renumber the mesh)", bw);
        return XT_E_UNSUPPORTED;
    }
    const int nb = (p + b - 1) / b;
    // only the lower triangle is read (K symmetric); b >= bandwidth puts every lower
    // entry in a diagonal block or the sub-diagonal block below it
    BandPrec *bp = new (std::nothrow) BandPrec;
    if (!bp) return XT_E_NOMEM;
    bp->p = p;
    bp->b = b;
    bp->nb = nb;
    bp->bw = bw;
    bp->keep64 = keep64;
    hipStream_t s = nullptr;
    int rc = factor(*bp, ia, ja, a, s);
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

bool band_has_fp64(const BandPrec *bp) { return bp->keep64; }

// Columns of the active load cases, gathered (out[row][a q + j] = in[row][cols[a] q + j],
// a < Ra) or scattered back (out[row][cols[a] q + j] = in[row][a q + j] for a < Ra, zero
// for the inactive cases a >= Ra), rows < p.
template <typename T>
__global__ void __launch_bounds__(256)
gather_cols_kernel(long long n, int q, int R, int Ra, const int *__restrict__ cols,
                   const T *__restrict__ in, T *__restrict__ out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;   // over p * Ra q
    if (i >= n) return;
    const int ca = Ra * q;
    const long long row = i / ca;
    const int e = (int)(i - row * ca), a = e / q, j = e - a * q;
    out[i] = in[row * (long long)R * q + (long long)cols[a] * q + j];
}
template <typename T>
__global__ void __launch_bounds__(256)
scatter_cols_kernel(long long n, int q, int R, int Ra, const int *__restrict__ cols,
                    const T *__restrict__ in, T *__restrict__ out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;   // over p * R q
    if (i >= n) return;
    const int c = R * q;
    const long long row = i / c;
    const int e = (int)(i - row * c), a = e / q, j = e - a * q;
    out[row * c + (long long)cols[a] * q + j] =
        a < Ra ? in[row * (long long)Ra * q + (long long)a * q + j] : T(0);
}

template <typename T>
int band_apply(BandPrec *bp, int R, int q, const T *minv, const T *x, T *y, hipStream_t s,
               int Ra, const int *dcols) {
    const int b = bp->b, nb = bp->nb;
    const int pi = sizeof(T) == 4 ? 1 : 0;
    if (pi == 0 && !bp->keep64) {
        set_error("band factor kept in FP32 only: no FP64 sweeps");
        return XT_E_UNSUPPORTED;
    }
    const bool compact = dcols && Ra >= 0 && Ra < R;
    if (compact && Ra == 0) {   // no active case: P^-1 0 = 0
        hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(ceil_div((long long)bp->p * q * R, 256)),
                           dim3(256), 0, s, (long long)bp->p * q * R, y);
        XT_HIP(hipGetLastError());
        return XT_OK;
    }
    const int c = q * (compact ? Ra : R), cfull = q * R;
    const long long npad = (long long)nb * b * c, nreal = (long long)bp->p * c;
    // sized for FP64 sweeps when the factor has them, and for all cases, so switching
    // precision or the active set does not reallocate (captured graphs hold these pointers)
    const size_t es = bp->keep64 ? sizeof(double) : sizeof(T);
    XT_TRY(bp->w1.reserve(es * (size_t)nb * b * cfull));
    XT_TRY(bp->w2.reserve(es * (size_t)nb * b * cfull));
    if (chain_split_on(cfull)) {   // the largest ns x c over the column counts up to cfull
        size_t most = 0;
        for (int cc = q; cc <= cfull; cc += q) {
            if (!chain_split_on(cc)) continue;
            int ks = 0, ns = 1;
            chain_split(b, cc, &ks, &ns);
            most = std::max(most, (size_t)ns * cc);
        }
        XT_TRY(bp->part.reserve(es * most * b));
    }
    T *w1 = bp->w1.as<T>(), *w2 = bp->w2.as<T>();
    if (compact) {   // temporal mix of every case into w2, the active columns into w1
        const T *src = x;
        if (minv) {
            XT_TRY(temporal_mix<T, T>(bp->p, q, R, minv, x, w2, s));
            src = w2;
        }
        hipLaunchKernelGGL(gather_cols_kernel<T>, dim3(ceil_div(nreal, 256)), dim3(256), 0, s,
                           nreal, q, R, Ra, dcols, src, w1);
    } else if (minv) {
        XT_TRY(temporal_mix<T, T>(bp->p, q, R, minv, x, w1, s));
    } else {
        XT_HIP(hipMemcpyAsync(w1, x, sizeof(T) * nreal, hipMemcpyDeviceToDevice, s));
    }
    if (npad > nreal)
        hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(ceil_div(npad - nreal, 256)), dim3(256), 0,
                           s, npad - nreal, w1 + nreal);
    const long long bb = (long long)b * b, bc = (long long)b * c;
    const T *Dl = bp->Dl[pi].as<T>(), *Du = bp->Du[pi].as<T>();
    const T *G = bp->G[pi].as<T>(), *H = bp->H[pi].as<T>();
    // pairing levels used by this apply: the factor's, or none when the columns alone
    // give the chain steps enough work (XT_TUNE_BAND_PAIR: 1 never, 2 always; built-in:
    // FP32 always, FP64 up to kPairMaxC64 columns)
    const int pk = tune_knob(XT_TUNE_BAND_PAIR);
    const int m = nb - 1,
              nlev = (pk == 2 || (pk == 0 && (sizeof(T) == 4 || c <= kPairMaxC64))) ? bp->nlev : 0;
    // One recurrence in chain coordinates: block v (0..m) at W + v S, level-L matrix of
    // step v at Mb[L] + (v - 1) MS[L] (level 0: G forward, H in reverse backward).
    auto chain = [&](int dir, T *w, const T *M0, hipStream_t st) -> int {
        if (m < 1) return XT_OK;
        T *W = dir == 0 ? w : w + (long long)m * bc;
        const long long S = dir == 0 ? bc : -bc;
        const T *Mb[kLevels + 1];
        long long MS[kLevels + 1];
        Mb[0] = dir == 0 ? M0 : M0 + (long long)(m - 1) * bb;
        MS[0] = dir == 0 ? bb : -bb;
        for (int L = 1; L <= nlev; ++L) {
            Mb[L] = bp->Lv[dir][L - 1][pi].as<T>();
            MS[L] = bb;
        }
        for (int L = 0; L < nlev; ++L) {   // even steps of level L: the next level's rhs
            const long long d = 1LL << L;
            XT_TRY(launch_gemm<T>(*bp, c, 0, Mb[L] + MS[L], 2 * MS[L], W + d * S, 2 * d * S,
                                  W + 2 * d * S, 2 * d * S, W + 2 * d * S, 2 * d * S,
                                  (m >> L) / 2, st));
        }
        const long long D = 1LL << nlev;   // the dependent chain of the top level
        for (int v = 1; v <= (m >> nlev); ++v)
            XT_TRY(launch_gemm<T>(*bp, c, 0, Mb[nlev] + (v - 1) * MS[nlev], 0, W + (v - 1) * D * S,
                                  0, W + v * D * S, 0, W + v * D * S, 0, 1, st));
        for (int L = nlev - 1; L >= 0; --L) {   // odd steps of level L, from the finished ones
            const long long d = 1LL << L;
            XT_TRY(launch_gemm<T>(*bp, c, 0, Mb[L], 2 * MS[L], W, 2 * d * S, W + d * S,
                                  2 * d * S, W + d * S, 2 * d * S, ((m >> L) + 1) / 2, st));
        }
        return XT_OK;
    };
    auto sweeps = [&](hipStream_t st) -> int {
        // forward: w2 = D w1 (all blocks), then w2_{i+1} -= G_i w2_i
        XT_TRY(launch_gemm<T>(*bp, c, 1, Dl, bb, w1, bc, nullptr, 0, w2, bc, nb, st));
        XT_TRY(chain(0, w2, G, st));
        // backward: w1 = D^T w2 (all blocks), then w1_i -= H_i w1_{i+1}
        XT_TRY(launch_gemm<T>(*bp, c, 2, Du, bb, w2, bc, nullptr, 0, w1, bc, nb, st));
        XT_TRY(chain(1, w1, H, st));
        return XT_OK;
    };
    hipStreamCaptureStatus capst = hipStreamCaptureStatusNone;
    XT_HIP(hipStreamIsCapturing(s, &capst));
    // (inside a caller's capture the sweeps join that graph as plain launches)
    if (graphs_on() && !prof_on(band_kind<T>()) && capst == hipStreamCaptureStatusNone) {
        const void *key[kKey] = {w1, w2, Dl, Du, G, H};
        for (int dir = 0; dir < 2; ++dir)
            for (int L = 0; L < kLevels; ++L) key[6 + dir * kLevels + L] = bp->Lv[dir][L][pi].ptr;
        SweepGraph *g = nullptr;
        for (SweepGraph &cand : bp->graph)
            if (cand.exec && cand.c == c && cand.pi == pi &&
                cand.tile == tune_knob(XT_TUNE_BAND_TILE) &&
                cand.smallc == tune_knob(XT_TUNE_BAND_SMALLC) &&
                cand.epoch == tune_epoch() &&
                cand.nl == nlev &&
                std::equal(key, key + kKey, cand.key)) {
                g = &cand;
                break;
            }
        if (!g) {
            if (!bp->cap) XT_HIP(hipStreamCreateWithFlags(&bp->cap, hipStreamNonBlocking));
            g = &bp->graph[bp->next_graph];
            bp->next_graph = (bp->next_graph + 1) % 4;
            if (g->exec) {
                (void)hipGraphExecDestroy(g->exec);
                g->exec = nullptr;
            }
            XT_HIP(hipStreamBeginCapture(bp->cap, hipStreamCaptureModeThreadLocal));
            const int rc = sweeps(bp->cap);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(bp->cap, &graph);
            if (rc != XT_OK || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                if (rc != XT_OK) return rc;
                XT_HIP(e);
            }
            const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) g->exec = nullptr;
            XT_HIP(ei);
            g->c = c;
            g->pi = pi;
            g->tile = tune_knob(XT_TUNE_BAND_TILE);
            g->smallc = tune_knob(XT_TUNE_BAND_SMALLC);
            g->epoch = tune_epoch();
            g->nl = nlev;
            std::copy(key, key + kKey, g->key);
        }
        XT_HIP(hipGraphLaunch(g->exec, s));
    } else {
        XT_TRY(sweeps(s));
        // algorithmic work of the apply: the unpaired block recurrences, each factor block
        // read once (D_i, D_i^T as triangles, G_i, H_i), each vector block read and written
        // once per product (plus the read of the block it is subtracted from)
        const double tri = (double)b * (b + 1) / 2, full = (double)b * b, vb = (double)b * c;
        prof_alg(band_kind<T>(),
                 sizeof(T) * (2.0 * (nb * tri + (nb - 1) * full) + 2.0 * (2.0 * nb + 3.0 * (nb - 1)) * vb),
                 2.0 * (2.0 * nb * tri * c + 2.0 * (nb - 1) * full * c));
    }
    XT_HIP(hipGetLastError());
    if (compact) {
        const long long nfull = (long long)bp->p * cfull;
        hipLaunchKernelGGL(scatter_cols_kernel<T>, dim3(ceil_div(nfull, 256)), dim3(256), 0, s,
                           nfull, q, R, Ra, dcols, (const T *)w1, y);
        XT_HIP(hipGetLastError());
    } else {
        XT_HIP(hipMemcpyAsync(y, w1, sizeof(T) * nreal, hipMemcpyDeviceToDevice, s));
    }
    return XT_OK;
}
template int band_apply<double>(BandPrec *, int, int, const double *, const double *, double *,
                                hipStream_t, int, const int *);
template int band_apply<float>(BandPrec *, int, int, const float *, const float *, float *,
                               hipStream_t, int, const int *);

}  // namespace xt
