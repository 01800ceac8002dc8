// Shared helpers of libxtfem_hip: error state, HIP checks, device buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <atomic>
#include <vector>

#include "../../include/xtfem.h"

namespace xt {

// Thread-local text of the last failure; the C ABI returns codes only.
void set_error(const char *fmt, ...);
const char *get_error();

#define XT_HIP(...)                                                                     \
    do {                                                                                \
        hipError_t e_ = (__VA_ARGS__);                                                  \
        if (e_ != hipSuccess) {                                                         \
            ::xt::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #__VA_ARGS__,         \
                            hipGetErrorString(e_));                                     \
            return XT_E_HIP;                                                            \
        }                                                                               \
    } while (0)

#define XT_TRY(...)                                                                     \
    do {                                                                                \
        int rc_ = (__VA_ARGS__);                                                               \
        if (rc_ != XT_OK) return rc_;                                                   \
    } while (0)

#define XT_ARG(cond, msg)                                                               \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::xt::set_error("argument error: %s", msg);                                 \
            return XT_E_ARG;                                                            \
        }                                                                               \
    } while (0)

// Count of device allocations and frees made through DevBuf: a captured HIP graph holds
// raw device pointers, so one recorded at an older count may name freed memory.
extern std::atomic<unsigned long long> g_dev_epoch;

// Owning device allocation (hipMalloc / hipFree); move-only.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            (void)hipFree(ptr);
            ++g_dev_epoch;
        }
        ptr = nullptr;
        bytes = 0;
    }
    // Grow to at least n bytes (contents not preserved).
    int reserve(size_t n) {
        if (n <= bytes) return XT_OK;
        release();
        if (n == 0) return XT_OK;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
            return XT_E_HIP;
        }
        ++g_dev_epoch;
        bytes = n;
        return XT_OK;
    }
    int upload(const void *src, size_t n, hipStream_t s = 0) {
        XT_TRY(reserve(n));
        if (n) XT_HIP(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s));
        return XT_OK;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Owning page-locked host allocation (hipHostMalloc / hipHostFree): staging for small
// host <-> device transfers, which then need no pageable bounce copy.  The caller keeps
// the contents until the copies that read them have completed.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    void release() { if (ptr) (void)hipHostFree(ptr); ptr = nullptr; bytes = 0; }
    // Grow to at least n bytes (contents not preserved); no copy may be in flight.
    int reserve(size_t n) {
        if (n <= bytes) return XT_OK;
        release();
        if (n == 0) return XT_OK;
        hipError_t e = hipHostMalloc(&ptr, n, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            set_error("hipHostMalloc(%zu) failed: %s", n, hipGetErrorString(e));
            return XT_E_HIP;
        }
        bytes = n;
        return XT_OK;
    }
    template <class T> T *at(size_t off) const {
        return reinterpret_cast<T *>(static_cast<char *>(ptr) + off);
    }
};

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace xt

// Internal entry points shared between translation units.
namespace xt {
// CSR on the device, 0-based.
struct DevCSR {
    int n = 0;
    long long nnz = 0;
    DevBuf rowptr, col, val64, val32;
};
int csr_upload_1based(int n, const int *ia, const int *ja, const double *a, DevCSR &out,
                      bool want32, hipStream_t s);

// Sizes of an operator handle.
int op_dims(xt_op op, int *p, int *q, int *nterms);
// Spatial rows of a preconditioner handle (-1 for none).
int prec_rows(xt_prec pr);
// q x q row-major inverse with partial pivoting; 0 on success, 1 if singular.
int small_inverse_rowmajor(int q, const double *A, double *Ainv);

// Kernel timing (xt_prof.hip): kinds are the XT_PROF_* codes of xtfem.h.
bool prof_on(int kind);
// Replay fixed launch sequences (the band sweeps) from captured HIP graphs.
bool graphs_on();
// Kernel-variant knob (XT_TUNE_*): 0 = the built-in choice.
int tune_knob(int knob);
long long tune_epoch();   // changes whenever a knob does
void prof_begin(int kind, hipStream_t s);
// bytes / flops the launch performs; alg: they are also its algorithmic share (else the
// launch site records the algorithmic totals with prof_alg)
void prof_end(int kind, hipStream_t s, double bytes, double flops, bool alg = true);
void prof_alg(int kind, double bytes, double flops);

// Rank-1 temporal split of a preconditioner that holds shifted factors
// (K + s_g M)^-1 beside K^-1.  For load case r with temporal matrices A_r, B_r let
// lambda_r be the dominant eigenvalue of C_r = A_r^-1 B_r, v_r, u_r its right and left
// eigenvectors (u_r . v_r = 1).  The preconditioner
//   P_r^-1 = [(I - v_r u_r^T) (x) K^-1 + v_r u_r^T (x) (K + s_g M)^-1] (A_r^-1 (x) I)
// applies K^-1 to the temporal mix (I - v u^T) A^-1 (passed as minv) and the shifted
// factor to the single column m2_r = u_r^T A_r^-1 X of the case, added back as v_r times
// the result.  Split cases are numbered t = 0..ns-1 group by group; the gathered column
// of case t sits at base[t] + dof * stride[t] (group g: a contiguous p x n_g block).
struct PrecSplit {
    int ns = 0;
    std::vector<int> goff, gfac;   // host: column offsets of the groups, their factor index
    unsigned long long sig = 0;    // hash of goff and gfac: the launches a split apply makes
    DevBuf dcase, dbase, dstride;  // device: int case[ns], long long base[ns], int stride[ns]
    DevBuf m2v[2];                 // device [ns][2q]: m2 (q) then v (q); FP64 [0], FP32 [1]
    // host sources of the uploads (kept until the next build: the copies are asynchronous)
    std::vector<int> hcase, hstride;
    std::vector<long long> hbase;
    std::vector<double> h64;
    std::vector<float> h32;
};
// Load cases whose columns the banded-factor sweeps must compute (the others' inputs are
// zero: instances frozen in the current GMRES cycle): per run of the factor (one for a
// single factor), Ra[g] active cases listed first in dcols[g R .. g R + R - 1] (device),
// the inactive ones after them.  Ra[g] == R: every case, no compaction.
struct ActiveCols {
    std::vector<int> Ra;
    const int *dcols = nullptr;
};
// DOF offsets of the runs of a banded-factor preconditioner ({0, p} for a single factor;
// empty for ILUT factors).
void prec_runs(xt_prec pr, std::vector<long long> &off);
// Shifted factors of a preconditioner: their number, the shifts when non-NULL, and
// whether they have FP64 copies (else the split serves the FP32 applies only).
int prec_shifts(xt_prec pr, const double **shifts, bool *fp64 = nullptr);
// Y = P^-1 X with the split above (sp may be NULL or empty: plain A^-1 (x) K^-1).
template <typename T>
int prec_apply_split(xt_prec pr, int R, int q, const T *minv, const PrecSplit *sp, const T *x,
                     T *y, hipStream_t s, const ActiveCols *ac = nullptr);
// Y = P^-1 X with the plain A^-1 (x) K^-1 (or ILUT) preconditioner.
template <typename T>
int prec_apply_t(xt_prec pr, int R, int q, const T *minv, const T *x, T *y, bool f32,
                 hipStream_t s, const ActiveCols *ac = nullptr);

// Batched slab solve (xt_slab_solve); hist: residual history of a one-case FP64 solve;
// xtol: forward-error test of the refinement (xt_slab_solve_tol).
int slab_solve_impl(xt_op op, xt_prec pr, int R, const double *ab, const double *b, double *x,
                    int im, double eps, int maxits, int mode, int *its, int *ierr,
                    double *relres, hipStream_t stream, std::vector<double> *hist,
                    double xtol = 0.0, int nseg = 1, const long long *seg_dof = nullptr);
// The preconditioner can apply in FP64 (a banded factor kept in FP32 only cannot).
bool prec_has_fp64(xt_prec pr);
// A number that identifies the preconditioner for the life of the process.
unsigned long long prec_id(xt_prec pr);

// Device workspace of the batched slab solver, owned by the operator handle and kept
// across calls (buffers only grow), so a solve allocates nothing once warm.
struct SolverWorkspace {
    DevBuf tm64, tm32, mi64, mi32;   // temporal matrices and inverses
    DevBuf V, Z;                     // Arnoldi basis, preconditioned basis (flexible)
    DevBuf w32a, w32b, z64, t64;     // work vectors
    DevBuf xprev;                    // iterate before a refinement step (forward-error test)
    // segments (independent meshes in one block-diagonal system): row chunks of the dot
    // products (never straddling a segment), the chunks of each segment, segment of each row
    DevBuf chunks, segchunk, segof;
    std::vector<int> hchunks, hsegchunk, hsegof;
    DevBuf partial, hdev, hdev2, ndev, coef;
    DevBuf reorth;                   // device flag: skip the second Gram-Schmidt pass
    DevBuf actcols;                  // active load cases per factor run (compaction)
    long long act_seq = 0;           // versions of actcols' lists
    // Givens state of a cycle in one device region (Hessenberg columns, rotations, rhs
    // of the least squares problem, last residual norms, counters, frozen flags), the
    // targets and the residual history; pinned host staging mirrors the state region
    DevBuf gstate, geps, ghist;
    PinnedBuf hstage;
    PrecSplit split;                 // temporal split of a shifted-factor preconditioner
    // Arnoldi steps replayed from HIP graphs: one per step index and variant, keyed on
    // every buffer and handle the step's launches use; captured on the second use of a
    // key (the first sizes the buffers), on a stream of its own
    struct StepGraph {
        std::vector<const void *> key;
        int uses = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<StepGraph> steps;
    unsigned long long step_epoch = 0;   // g_dev_epoch the cached step graphs were made at
    // GMRES-IR precision probe for the preconditioner ir_prec (prec_id): phase 0 next
    // solve runs FP32 cycles first, 1 FP64 only, 2 the faster (ir_fp64) for ir_left more
    unsigned long long ir_prec = 0;
    int ir_phase = 0, ir_left = 0;
    double ir_cost_a = 0.0;
    bool ir_fp64 = false;
    long long last_it32 = 0, last_it64 = 0;   // iterations of the last solve by sweep precision
    hipStream_t cap = nullptr;
    SolverWorkspace() = default;
    SolverWorkspace(const SolverWorkspace &) = delete;
    SolverWorkspace &operator=(const SolverWorkspace &) = delete;
    ~SolverWorkspace() {
        for (StepGraph &g : steps)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (cap) (void)hipStreamDestroy(cap);
    }
};
SolverWorkspace *op_workspace(xt_op op);

// Exact block-tridiagonal factor of a banded SPD matrix (xt_band.hip).
struct BandPrec;
int band_build(int p, const int *ia, const int *ja, const double *a, int block, BandPrec **out,
               bool keep64 = true);
bool band_has_fp64(const BandPrec *bp);
void band_delete(BandPrec *bp);
void band_dims(const BandPrec *bp, int *b, int *nb, int *bw);
int band_refactor(BandPrec *bp, const int *dia, const int *dja, const double *da,
                  hipStream_t s, int first_row = 0);
// Ra >= 0 with dcols (device, R entries, the Ra active cases first): the sweeps run on the
// active cases' columns only, gathered after the temporal mix, and y is zero elsewhere.
template <typename T>
int band_apply(BandPrec *bp, int R, int q, const T *minv, const T *x, T *y, hipStream_t s,
               int Ra = -1, const int *dcols = nullptr);

// Temporal mixing Y[dof][r*q+i] = sum_j M_r(i,j) X[dof][r*q+j] (M row-major [R][q][q]).
template <typename TI, typename TO>
int temporal_mix(int p, int q, int R, const TI *M, const TI *x, TO *y, hipStream_t s);
}  // namespace xt
