/*
 * xtfem.h -- C ABI of the MI355X space-time solve path (libxtfem_hip.so).
 *
 * Two groups of entry points:
 *
 * 1. Drop-in replacements for the reference subroutines on the path.  Same
 *    argument meaning and order as the Fortran routines (implicit interfaces,
 *    so Fortran passes every scalar by reference; these C entry points take
 *    scalars by value and a thin iso_c_binding module adapts them), 1-based
 *    CSR/MSR arrays, column-major q x q temporal matrices and temporal-major
 *    space-time vectors x[j*p + dof].  Arrays are host memory; the work runs
 *    on the current HIP device.  They return 0 on success and a negative
 *    XT_E* code on a library failure; the reference's own ierr conventions
 *    are reported through the ierr output arguments where the reference has
 *    one.
 *
 * 2. Device-resident handle API (operators, preconditioners, batched slab
 *    solves, damage state) for callers that keep data in HBM across slabs.
 *    Vectors there use the interleaved layout X[dof][r*q + j] (all q*R
 *    columns of one spatial DOF contiguous; r = right-hand side, j = temporal
 *    function) and are device pointers.
 *
 * No function throws; no function frees caller memory.
 */
#ifndef XTFEM_H
#define XTFEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define XT_OK 0
#define XT_E_ARG -1       /* invalid argument (sizes, null pointers)            */
#define XT_E_HIP -2       /* HIP runtime error (allocation, launch, copy)       */
#define XT_E_SINGULAR -3  /* singular q x q temporal matrix (reference: STOP)   */
#define XT_E_NOMEM -4     /* host allocation failure                            */
#define XT_E_UNSUPPORTED -5

/* Last error message of the calling thread (never NULL). */
const char *xt_last_error(void);

/* ---- 1. reference-signature entry points ------------------------------------ */

/* y = (A (x) K + B (x) M) x.  Replaces kronspmv, kronecker.f90:19-85. */
int xt_kronspmv(const double *a, const double *b, int q,
                const int *ki, const int *kj, const double *kv,
                const int *mi, const int *mj, const double *mv,
                int p, const double *x, double *y);

/* y := alpha (A (x) K + B (x) M) x + beta y.  Replaces kronspgemv,
 * kronecker.f90:88-127 (and mpikronspgemv, kronecker_mumps.f90:76-118, on one
 * device). */
int xt_kronspgemv(const double *a, const double *b, double alpha, double beta, int q,
                  const int *ki, const int *kj, const double *kv,
                  const int *mi, const int *mj, const double *mv,
                  int p, const double *x, double *y);

/* y = A x, CSR.  Replaces amux, ilut.f:3-44. */
int xt_amux(int n, const double *x, double *y, const double *a, const int *ja,
            const int *ia);

/* Dual-threshold ILU of a CSR matrix into MSR alu/jlu/ju.  Replaces ilut,
 * ilut.f:152-495 (host C++; same algorithm, same output format and ierr).
 * Work arrays w (n+1) and jw (2n) as in the reference. */
int xt_ilut(int n, const double *a, const int *ja, const int *ia, int lfil,
            double droptol, double *alu, int *jlu, int *ju, int iwk, double *w,
            int *jw, int *ierr);

/* (LU) x = y with ilut's MSR factors.  Replaces lusol, ilut.f:2383-2432. */
int xt_lusol(int n, const double *y, double *x, const double *alu, const int *jlu,
             const int *ju);

/* Solve (A (x) LU) y = x.  Replaces kronlusol, kronecker.f90:130-178.
 * x and y may alias. */
int xt_kronlusol(const double *a, int q, const double *alu, const int *jlu,
                 const int *ju, int p, const double *x, double *y);

/* Right-preconditioned restarted GMRES on (A (x) K + B (x) M) x = rhs with the
 * Kronecker ILUT preconditioner A (x) LU.  Replaces kronpgmres,
 * kronecker.f90:181-365: rhs is overwritten, sol is the initial guess on entry
 * and the solution on exit, eps is relative to the initial residual, ierr is
 * 0 (converged), 1 (maxits), -1 (zero initial residual).  The reference writes
 * its residual history to Fortran unit iout; here it is returned in hist
 * (its/ro0 ratios, up to hist_cap entries, count in *nhist) and the Fortran
 * binding writes it in the reference's format. */
int xt_kronpgmres(const double *a, const double *b, int q,
                  const int *ki, const int *kj, const double *kv,
                  const int *mi, const int *mj, const double *mv, int p,
                  const double *alu, const int *jlu, const int *ju,
                  double *rhs, double *sol, int im, double eps, int maxits,
                  double *hist, int hist_cap, int *nhist, int *its, int *ierr);

/* Two-scale fatigue damage over nip temporal points per Gauss point.
 * Replaces kokkos_host_, kokkos_host.cpp:37-190 (functor kokkos_kernel.h:44-344):
 * same arguments, same flattened layouts; state arrays are updated in place. */
int xt_kokkos_host(int rank, int ne, int nn, int ndofn, int nnse, int nnte, int ngpe,
                   int ncmp, int nip, const int *nodes, const double *Nx,
                   const double *Ny, const double *Nz, const double *Nt,
                   const double *sol, const double *ym, const double *nu,
                   double *eps_old, double *eps_pl, double *sig_eff, double *sig_back,
                   double *p, double *D, double *ws, double *del_po, double *sig_d);

/* The reference's own C symbols of the damage boundary (kokkos_host.cpp:14-24,
 * 37-41), with its argument list (scalars by value, as kokkos_damage.f90:86-89 passes
 * them with %VAL): kokkos_host_ runs xt_kokkos_host and aborts with a message on a
 * library failure (the reference has no error return); kokkos_init_ does nothing,
 * kokkos_finish_ synchronises the device. */
void kokkos_init_(void);
void kokkos_finish_(void);
void kokkos_host_(int rank, int ne, int nn, int ndofn, int nnse, int nnte, int ngpe, int ncmp,
                  int nip, int *nodes, double *Nx, double *Ny, double *Nz, double *Nt,
                  double *sol, double *ym, double *nu, double *eps_old, double *eps_pl,
                  double *sig_eff, double *sig_back, double *p, double *D, double *ws,
                  double *del_po, double *sig_d);

/* Element status from damage: element e fails iff any of its ngpe Gauss points
 * has D >= Dc.  Replaces chkelesta, delelement.f90:109-136 (device reduction).
 * nefail receives the failed (1-based) element ids in increasing order. */
int xt_chkelesta(int ne, int ngpe, double Dc, const double *D, int *elesta,
                 int *gpsta, int *nef, int *nefail);

/* K := K - sum of the element stiffness matrices of the nef failed elements
 * nefail (1-based ids), each from elekmtx (elesubs.f90:207-251) with the one
 * 6x6 material matrix stiff for all its Gauss points.  Replaces delelement,
 * delelement.f90:1-59 (same arguments; xconn ne x nnse and xcoord nn x 3
 * column-major, dns/dnxyz/sd/wqs as setshapefunc_m and setsdstiff give them).
 * K is 1-based CSR with knz entries, updated in place with the reference's
 * cs_add semantics (sorted rows, exact zeros dropped).  kj and kv must have
 * room for the result, at most knz + nef*(3*nnse)^2 entries (couplings that
 * boundary conditions removed from K come back with -K_e); a K assembled from
 * these elements never needs more than the assembly's own entry count, which
 * the reference's driver allocates (main3d.f90:757-760).  Hexahedra with ndofn = 3,
 * ncmp = 6, ngpe = nquads^3 only (XT_E_UNSUPPORTED otherwise); a singular
 * element Jacobian gives XT_E_SINGULAR where the reference STOPs. */
int xt_delelement(int nn, int ne, int ndofn, int nnse, int ngpe, int ncmp, int nquads,
                  const int *xconn, const double *xcoord, const double *dns,
                  const double *dnxyz, const double *sd, const double *stiff,
                  const double *wqs, int nef, const int *nefail, int *ki, int *kj,
                  double *kv, int ndof, int *knz);

/* necnt[node] -= 1 for every node of every failed element.  Replaces chkcnsta,
 * delelement.f90:139-161. */
int xt_chkcnsta(int nn, int ne, int nnse, const int *xconn, int nef, const int *nefail,
                int *necnt);

/* Nodes of the failed elements whose surviving-element count is 0, each once,
 * in the order the element list first reaches them; ndfail has nn entries
 * (zero-filled past ndf).  Replaces chkndsta, delelement.f90:164-201. */
int xt_chkndsta(int nn, int ne, int nnse, const int *xconn, int nef, const int *nefail,
                const int *necnt, int *ndf, int *ndfail);

/* K := K + I on the ndofn DOFs of each orphan node (cs_add semantics).  kj and
 * kv must hold knz + ndf*ndofn entries.  Replaces adddiagonal,
 * delelement.f90:204-247. */
int xt_adddiagonal(int nnse, int ndofn, int ndf, const int *ndfail, int *ki, int *kj,
                   double *kv, int ndof, int *knz);

/* 1-0 method on the DOFs with bc != 0: their rows and columns are cleared, the
 * diagonal set to 1, exact zeros dropped, in place.  Replaces csrapplybc,
 * spmtx.f90:268-307. */
int xt_csrapplybc(int *ia, int *ja, double *a, int n, int *nz, const int *bc);

/* Values of the 1-based CSR (ia, ja, a) of order n in the pattern (pia, pja) of the same
 * order (both with rows sorted by column): out[k] = a at pattern entry k, 0 where the
 * matrix has no entry; XT_E_ARG when the matrix has an entry outside the pattern.  For
 * the matrices a deletion event leaves (delelement / adddiagonal / csrapplybc only drop
 * entries of the assembled pattern or set its diagonal), so xt_op_set_values and
 * xt_prec_refactor_band update the device handles in place.  No counterpart in the
 * reference, which applies the constraints to the edited K and M (csrapplybc,
 * main3d.f90:684-685) and refactors with MUMPS (main3d.f90:687). */
int xt_csr_values_in_pattern(int n, const int *pia, const int *pja, const int *ia,
                             const int *ja, const double *a, double *out);

/* ---- 2. device-resident handle API ------------------------------------------- */

typedef struct xt_op_s *xt_op;        /* Kronecker operator sum_k T_k (x) S_k   */
typedef struct xt_prec_s *xt_prec;    /* ILUT factors + level schedule          */

/* Upload nterms spatial matrices (1-based CSR, host) of order p.  FP64 and FP32
 * copies are kept on the device.  q <= 8, nterms <= 3. */
int xt_op_create(int p, int q, int nterms, const int *const *rowptr,
                 const int *const *colidx, const double *const *vals, xt_op *out);
void xt_op_destroy(xt_op op);
/* nnz of term k (CSR form). */
long long xt_op_nnz(xt_op op, int k);
/* Kernel selection: 0 (default) = node-block form when the operator has one (two
 * terms K, M of a 3-DOF mesh with diagonal M blocks, q = 6), 1 = CSR form. */
int xt_op_set_format(xt_op op, int format);
/* Format used by xt_op_apply: 0 node blocks, 1 CSR. */
int xt_op_format(xt_op op);
/* New values of every term, in the CSR pattern the operator was created with (device
 * arrays of nnz(k) doubles, e.g. from xt_asm_values): both copies and the node blocks
 * are updated on the device.  No counterpart in the reference, which rebuilds its
 * matrices every slab (main3d.f90:752-783). */
int xt_op_set_values(xt_op op, const double *const *vals, void *stream);
/* Stored node blocks (real != 0: without slice padding); 0 without node-block form. */
long long xt_op_blocks(xt_op op, int real);

/* Y := alpha * sum_k S_k X T_{k,r}^T + beta * Y on R right-hand sides.
 * tmat: device array [R][nterms][q][q] (row-major T_{k,r}(i,j)), same dtype as
 * x/y.  fp32 != 0 selects the FP32 values and float X/Y.  stream: hipStream_t. */
int xt_op_apply(xt_op op, int R, const void *tmat, double alpha, double beta,
                const void *x, void *y, int fp32, void *stream);

/* Upload ILUT factors (1-based MSR from ilut) and build the level schedule. */
int xt_prec_create(int p, const double *alu, const int *jlu, const int *ju, xt_prec *out);
void xt_prec_destroy(xt_prec pr);
/* Exact factor K^-1 of a symmetric positive definite banded K (1-based CSR, host;
 * the lower triangle is used), applied by block sweeps on the device.  Stands in
 * for the reference's live preconditioner, the exact MUMPS factor of K
 * (mumps_solver, mumps.f90:22-145; mpilusol, kronecker_mumps.f90:120-183), and is
 * ilut's lfil >= n, droptol = 0 limit (ilut.f:2169-2170).  K is cut into diagonal
 * blocks of max(block, bandwidth) rows rounded up to a multiple of 64 (at most
 * 4096: number the mesh along its longest axis last, or renumber it by reverse
 * Cuthill-McKee); factorization in FP64 on the
 * device.  XT_E_SINGULAR if K is not positive definite. */
int xt_prec_create_band(int p, const int *rowptr, const int *colidx, const double *vals,
                        int block, xt_prec *out);
/* Exact factor K^-1 as xt_prec_create_band, plus the factors (K + s_g M)^-1 for the
 * nshift positive shifts s_g (same blocks; K and M 1-based CSR, host, rows sorted by
 * column).  A slab solve with this preconditioner splits each load case's temporal
 * coupling C_r = A_r^-1 B_r: its dominant eigenvalue lambda_r (real, positive: the
 * mass-like temporal mode that grows with the load frequency) with right and left
 * eigenvectors v_r, u_r (u_r . v_r = 1) is preconditioned by the factor whose shift is
 * nearest lambda_r (log scale; none when lambda_r is below a quarter of the smallest
 * shift), the rest by K^-1:
 *   P_r^-1 = [(I - v_r u_r^T) (x) K^-1 + v_r u_r^T (x) (K + s M)^-1] (A_r^-1 (x) I).
 * With s = lambda_r this inverts the mode's spatial operator K + lambda_r M exactly.
 * No counterpart in the reference, whose preconditioner A^-1 (x) K^-1 (mpilusol,
 * kronecker_mumps.f90:120-183) drops the mass term B (x) M: at load frequencies where
 * lambda_r reaches the structure's lowest K^-1 M eigenvalues, GMRES with a factor
 * rounded to FP32 stagnates. */
#define XT_SHIFT_FP32_ONLY 1   /* flags: keep the shifted factors in FP32 only (a third of
                                  the memory); FP64 applies then use A^-1 (x) K^-1 */
#define XT_BAND_FP32_ONLY 2    /* flags: keep the factor of K itself in FP32 only: the
                                  preconditioner then applies in FP32 only (GMRES-IR without
                                  FP64 fallback, FGMRES with FP32 sweeps); nshift may be 0 */
int xt_prec_create_band_shifted(int p, const int *krow, const int *kcol, const double *kval,
                                const int *mrow, const int *mcol, const double *mval,
                                int nshift, const double *shifts, int block, int flags,
                                xt_prec *out);
/* Exact factor K^-1 of a block-diagonal K whose diagonal blocks are nseg independent
 * systems (meshes) on the consecutive DOF ranges seg_dof[s]..seg_dof[s+1]-1 (seg_dof
 * [nseg+1], seg_dof[0] = 0, seg_dof[nseg] = p; 1-based CSR on the host, no entry may
 * couple two ranges: XT_E_ARG).  The ranges are grouped into runs of consecutive
 * systems and each run gets a banded factor of its own, with the block size of its
 * widest band, so a batch of uneven meshes does not pad every mesh to the widest band
 * of the batch (order the systems by bandwidth for the fewest, tightest runs).  The
 * runs are chosen to minimise the factor bytes an apply streams plus a per-run launch
 * cost.  flags: XT_BAND_FP32_ONLY as for xt_prec_create_band_shifted.  The temporal
 * split (shifted factors) and xt_prec_refactor_band are not available for it.  No
 * counterpart in the reference, which factors one mesh per process (mumps.f90:22-145). */
int xt_prec_create_band_seg(int p, const int *rowptr, const int *colidx, const double *vals,
                            int nseg, const int *seg_dof, int block, int flags, xt_prec *out);
/* Runs of a preconditioner from xt_prec_create_band_seg (1 for any other banded factor,
 * 0 for ILUT): ngroup, and when non-NULL the first DOF of each run and the end
 * (goff [ngroup+1]) and each run's block size (gblock [ngroup]). */
int xt_prec_band_groups(xt_prec pr, int *ngroup, int *goff, int *gblock);
/* Number of shifted factors of a preconditioner and (when shifts != NULL) their shifts. */
int xt_prec_shifts(xt_prec pr, int *nshift, double *shifts);
/* Refactor a banded-factor preconditioner with new values of K (1-based CSR in device
 * memory, within the band it was created for: XT_E_ARG otherwise).  Replaces the
 * per-slab refactorization mumps_solver(job = 1), main3d.f90:783. */
int xt_prec_refactor_band(xt_prec pr, const int *rowptr, const int *colidx,
                          const double *vals, void *stream);
/* xt_prec_refactor_band for a K that changed only in rows and columns >= first_row
 * (0-based; the caller guarantees it, e.g. the DOFs of deleted elements and of the nodes
 * they orphan): the factor's blocks before first_row / block are kept and the rest, with
 * the recurrence and pairing matrices that depend on them, recomputed from the kept
 * D_{j-1} and K (equal to a full refactorization up to rounding).  A factor kept in FP32
 * only is refactored whole.  Plain banded factors only (no shifted or grouped factors).
 * Replaces the refactorization after element deletion, mumps_solver(job = 1),
 * main3d.f90:687. */
int xt_prec_refresh_band(xt_prec pr, const int *rowptr, const int *colidx,
                         const double *vals, int first_row, void *stream);
/* Block size, number of blocks and bandwidth of a banded-factor preconditioner (of a
 * grouped one: the largest block, the blocks of all runs, the largest bandwidth). */
int xt_prec_band_info(xt_prec pr, int *block, int *nblocks, int *bandwidth);
/* Sequential steps of the two sweeps: ILUT levels, or blocks of a banded factor. */
int xt_prec_levels(xt_prec pr, int *nlev_lower, int *nlev_upper);

/* ---- spatial assembly on the device ------------------------------------------ */

typedef struct xt_asm_s *xt_asm;      /* mesh, global patterns, element buffers  */

/* Symbolic assembly of a C3D8 mesh, once per mesh: xcoord [nn][3] and xconn [ne][8]
 * (1-based node ids, shape3x node order) row-major host arrays; fixed [3*nn] nonzero
 * for the constrained DOFs (DOF 3*(node-1) + component).  Builds the global patterns
 * of K (all couplings of each element's DOFs) and M (equal components only) and, for
 * each global entry, its element contributions in element order.  XT_E_ARG for an
 * element with a non-positive Jacobian determinant. */
int xt_asm_create(int nn, int ne, const double *xcoord, const int *xconn, const int *fixed,
                  xt_asm *out);
void xt_asm_destroy(xt_asm h);
/* Host copy of a pattern (which 0: K, 1: M, 2: thermal K_t and M_t, after
 * xt_asm_thermal): nnz, then 1-based rowptr [3*nn+1, or nn+1 for which 2] and colidx
 * [nnz] when non-NULL. */
int xt_asm_pattern(xt_asm h, int which, long long *nnz, int *rowptr, int *colidx);
/* Numeric assembly on the device for per-Gauss-point E (ym) and nu (device [8*ne],
 * element-major) and density rho: the element matrices of elemtx_m
 * (elesubs.f90:2-58, stiffness per point as setsdstiff builds it), summed per entry in
 * element order as tricsr does (spmtx.f90:359-407), the 1-0 method of csrapplybc
 * (spmtx.f90:268-307) on the fixed DOFs, and the shape-derivative tables Nx, Ny, Nz
 * [8*ne][8].  Replaces mpiglbmtx_m + tricsr + csrapplybc (mpisubs.f90:69-139,
 * main3d.f90:752-781).  Bit-identical values; entries that sum to zero stay in the
 * pattern (the reference drops them). */
int xt_asm_run(xt_asm h, const double *ym, const double *nu, double rho, void *stream);
/* Device arrays of the last xt_asm_run: which 0 K values, 1 M values (in the patterns'
 * order), 2/3/4 Nx/Ny/Nz; of xt_asm_thermal: 5 K_t and 6 M_t with the 1-0 method on the
 * fixed temperatures, 7 K_t and 8 M_t without it (the ki0/kv0 copies of
 * main3d.f90:444-449). */
int xt_asm_values(xt_asm h, int which, const double **vals);
/* Copy array `which` (as xt_asm_values) into device memory dst (nnz or 64*ne doubles). */
int xt_asm_copy(xt_asm h, int which, double *dst, void *stream);
/* Device copies of a pattern's 1-based column indices and 0-based rows of its entries. */
int xt_asm_device_pattern(xt_asm h, int which, const int **colidx, const int **row);
/* Thermal conductivity K_t and capacity M_t (one DOF a node) on the device: elemtx_t
 * for every element (elesubs.f90:60-116) summed per entry in element order as
 * mpiglbmtx_t + tricsr build them (mpisubs.f90:141-188, main3d.f90:365-375), then the
 * 1-0 method of csrapplybc on the nodes with tfixed [nn] (host) nonzero
 * (main3d.f90:452-453).  Replaces that setup block.  The first call builds the
 * node-level pattern; later calls recompute the values.  Bit-identical values. */
int xt_asm_thermal(xt_asm h, const int *tfixed, double kappa, double cp, double rho,
                   void *stream);
/* One slab's temperature-dependent material and thermal load on the device
 * (main3d.f90:758-768): material_update (thermoelastic.f90:1-46) writes E, nu [8*ne]
 * to ym, nu; calc_thermo_stress (thermoelastic.f90:48-78) and elevec_m
 * (elesubs.f90:119-167) per element, summed per DOF in element order as mpiglbmtx_m
 * sums vv (mpisubs.f90:124-131), write vv [3*nn].  T: device FP64 with the temperature
 * of node n (0-based) at T[n*tstride + toff] (0 <= toff < tstride), e.g. the third
 * temporal node of an interleaved thermal solution (tstride 3R, toff 3r + 2).
 * alpha, Tref, ym0, nu0 as the driver's (main3d.f90:145-157).  Bit-identical. */
int xt_asm_thermo_load(xt_asm h, const double *T, long long tstride, long long toff, double Tref,
                       double alpha, double ym0, double nu0, double *ym, double *nu, double *vv,
                       void *stream);

/* Y = (LU)^-1 X Minv_r^T on R right-hand sides (Minv: device [R][q][q],
 * row-major, the inverse temporal matrix; NULL = identity).  x and y must not
 * alias.  fp32 selects the FP32 factor copy. */
int xt_prec_apply(xt_prec pr, int R, int q, const void *minv, const void *x, void *y,
                  int fp32, void *stream);

/* xt_prec_apply for ILUT factors without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The temporal
 * mix is staged in a fixed per-device array of XT_NOALLOC_SCRATCH_BYTES held until the
 * process ends, so p*q*R*sizeof(value) must not exceed it when minv != NULL (XT_E_ARG
 * otherwise); no kernel timing.  Banded factors: XT_E_UNSUPPORTED.  Every call on one
 * device shares that array: calls that overlap on one device (other streams or threads)
 * race.  Same results as xt_prec_apply (the same kernels). */
#define XT_NOALLOC_SCRATCH_BYTES (16u << 20)
int xt_prec_apply_noalloc(xt_prec pr, int R, int q, const void *minv, const void *x, void *y,
                          int fp32, void *stream);

/* Layout conversion between temporal-major x[r][j*p + dof] (R stacked
 * reference vectors) and interleaved X[dof][r*q + j], FP64. */
int xt_to_interleaved(int p, int q, int R, const double *src, double *dst, void *stream);
int xt_from_interleaved(int p, int q, int R, const double *src, double *dst, void *stream);

/* Batched slab solve.  Solves (A_r (x) K + B_r (x) M) x_r = b_r for r < R with
 * GMRES(im) right-preconditioned by A_r (x) P, P the ILUT factors or the exact banded
 * factor of K (preconditioner handle).
 *   mode 0: FP64 throughout (plain FP64 GPU GMRES, the reference's algorithm);
 *   mode 1: GMRES iterative refinement: FP64 true residual b - A x outside, one
 *           flexible GMRES(im) cycle per refinement step inside with the FP32
 *           operator and FP32 preconditioner sweeps (FP64 Arnoldi basis,
 *           orthogonalisation and Givens least squares);
 *   mode 2: FP64 flexible GMRES(im) with the FP32 preconditioner sweeps.
 * ab: host [R][2][q][q] column-major A_r, B_r (as passed to kronpgmres).
 * b, x: device, interleaved, FP64.  x is the initial guess on entry.
 * Outputs per r (host arrays of length R): its, ierr, relres (true
 * ||b - A x|| / ||b||, FP64).  eps is relative to the initial residual.  pr must have
 * op's p rows (XT_E_ARG otherwise, before any device work). */
int xt_slab_solve(xt_op op, xt_prec pr, int R, const double *ab, const double *b,
                  double *x, int im, double eps, int maxits, int mode, int *its,
                  int *ierr, double *relres, void *stream);

/* xt_slab_solve with a forward-error test for mode 1: a case converges when its true
 * relative residual meets eps and its last refinement correction is at most xtol times
 * its solution norm (xtol = 0: the residual test alone, as xt_slab_solve).  At load
 * frequencies where the slab operator is ill-conditioned the residual bounds the error
 * of the displacement only loosely; the correction of a converging refinement does
 * (each step contracts the error).  Modes 0 and 2 ignore xtol.  No counterpart in the
 * reference (kronpgmres stops on the residual, kronecker.f90:312). */
int xt_slab_solve_tol(xt_op op, xt_prec pr, int R, const double *ab, const double *b,
                      double *x, int im, double eps, double xtol, int maxits, int mode,
                      int *its, int *ierr, double *relres, void *stream);

/* Batched slab solve of nseg independent systems stacked in one block-diagonal
 * operator (e.g. meshes of different sizes: their K and M concatenated, one banded
 * factor of the block-diagonal K), rows seg_dof[s] .. seg_dof[s+1]-1 for system s
 * (host array of nseg + 1 increasing values, 0 and p at the ends).  Every (system s,
 * load case r) is its own instance g = s*R + r: its own Krylov space, Hessenberg,
 * rotations, residual norms and convergence (a converged instance is masked: its basis
 * columns are zero and its solution frozen), while every kernel of a step covers all
 * instances at once.  The temporal matrices are per load case, shared by the systems.
 * its, ierr, relres: host arrays of nseg*R.  Otherwise as xt_slab_solve_tol.  No
 * counterpart in the reference, which solves one mesh per run. */
int xt_slab_solve_seg(xt_op op, xt_prec pr, int nseg, const long long *seg_dof, int R,
                      const double *ab, const double *b, double *x, int im, double eps,
                      double xtol, int maxits, int mode, int *its, int *ierr, double *relres,
                      void *stream);

/* mpipgmres (kronecker_mumps.f90:186-415): the reference's live solve path, GMRES
 * right-preconditioned by A^-1 (x) K^-1 with K^-1 from the MUMPS factor, on host arrays.
 * Same arguments as the reference with the preconditioner handle pr in place of
 * mumps_par (the driver's mumps_solver job 0, main3d.f90:550, becomes
 * xt_prec_create_band; job 1 after re-assembly or deletion, main3d.f90:687 and 783,
 * becomes xt_prec_refactor_band or xt_prec_refresh_band; any preconditioner handle
 * works) and no MPI (one device holds all of K and M).  The operator is built from
 * the CSR arrays per call, rhs and sol staged to the device and back; FP64 throughout,
 * the reference's algorithm (xt_slab_solve mode 0).  Outputs and residual history as
 * xt_kronpgmres. */
int xt_mpipgmres(const double *a, const double *b, int q,
                 const int *ki, const int *kj, const double *kv,
                 const int *mi, const int *mj, const double *mv, int p,
                 xt_prec pr, double *rhs, double *sol, int im, double eps, int maxits,
                 double *hist, int hist_cap, int *nhist, int *its, int *ierr);

/* Iterations of the last slab solve on op (any of the entry points above), summed over
 * its instances: it32 those whose preconditioner sweeps ran in FP32 (GMRES-IR's FP32
 * cycles, mode 2), it64 those in FP64 (mode 0, GMRES-IR's FP64 cycles).  No counterpart
 * in the reference (one precision). */
int xt_slab_stats(xt_op op, long long *it32, long long *it64);

/* chkelesta (delelement.f90:109-136) on device-resident state: failed[e] = 1 iff
 * some Gauss point of element e has D >= Dc, else 0; for a failed element
 * elesta[e] and its gpsta entries are set to 0 (others untouched).  Device
 * pointers, ne*ngpe Gauss points element-major. */
int xt_elestatus_device(int ne, int ngpe, double Dc, const double *D, int *elesta,
                        int *gpsta, int *failed, void *stream);

/* D = ws = 0 at every Gauss point with gpsta == 0 (kokkos_damage.f90:102-107),
 * device pointers. */
int xt_damage_zero_failed(int ngp, const int *gpsta, double *D, double *ws, void *stream);

/* Damage kernel on device-resident state (kokkos_kernel.h:44-344), one lane per
 * Gauss point, nip temporal points per launch.  All pointers are device
 * pointers with kokkos_host_'s layouts; sol is temporal-major FP64. */
int xt_damage_device(int ne, int nn, int nip, const int *nodes, const double *Nx,
                     const double *Ny, const double *Nz, const double *Nt,
                     const double *sol, const double *ym, const double *nu,
                     double *eps_old, double *eps_pl, double *sig_eff, double *sig_back,
                     double *p, double *D, double *ws, double *del_po, double *sig_d,
                     void *stream);

/* xt_damage_device with other constants of the damage law: params[7] = sig_f, sig_u,
 * eps_D, Cy, h, S, s (damage.h:9-17 names; NULL: the reference's 180, 577, 0.08, 1740,
 * 0.2, 0.5, 0.5).  The reference compiles them in; a material with another fatigue
 * limit (e.g. a 350 MPa yield) needs them at run time. */
int xt_damage_device_params(int ne, int nn, int nip, const int *nodes, const double *Nx,
                            const double *Ny, const double *Nz, const double *Nt,
                            const double *sol, const double *ym, const double *nu,
                            double *eps_old, double *eps_pl, double *sig_eff, double *sig_back,
                            double *p, double *D, double *ws, double *del_po, double *sig_d,
                            const double *params, void *stream);

/* Right-hand side of a slab for R load cases (main3d.f90:803-823):
 *   b[dof][r*q + j] := fixed[dof] ? 0 : b[dof][r*q + j] + coef[r*q + j] * rext_r[dof]
 * i.e. fextc_r(j) * amplitude_r * rext_r added to the jump term already in b (from
 * xt_op_apply with the ktn1/mtn1 temporal matrices), then the fixed DOFs zeroed.
 * coef: device [R*q]; rext: device, rext_r = rext + r * rext_ld (rext_ld = 0: one
 * load vector for every case, else >= p); fixed: device uint8 [p] or NULL.  FP64. */
int xt_slab_rhs(int p, int q, int R, const double *coef, const double *rext, long long rext_ld,
                const unsigned char *fixed, double *b, void *stream);

/* ---- kernel timing ----------------------------------------------------------- */
#define XT_PROF_SPMM 0     /* Kronecker operator (node-block or CSR kernel)         */
#define XT_PROF_BAND 1     /* banded-factor block sweeps in FP32                    */
#define XT_PROF_TRSV 2     /* ILUT level-scheduled triangular solves                */
#define XT_PROF_DAMAGE 3   /* damage kernel                                         */
#define XT_PROF_BAND64 4   /* banded-factor block sweeps in FP64                    */
#define XT_PROF_NKINDS 5
/* Enable per-launch HIP-event timing for the kinds in mask (bit k = kind k) and
 * clear the counters; 0 disables and releases the events.  Process-wide; one host
 * thread. */
int xt_prof_enable(unsigned mask);
/* Sum of per-launch kernel durations (ms), number of launches, and the launches'
 * algorithmic bytes and flops since xt_prof_enable.  Synchronises on the last event. */
int xt_prof_read(int kind, double *ms, long long *launches, double *bytes, double *flops);

/* Algorithmic bytes and flops recorded since xt_prof_enable: the work of the unpaired
 * computation.  The banded-factor sweeps perform more flops than that (the pairing
 * products of their recurrences); for the other kinds it equals xt_prof_read's. */
int xt_prof_read_alg(int kind, double *bytes, double *flops);

/* Replay the banded factor's sweeps (2 + 2 (nb - 1) launches of fixed buffers) from a
 * HIP graph captured on first use per column count and precision (on = 1, the
 * default), or launch them one by one (0).  No counterpart in the reference (launch
 * overhead control).  Process-wide. */
int xt_graphs_enable(int on);

/* Kernel-variant knobs for the autotuner (xtfem_amd.tune times the candidates on the
 * caller's operator and preconditioner and sets the fastest).  0 = the built-in choice.
 * Process-wide; set them between solves, not during one.  No counterpart in the
 * reference.
 *   XT_TUNE_SPMM_UNR      node-block SpMM: entries per pass (loads in flight)   2, 4, 8
 *   XT_TUNE_SPMM_WPB      node-block SpMM: waves per workgroup                  1, 2, 4
 *   XT_TUNE_SPMM_R1SPLIT  node-block SpMM, one load case: lanes per node row    4, 8
 *   XT_TUNE_BAND_TILE     banded-factor sweeps, batched passes: tile rows       64, 128
 *                         (128: FP32 with block rows a multiple of 128)
 *   XT_TUNE_GMRES_CHECK   slab solve: Arnoldi steps between host convergence checks 1, 2, 4
 *   XT_TUNE_SPMM_LDS      node-block SpMM, several load cases: the LDS-staged kernel
 *                         (neighbour rows of X staged per slice) off 1, on 2 (where the
 *                         neighbour rows fit), built-in: off
 *   XT_TUNE_BAND_SMALLC   banded-factor sweeps with at most 16 columns: VALU rows kernel 1,
 *                         matrix-core kernels with 64-column tiles in batched passes 2,
 *                         with 16-column tiles 3; built-in: rows kernel below 1024-row
 *                         blocks (FP32) or 512-row blocks (FP64), else 3    1, 2, 3
 *   XT_TUNE_IR_FP32_CYCLE GMRES-IR: iterations an FP32 inner cycle gets to meet its target
 *                         before the case moves to FP64 cycles              1, 2, 4, 8
 *                         (built-in 2)
 *   XT_TUNE_IR_MIN_GAIN   GMRES-IR: least reduction of the true residual an FP32
 *                         refinement step must reach for the case to stay in FP32
 *                         cycles (when the preconditioner has FP64 sweeps)  2, 8, 32, 128
 *                         (built-in 16)
 *   XT_TUNE_IR_SKIP32     GMRES-IR precision probe per operator and preconditioner (with
 *                         FP64 sweeps): one solve FP32 cycles first, the next FP64 cycles
 *                         only, the following n - 2 the faster of the two by wall time,
 *                         then again; 1: FP32 cycles first always           1, 4, 16, 64
 *                         (built-in 16)
 *   XT_TUNE_BAND_CHAIN    banded-factor sweeps, the dependent chain steps with more than
 *                         16 columns: one workgroup per 16 x 16 output tile 1, 64 x 64
 *                         tiles split over k into about 512 / 256 / 1024 workgroups with
 *                         a fixed-order sum 2 / 3 / 4 (built-in 4)          1, 2, 3, 4
 *   XT_TUNE_BAND_LEVELS   banded factor: pairing levels of the block recurrences at most
 *                         v - 1 (read when a factor is built; built-in 4)   1, 2, 3, 5
 *   XT_TUNE_GMRES_REORTH  Arnoldi orthogonalisation: classical Gram-Schmidt twice always
 *                         1, the second pass only where the first lost more than half of
 *                         |w|^2 for some instance 2 (built-in 2)            1, 2
 *   XT_TUNE_GMRES_COMPACT banded-factor sweeps of a batched solve over the load cases still
 *                         iterating in the cycle only 2, over all 1 (built-in 2)   1, 2
 *   XT_TUNE_BAND_CHAIN_FEW as XT_TUNE_BAND_CHAIN for chain steps with at most 16 columns on
 *                         the matrix cores: one 16 x 16 tile per workgroup 1, split-k 64 x
 *                         64 tiles 2 (built-in 1)                            1, 2
 *   XT_TUNE_BAND_PAIR     banded-factor sweeps: the factor's pairing levels never 1,
 *                         always 2 (built-in: FP32 always, FP64 up to 64 columns)   1, 2
 * XT_E_ARG for a value a knob does not take. */
#define XT_TUNE_SPMM_UNR 0
#define XT_TUNE_SPMM_WPB 1
#define XT_TUNE_SPMM_R1SPLIT 2
#define XT_TUNE_BAND_TILE 3
#define XT_TUNE_GMRES_CHECK 4
#define XT_TUNE_SPMM_LDS 5
#define XT_TUNE_BAND_SMALLC 6
#define XT_TUNE_IR_FP32_CYCLE 7
#define XT_TUNE_IR_MIN_GAIN 8
#define XT_TUNE_IR_SKIP32 9
#define XT_TUNE_BAND_CHAIN 10
#define XT_TUNE_BAND_LEVELS 11
#define XT_TUNE_GMRES_REORTH 12
#define XT_TUNE_GMRES_COMPACT 13
#define XT_TUNE_BAND_CHAIN_FEW 14
#define XT_TUNE_BAND_PAIR 15
#define XT_TUNE_NKNOBS 16
int xt_tune_set(int knob, int value);
/* Current value of a knob (0 = built-in choice), -1 for an unknown knob. */
int xt_tune_get(int knob);

/* Synchronise the device (all streams of this process). */
int xt_device_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* XTFEM_H */
