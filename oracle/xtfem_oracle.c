/*
 * xtfem_oracle.c -- CPU restatement of the XTFEM space-time solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * the smoke check and the CPU-baseline leg of bench.py may load it.  The
 * product (xtfem_amd/) never links, imports or calls it.
 *
 * Every function restates one routine of the reference (pourmatin/XTFEM)
 * in plain C99, FP64, keeping the reference's loop order so that results
 * agree with the compiled reference to rounding.  Conventions follow the
 * reference exactly:
 *   - CSR arrays are 1-based (ia[0] == 1), as produced by tricsr/csrapplybc
 *     (spmtx.f90:359-454);
 *   - q x q temporal matrices are column-major (Fortran a(i,j) == a[i + j*q]);
 *   - space-time vectors are temporal-major: x[j*p + dof], j = 0..q-1
 *     (kronecker.f90:64, kokkos_kernel.h:67).
 *
 * Build: cc -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * Contraction is disabled so each multiply and add rounds separately, as in
 * the reference's x86-64 build without FMA.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* SPARSKIT amux: y = A x, CSR, row-sequential accumulation.  ilut.f:3-44     */
/* ------------------------------------------------------------------------ */
void or_amux(int n, const double *x, double *y, const double *a,
             const int *ja, const int *ia)
{
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k)
            acc = acc + a[k] * x[ja[k] - 1];
        y[i] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* kronspmv: y = (A (x) K + B (x) M) x.  kronecker.f90:19-85                  */
/* u_j = K x_j for all j, then y_i += a(i,j) u_j (i outer, j inner); same     */
/* for M with b.                                                             */
/* ------------------------------------------------------------------------ */
void or_kronspmv(const double *a, const double *b, int q,
                 const int *ki, const int *kj, const double *kv,
                 const int *mi, const int *mj, const double *mv,
                 int p, const double *x, double *y)
{
    size_t n = (size_t)p * q;
    double *u = (double *)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) y[k] = 0.0;
    for (int i = 0; i < q; ++i) or_amux(p, x + (size_t)i * p, u + (size_t)i * p, kv, kj, ki);
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) {
            double s = a[i + j * q];
            double *yi = y + (size_t)i * p;
            const double *uj = u + (size_t)j * p;
            for (int k = 0; k < p; ++k) yi[k] = yi[k] + s * uj[k];
        }
    for (int i = 0; i < q; ++i) or_amux(p, x + (size_t)i * p, u + (size_t)i * p, mv, mj, mi);
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) {
            double s = b[i + j * q];
            double *yi = y + (size_t)i * p;
            const double *uj = u + (size_t)j * p;
            for (int k = 0; k < p; ++k) yi[k] = yi[k] + s * uj[k];
        }
    free(u);
}

/* kronspgemv: y := alpha (A(x)K + B(x)M) x + beta y.  kronecker.f90:88-127   */
void or_kronspgemv(const double *a, const double *b, double alpha, double beta, int q,
                   const int *ki, const int *kj, const double *kv,
                   const int *mi, const int *mj, const double *mv,
                   int p, const double *x, double *y)
{
    size_t n = (size_t)p * q;
    double *z = (double *)malloc(n * sizeof(double));
    or_kronspmv(a, b, q, ki, kj, kv, mi, mj, mv, p, x, z);
    for (size_t k = 0; k < n; ++k) y[k] = alpha * z[k] + beta * y[k];
    free(z);
}

/* ------------------------------------------------------------------------ */
/* qsplit: quick-split so |a(i)| >= |a(ncut)| for i < ncut and <= after.     */
/* ilut.f:2487-2545.  a, ind are 1-based views of length n.                  */
/* ------------------------------------------------------------------------ */
static void or_qsplit(double *a, int *ind, int n, int ncut)
{
    /* shift to 1-based indexing to follow the reference literally */
    double *A = a - 1;
    int *I = ind - 1;
    int first = 1, last = n;
    if (ncut < first || ncut > last) return;
    for (;;) {
        int mid = first;
        double abskey = fabs(A[mid]);
        for (int j = first + 1; j <= last; ++j) {
            if (fabs(A[j]) > abskey) {
                mid = mid + 1;
                double tmp = A[mid]; int itmp = I[mid];
                A[mid] = A[j]; I[mid] = I[j];
                A[j] = tmp; I[j] = itmp;
            }
        }
        double tmp = A[mid]; A[mid] = A[first]; A[first] = tmp;
        int itmp = I[mid]; I[mid] = I[first]; I[first] = itmp;
        if (mid == ncut) return;
        if (mid > ncut) last = mid - 1; else first = mid + 1;
    }
}

/* ------------------------------------------------------------------------ */
/* ilut: Saad's dual-threshold ILU.  ilut.f:152-495.                          */
/* Inputs a/ja/ia: 1-based CSR.  Outputs MSR alu/jlu (1-based, diagonal       */
/* inverted in alu[0..n-1], jlu[0] = n+2) and ju (1-based start of U).        */
/* Work arrays w (n+1) and jw (2n) are supplied by the caller as in the       */
/* reference.  Returns ierr with the reference's meaning.                     */
/* ------------------------------------------------------------------------ */
int or_ilut(int n, const double *a, const int *ja, const int *ia, int lfil,
            double droptol, double *alu, int *jlu, int *ju, int iwk,
            double *w, int *jw)
{
    /* 1-based views */
    const double *A = a - 1; const int *JA = ja - 1; const int *IA = ia - 1;
    double *ALU = alu - 1; int *JLU = jlu - 1; int *JU = ju - 1;
    double *W = w - 1; int *JW = jw - 1;
    if (lfil < 0) return -4;                                   /* ilut.f:239 */
    int ju0 = n + 2;
    JLU[1] = ju0;
    for (int j = 1; j <= n; ++j) JW[n + j] = 0;
    for (int ii = 1; ii <= n; ++ii) {                           /* ilut.f:256 */
        int j1 = IA[ii], j2 = IA[ii + 1] - 1;
        double tnorm = 0.0;
        for (int k = j1; k <= j2; ++k) tnorm = tnorm + fabs(A[k]);
        if (tnorm == 0.0) return -5;
        /* real(j2-j1+1) is a default (FP32) real; exact for any row length */
        tnorm = tnorm / (double)(float)(j2 - j1 + 1);
        int lenu = 1, lenl = 0;
        JW[ii] = ii; W[ii] = 0.0; JW[n + ii] = ii;
        for (int j = j1; j <= j2; ++j) {                        /* ilut.f:278 */
            int k = JA[j]; double t = A[j];
            if (k < ii) {
                lenl = lenl + 1; JW[lenl] = k; W[lenl] = t; JW[n + k] = lenl;
            } else if (k == ii) {
                W[ii] = t;
            } else {
                lenu = lenu + 1; int jpos = ii + lenu - 1;
                JW[jpos] = k; W[jpos] = t; JW[n + k] = jpos;
            }
        }
        int jj = 0, len = 0;
        for (;;) {                                              /* ilut.f:300 */
            jj = jj + 1;
            if (jj > lenl) break;
            int jrow = JW[jj], k = jj;
            for (int j = jj + 1; j <= lenl; ++j)
                if (JW[j] < jrow) { jrow = JW[j]; k = j; }
            if (k != jj) {
                int j = JW[jj]; JW[jj] = JW[k]; JW[k] = j;
                JW[n + jrow] = jj; JW[n + j] = k;
                double s = W[jj]; W[jj] = W[k]; W[k] = s;
            }
            JW[n + jrow] = 0;
            double fact = W[jj] * ALU[jrow];                    /* ilut.f:338 */
            if (fabs(fact) <= droptol) continue;                /* absolute drop */
            for (int kk = JU[jrow]; kk <= JLU[jrow + 1] - 1; ++kk) {
                double s = fact * ALU[kk];
                int j = JLU[kk];
                int jpos = JW[n + j];
                if (j >= ii) {
                    if (jpos == 0) {
                        lenu = lenu + 1;
                        if (lenu > n) return -1;
                        int i = ii + lenu - 1;
                        JW[i] = j; JW[n + j] = i; W[i] = -s;
                    } else {
                        W[jpos] = W[jpos] - s;
                    }
                } else {
                    if (jpos == 0) {
                        lenl = lenl + 1;
                        if (lenl > n) return -1;
                        JW[lenl] = j; JW[n + j] = lenl; W[lenl] = -s;
                    } else {
                        W[jpos] = W[jpos] - s;
                    }
                }
            }
            len = len + 1; W[len] = fact; JW[len] = jrow;
        }
        for (int k = 1; k <= lenu; ++k) JW[n + JW[ii + k - 1]] = 0;  /* ilut.f:400 */
        lenl = len;
        len = lenl < lfil ? lenl : lfil;
        or_qsplit(&W[1], &JW[1], lenl, len);
        for (int k = 1; k <= len; ++k) {                        /* ilut.f:414 */
            if (ju0 > iwk) return -2;
            ALU[ju0] = W[k]; JLU[ju0] = JW[k]; ju0 = ju0 + 1;
        }
        JU[ii] = ju0;
        len = 0;
        for (int k = 1; k <= lenu - 1; ++k) {                   /* ilut.f:429 */
            if (fabs(W[ii + k]) > droptol * tnorm) {
                len = len + 1; W[ii + len] = W[ii + k]; JW[ii + len] = JW[ii + k];
            }
        }
        lenu = len + 1;
        len = lenu < lfil ? lenu : lfil;
        or_qsplit(&W[ii + 1], &JW[ii + 1], lenu - 1, len);
        if (len + ju0 > iwk) return -3;                         /* ilut.f:443 */
        for (int k = ii + 1; k <= ii + len - 1; ++k) {
            JLU[ju0] = JW[k]; ALU[ju0] = W[k]; ju0 = ju0 + 1;
        }
        /* zero pivot: (0.0001 + droptol)*tnorm with 0.0001 an FP32 literal */
        if (W[ii] == 0.0) W[ii] = ((double)0.0001f + droptol) * tnorm;
        ALU[ii] = 1.0 / W[ii];                                  /* ilut.f:457 */
        JLU[ii + 1] = ju0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* lusol: (LU) x = y with the MSR factors of ilut.  ilut.f:2383-2432.         */
/* In place allowed (x == y).                                                */
/* ------------------------------------------------------------------------ */
void or_lusol(int n, const double *y, double *x, const double *alu,
              const int *jlu, const int *ju)
{
    for (int i = 0; i < n; ++i) {
        double xi = y[i];
        for (int k = jlu[i] - 1; k < ju[i] - 1; ++k) xi = xi - alu[k] * x[jlu[k] - 1];
        x[i] = xi;
    }
    for (int i = n - 1; i >= 0; --i) {
        double xi = x[i];
        for (int k = ju[i] - 1; k < jlu[i + 1] - 1; ++k) xi = xi - alu[k] * x[jlu[k] - 1];
        x[i] = alu[i] * xi;
    }
}

/* ------------------------------------------------------------------------ */
/* inverse: kronecker.f90:368-395 calls LAPACK DGETRF + DGETRI (third-party,  */
/* LAPACK 3.6.1).  Restated as LU with partial pivoting (first maximal |a|,   */
/* as IDAMAX) followed by column-wise solves of A X = I; equal to the LAPACK  */
/* result to rounding.  Returns 0, or 1 if A is singular (reference: STOP).   */
/* ------------------------------------------------------------------------ */
int or_inverse(int n, const double *A, double *Ainv)
{
    double lu[64]; int piv[8];
    if (n > 8) return 2;
    memcpy(lu, A, sizeof(double) * n * n);
    for (int k = 0; k < n; ++k) {
        int pr = k; double mx = fabs(lu[k + k * n]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(lu[i + k * n]) > mx) { mx = fabs(lu[i + k * n]); pr = i; }
        piv[k] = pr;
        if (lu[pr + k * n] == 0.0) return 1;
        if (pr != k)
            for (int j = 0; j < n; ++j) { double t = lu[k + j * n]; lu[k + j * n] = lu[pr + j * n]; lu[pr + j * n] = t; }
        double r = 1.0 / lu[k + k * n];
        for (int i = k + 1; i < n; ++i) lu[i + k * n] *= r;
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) lu[i + j * n] -= lu[i + k * n] * lu[k + j * n];
    }
    for (int c = 0; c < n; ++c) {
        double e[8];
        for (int i = 0; i < n; ++i) e[i] = (i == c) ? 1.0 : 0.0;
        for (int k = 0; k < n; ++k) if (piv[k] != k) { double t = e[k]; e[k] = e[piv[k]]; e[piv[k]] = t; }
        for (int i = 0; i < n; ++i) for (int k = 0; k < i; ++k) e[i] -= lu[i + k * n] * e[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k) e[i] -= lu[i + k * n] * e[k];
            e[i] /= lu[i + i * n];
        }
        for (int i = 0; i < n; ++i) Ainv[i + c * n] = e[i];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* kronlusol: solve (A (x) LU) y = x.  kronecker.f90:130-178.                 */
/* w_j = (LU)^-1 x_j, then y_i = sum_j inva(i,j) w_j.  x and y may alias.     */
/* ------------------------------------------------------------------------ */
int or_kronlusol(const double *a, int q, const double *alu, const int *jlu,
                 const int *ju, int p, const double *x, double *y)
{
    double inva[64];
    if (or_inverse(q, a, inva)) return 1;
    size_t n = (size_t)p * q;
    double *w = (double *)calloc(n, sizeof(double));
    for (int i = 0; i < q; ++i) or_lusol(p, x + (size_t)i * p, w + (size_t)i * p, alu, jlu, ju);
    for (size_t k = 0; k < n; ++k) y[k] = 0.0;
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) {
            double s = inva[i + j * q];
            double *yi = y + (size_t)i * p;
            const double *wj = w + (size_t)j * p;
            for (int k = 0; k < p; ++k) yi[k] = yi[k] + s * wj[k];
        }
    free(w);
    return 0;
}

/* BLAS-1 (reference BLAS 3.6.0; third party): plain left-to-right sums.     */
static double or_ddot(size_t n, const double *x, const double *y)
{
    double s = 0.0;
    for (size_t k = 0; k < n; ++k) s = s + x[k] * y[k];
    return s;
}
static double or_dnrm2(size_t n, const double *x) { return sqrt(or_ddot(n, x, x)); }
static void or_daxpy(size_t n, double t, const double *x, double *y)
{
    for (size_t k = 0; k < n; ++k) y[k] = y[k] + t * x[k];
}

/* ------------------------------------------------------------------------ */
/* kronpgmres: right-preconditioned restarted GMRES(im) with MGS.             */
/* kronecker.f90:181-365.  rhs is destroyed (workspace), sol is the initial   */
/* guess on entry and the solution on exit.  The residual history that the    */
/* reference writes to unit iout ('its', ro/ro0) is returned in hist[] (at    */
/* most hist_cap entries, *nhist set to the count).  Returns ierr:            */
/* 0 converged, 1 maxits reached, -1 zero initial residual, 2 singular A.     */
/* ------------------------------------------------------------------------ */
/* Preconditioner of the GMRES restatement: y = (A (x) P)^-1 x, x may alias y. */
typedef int (*or_prec_fn)(const double *a, int q, const void *ctx, int p, const double *x,
                          double *y);
struct or_ilut_ctx { const double *alu; const int *jlu, *ju; };
static int or_prec_ilut(const double *a, int q, const void *ctx, int p, const double *x,
                        double *y)
{
    const struct or_ilut_ctx *c = (const struct or_ilut_ctx *)ctx;
    return or_kronlusol(a, q, c->alu, c->jlu, c->ju, p, x, y);
}

static int or_kronpgmres_impl(const double *a, const double *b, int q,
                              const int *ki, const int *kj, const double *kv,
                              const int *mi, const int *mj, const double *mv, int p,
                              or_prec_fn prec, const void *pctx,
                              double *rhs, double *sol, int im, double eps, int maxits,
                              double *hist, int hist_cap, int *nhist, int *its_out)
{
    size_t n = (size_t)p * q;
    double *hh = (double *)calloc((size_t)(im + 1) * im, sizeof(double));
    double *c = (double *)calloc(im, sizeof(double));
    double *s = (double *)calloc(im, sizeof(double));
    double *rs = (double *)calloc(im + 1, sizeof(double));
    double *vv = (double *)calloc(n * (im + 1), sizeof(double));
#define HH(r, cc) hh[(r) + (size_t)(cc) * (im + 1)]
#define VV(cc) (vv + (size_t)(cc) * n)
    const double epsmac = 1.0e-16;
    int its = 0, ierr = 0, nh = 0, i = 0;
    double ro, t, eps1 = 0.0, gam;
    double ro0 = or_dnrm2(n, rhs);                              /* :247 */
    or_kronspmv(a, b, q, ki, kj, kv, mi, mj, mv, p, sol, VV(0));
    for (size_t j = 0; j < n; ++j) VV(0)[j] = rhs[j] - VV(0)[j];
    for (;;) {                                                   /* label 20 */
        ro = or_dnrm2(n, VV(0));
        if (its == 0 && nh < hist_cap) hist[nh++] = ro / ro0;
        if (ro == 0.0) { ierr = -1; break; }
        t = 1.0 / ro;
        for (size_t j = 0; j < n; ++j) VV(0)[j] = VV(0)[j] * t;
        if (its == 0) eps1 = eps * ro;
        rs[0] = ro;
        i = 0;
        for (;;) {                                               /* label 4 */
            i = i + 1; its = its + 1;
            int i1 = i + 1;
            if (prec(a, q, pctx, p, VV(i - 1), rhs)) { ierr = 2; goto done; }
            or_kronspmv(a, b, q, ki, kj, kv, mi, mj, mv, p, rhs, VV(i1 - 1));
            for (int j = 1; j <= i; ++j) {                       /* MGS :277 */
                t = or_ddot(n, VV(j - 1), VV(i1 - 1));
                HH(j - 1, i - 1) = t;
                or_daxpy(n, -t, VV(j - 1), VV(i1 - 1));
            }
            t = or_dnrm2(n, VV(i1 - 1));
            HH(i1 - 1, i - 1) = t;
            if (t != 0.0) {
                t = 1.0 / t;
                for (size_t k = 0; k < n; ++k) VV(i1 - 1)[k] = VV(i1 - 1)[k] * t;
            }
            if (i != 1) {                                        /* :293 */
                for (int k = 2; k <= i; ++k) {
                    int k1 = k - 1;
                    t = HH(k1 - 1, i - 1);
                    HH(k1 - 1, i - 1) = c[k1 - 1] * t + s[k1 - 1] * HH(k - 1, i - 1);
                    HH(k - 1, i - 1) = -s[k1 - 1] * t + c[k1 - 1] * HH(k - 1, i - 1);
                }
            }
            gam = sqrt(HH(i - 1, i - 1) * HH(i - 1, i - 1) + HH(i1 - 1, i - 1) * HH(i1 - 1, i - 1));
            if (gam == 0.0) gam = epsmac;
            c[i - 1] = HH(i - 1, i - 1) / gam;
            s[i - 1] = HH(i1 - 1, i - 1) / gam;
            rs[i1 - 1] = -s[i - 1] * rs[i - 1];
            rs[i - 1] = c[i - 1] * rs[i - 1];
            HH(i - 1, i - 1) = c[i - 1] * HH(i - 1, i - 1) + s[i - 1] * HH(i1 - 1, i - 1);
            ro = fabs(rs[i1 - 1]);
            if (nh < hist_cap) hist[nh++] = ro / ro0;
            if (!(i < im && ro > eps1)) break;
        }
        rs[i - 1] = rs[i - 1] / HH(i - 1, i - 1);               /* :314 */
        for (int ii = 2; ii <= i; ++ii) {
            int k = i - ii + 1, k1 = k + 1;
            t = rs[k - 1];
            for (int j = k1; j <= i; ++j) t = t - HH(k - 1, j - 1) * rs[j - 1];
            rs[k - 1] = t / HH(k - 1, k - 1);
        }
        t = rs[0];
        for (size_t k = 0; k < n; ++k) rhs[k] = VV(0)[k] * t;
        for (int j = 2; j <= i; ++j) {
            t = rs[j - 1];
            for (size_t k = 0; k < n; ++k) rhs[k] = rhs[k] + t * VV(j - 1)[k];
        }
        if (prec(a, q, pctx, p, rhs, rhs)) { ierr = 2; goto done; }
        for (size_t k = 0; k < n; ++k) sol[k] = sol[k] + rhs[k];
        if (ro <= eps1) { ierr = 0; break; }                     /* :341 */
        if (its >= maxits) { ierr = 1; break; }
        for (int j = 1; j <= i; ++j) {                           /* :344 */
            int jj = i + 1 - j + 1;
            rs[jj - 2] = -s[jj - 2] * rs[jj - 1];
            rs[jj - 1] = c[jj - 2] * rs[jj - 1];
        }
        for (int j = 1; j <= i + 1; ++j) {
            t = rs[j - 1];
            if (j == 1) t = t - 1.0;
            or_daxpy(n, t, VV(j - 1), VV(0));
        }
    }
done:
    *nhist = nh;
    if (its_out) *its_out = its;
    free(hh); free(c); free(s); free(rs); free(vv);
#undef HH
#undef VV
    return ierr;
}

int or_kronpgmres(const double *a, const double *b, int q,
                  const int *ki, const int *kj, const double *kv,
                  const int *mi, const int *mj, const double *mv, int p,
                  const double *alu, const int *jlu, const int *ju,
                  double *rhs, double *sol, int im, double eps, int maxits,
                  double *hist, int hist_cap, int *nhist, int *its_out)
{
    struct or_ilut_ctx c = {alu, jlu, ju};
    return or_kronpgmres_impl(a, b, q, ki, kj, kv, mi, mj, mv, p, or_prec_ilut, &c, rhs, sol,
                              im, eps, maxits, hist, hist_cap, nhist, its_out);
}

/* ------------------------------------------------------------------------ */
/* Exact factor of K for the live-path semantics.  The reference's live       */
/* solver preconditions with the exact MUMPS factor of K (mumps.f90:22-145,   */
/* applied by mpilusol, kronecker_mumps.f90:120-183); ilut with lfil >= n and */
/* droptol = 0 is the same exact factor (ilut.f:2169-2170).  For a banded SPD */
/* K (half-bandwidth w) this is the textbook band Cholesky K = L L^T          */
/* (right-looking, column by column).  Band storage: lb[(i-j) + j*(w+1)] =    */
/* L(i,j) for 0 <= i-j <= w, filled on entry with the lower band of K.        */
/* Returns 0, or j+1 if the leading minor of order j+1 is not positive.       */
/* ------------------------------------------------------------------------ */
int or_bandchol(int n, int w, double *lb)
{
    const size_t ld = (size_t)w + 1;
    for (int j = 0; j < n; ++j) {
        double *cj = lb + (size_t)j * ld;
        if (!(cj[0] > 0.0)) return j + 1;
        const double d = sqrt(cj[0]);
        cj[0] = d;
        const int m = (n - 1 - j) < w ? (n - 1 - j) : w;
        for (int t = 1; t <= m; ++t) cj[t] = cj[t] / d;
        for (int k = 1; k <= m; ++k) {          /* column j+k -= L(j+k,j) L(j+k..,j) */
            double *ck = lb + (size_t)(j + k) * ld;
            const double lkj = cj[k];
            for (int t = k; t <= m; ++t) ck[t - k] = ck[t - k] - cj[t] * lkj;
        }
    }
    return 0;
}

/* (L L^T) x = y with or_bandchol's factor; x may alias y. */
void or_bandsol(int n, int w, const double *lb, const double *y, double *x)
{
    const size_t ld = (size_t)w + 1;
    if (x != y) memcpy(x, y, sizeof(double) * (size_t)n);
    for (int j = 0; j < n; ++j) {                /* L z = y, column oriented */
        const double *cj = lb + (size_t)j * ld;
        x[j] = x[j] / cj[0];
        const int m = (n - 1 - j) < w ? (n - 1 - j) : w;
        for (int t = 1; t <= m; ++t) x[j + t] = x[j + t] - cj[t] * x[j];
    }
    for (int j = n - 1; j >= 0; --j) {           /* L^T x = z, dot oriented */
        const double *cj = lb + (size_t)j * ld;
        const int m = (n - 1 - j) < w ? (n - 1 - j) : w;
        double s = x[j];
        for (int t = 1; t <= m; ++t) s = s - cj[t] * x[j + t];
        x[j] = s / cj[0];
    }
}

/* (A (x) K)^-1 x with the band factor, in kronlusol's order (kronecker.f90:130-178):
   w_j = K^-1 x_j, then y_i = sum_j inva(i,j) w_j.  x and y may alias. */
struct or_band_ctx { const double *lb; int w; };
int or_kronbandsol(const double *a, int q, const double *lb, int w, int p, const double *x,
                   double *y)
{
    double inva[64];
    if (or_inverse(q, a, inva)) return 1;
    size_t n = (size_t)p * q;
    double *wv = (double *)calloc(n, sizeof(double));
    for (int i = 0; i < q; ++i) or_bandsol(p, w, lb, x + (size_t)i * p, wv + (size_t)i * p);
    for (size_t k = 0; k < n; ++k) y[k] = 0.0;
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) {
            double s = inva[i + j * q];
            double *yi = y + (size_t)i * p;
            const double *wj = wv + (size_t)j * p;
            for (int k = 0; k < p; ++k) yi[k] = yi[k] + s * wj[k];
        }
    free(wv);
    return 0;
}
static int or_prec_band(const double *a, int q, const void *ctx, int p, const double *x,
                        double *y)
{
    const struct or_band_ctx *c = (const struct or_band_ctx *)ctx;
    return or_kronbandsol(a, q, c->lb, c->w, p, x, y);
}

/* kronpgmres (kronecker.f90:181-365) with the exact band factor in place of the
   ILUT factors: the live path's mpipgmres semantics with K^-1 exact. */
int or_kronpgmres_band(const double *a, const double *b, int q,
                       const int *ki, const int *kj, const double *kv,
                       const int *mi, const int *mj, const double *mv, int p,
                       const double *lb, int w,
                       double *rhs, double *sol, int im, double eps, int maxits,
                       double *hist, int hist_cap, int *nhist, int *its_out)
{
    struct or_band_ctx c = {lb, w};
    return or_kronpgmres_impl(a, b, q, ki, kj, kv, mi, mj, mv, p, or_prec_band, &c, rhs, sol,
                              im, eps, maxits, hist, hist_cap, nhist, its_out);
}

/* kronpgmres (kronecker.f90:181-365) with the spatial solve K^-1 supplied by the
   caller: solve(ctx, p, q, x, w) writes w_j = K^-1 x_j for the q temporal blocks of x
   (any exact factorization of K: the live path's MUMPS role); the temporal mix
   y_i = sum_j inva(i,j) w_j is kronlusol's (kronecker.f90:171-176). */
typedef int (*or_spatial_fn)(void *ctx, int p, int q, const double *x, double *w);
struct or_cb_ctx { or_spatial_fn solve; void *ctx; };
static int or_prec_cb(const double *a, int q, const void *ctx, int p, const double *x,
                      double *y)
{
    const struct or_cb_ctx *c = (const struct or_cb_ctx *)ctx;
    double inva[64];
    if (or_inverse(q, a, inva)) return 1;
    size_t n = (size_t)p * q;
    double *wv = (double *)calloc(n, sizeof(double));
    if (c->solve(c->ctx, p, q, x, wv)) { free(wv); return 1; }
    for (size_t k = 0; k < n; ++k) y[k] = 0.0;
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) {
            double s = inva[i + j * q];
            double *yi = y + (size_t)i * p;
            const double *wj = wv + (size_t)j * p;
            for (int k = 0; k < p; ++k) yi[k] = yi[k] + s * wj[k];
        }
    free(wv);
    return 0;
}
int or_kronpgmres_cb(const double *a, const double *b, int q,
                     const int *ki, const int *kj, const double *kv,
                     const int *mi, const int *mj, const double *mv, int p,
                     or_spatial_fn solve, void *sctx,
                     double *rhs, double *sol, int im, double eps, int maxits,
                     double *hist, int hist_cap, int *nhist, int *its_out)
{
    struct or_cb_ctx c = {solve, sctx};
    return or_kronpgmres_impl(a, b, q, ki, kj, kv, mi, mj, mv, p, or_prec_cb, &c, rhs, sol,
                              im, eps, maxits, hist, hist_cap, nhist, its_out);
}

/* ------------------------------------------------------------------------ */
/* stcoeff_m: closed-form 6x6 temporal matrices of one slab.                  */
/* stcoeff.f90:62-356.  Outputs column-major 6x6.  Default-real literals of   */
/* the reference (2./3., 1./6., ...) are rounded to FP32 and then widened,    */
/* exactly as the Fortran compiler evaluates them (F32 below).               */
/* ------------------------------------------------------------------------ */
#define F32(x) ((double)(float)(x))
void or_stcoeff_m(double t1, double dt, double w, double *k, double *m,
                  double *ktn, double *mtn, double *ktn1, double *mtn1)
{
#define K(i, j) k[((i) - 1) + ((j) - 1) * 6]
#define M(i, j) m[((i) - 1) + ((j) - 1) * 6]
#define KTN(i, j) ktn[((i) - 1) + ((j) - 1) * 6]
#define MTN(i, j) mtn[((i) - 1) + ((j) - 1) * 6]
#define KTN1(i, j) ktn1[((i) - 1) + ((j) - 1) * 6]
#define MTN1(i, j) mtn1[((i) - 1) + ((j) - 1) * 6]
    const double dt2 = dt * dt, dt3 = dt2 * dt, dt4 = dt2 * dt2;
    const double w2 = w * w, w3 = w2 * w, w4 = w2 * w2;
    const double S1 = sin(t1 * w), C1 = cos(t1 * w);
    const double S2 = sin((dt + t1) * w), C2 = cos((dt + t1) * w);
    const double Sh = sin((dt / 2. + t1) * w);
    (void)C1;
    K(1, 1) = -0.5;
    K(1, 2) = -F32(2.0f / 3.0f);
    K(1, 3) = F32(1.0f / 6.0f);
    K(1, 4) = (-6. * dt * w * (-12. + dt2 * w2) * C1 + 24. * dt * w * C2
               + (96. - 26. * dt2 * w2 + dt4 * w4) * S1 + 2. * (-48. + dt2 * w2) * S2) / (2. * dt4 * w4);
    K(1, 5) = (2. * (dt * w * (-84. * C1 - 60. * C2 + dt * w * (18. * S1 + dt2 * w2 * Sh - 6. * S2))
                     - 144. * (S1 - S2))) / (3. * dt4 * w4);
    K(1, 6) = (120. * dt * w * C1 - 6. * dt * w * (-28. + dt2 * w2) * C2 - 18. * (-16. + dt2 * w2) * S1
               - (288. - 42. * dt2 * w2 + dt4 * w4) * S2) / (6. * dt4 * w4);
    K(2, 1) = F32(2.0f / 3.0f);
    K(2, 2) = 0.0;
    K(2, 3) = -F32(2.0f / 3.0f);
    K(2, 4) = (-2. * (-6. * dt * w * (-16. + dt2 * w2) * C1 + 48. * dt * w * C2
                      + (-24. + dt2 * w2) * ((-6. + dt2 * w2) * S1 + 6. * S2))) / (3. * dt4 * w4);
    K(2, 5) = (16. * (6. * dt * w * C1 + 6. * dt * w * C2 - (-12. + dt2 * w2) * (S1 - S2))) / (dt4 * w4);
    K(2, 6) = (2. * (-48. * dt * w * C1 + 6. * dt * w * (-16. + dt2 * w2) * C2
                     + (-24. + dt2 * w2) * (6. * S1 + (-6. + dt2 * w2) * S2))) / (3. * dt4 * w4);
    K(3, 1) = -F32(1.0f / 6.0f);
    K(3, 2) = F32(2.0f / 3.0f);
    K(3, 3) = 1. / 2.;
    K(3, 4) = (-6. * dt * w * (-28. + dt2 * w2) * C1 + 120. * dt * w * C2
               + (288. - 42. * dt2 * w2 + dt4 * w4) * S1 + 18. * (-16. + dt2 * w2) * S2) / (6. * dt4 * w4);
    K(3, 5) = (-2. * (144. * S1 + dt * w * (60. * C1 + 84. * C2 - 6. * dt * w * S1 + dt3 * w3 * Sh)
                      + 18. * (-8. + dt2 * w2) * S2)) / (3. * dt4 * w4);
    K(3, 6) = (96. * (S1 - S2) + dt * w * (24. * C1 + (72. - 6. * dt2 * w2) * C2
                                           + dt * w * (-2. * S1 + (26. - dt2 * w2) * S2))) / (2. * dt4 * w4);
    K(4, 1) = -K(1, 4); K(4, 2) = -K(2, 4); K(4, 3) = -K(3, 4); K(4, 4) = 0.;
    {
        const double S2t = sin(2. * t1 * w), S2e = sin(2. * (dt + t1) * w);
        K(4, 5) = (576. - 72. * dt2 * w2 - 4. * dt4 * w4 + 6. * (-96. + 13. * dt2 * w2) * C1 * C1
                   + 576. * S1 * S1 - 78. * dt2 * w2 * S1 * S1 + 339. * dt * w * S2t
                   - 6. * dt3 * w3 * S2t + 96. * dt * w * C2 * (5. * S1 - 4. * Sh)
                   + 48 * dt * w * (-16 + dt2 * w2) * C1 * Sh - 1152. * S1 * Sh
                   + 240. * dt2 * w2 * S1 * Sh - 8. * dt4 * w4 * S1 * Sh - 1152. * S1 * S2
                   + 48. * dt2 * w2 * S1 * S2 + 1152. * Sh * S2 - 48. * dt2 * w2 * Sh * S2
                   - 3. * dt * w * S2e) / (12. * dt4 * w4);
        const double Cd = cos(dt * w), Sd = sin(dt * w);
        const double C2t = cos(2. * t1 * w), C2e = cos(2. * (dt + t1) * w), Cm = cos((dt + 2. * t1) * w);
        K(4, 6) = (-1152. + 72. * dt2 * w2 + 2. * dt4 * w4 + 2. * (576. - 84. * dt2 * w2 + dt4 * w4) * Cd
                   - 6. * (-96. + 7. * dt2 * w2) * C2t + 576. * C2e - 42. * dt2 * w2 * C2e
                   - 1152. * Cm + 168. * dt2 * w2 * Cm - 2. * dt4 * w4 * Cm
                   + 672. * dt * w * Sd - 24. * dt3 * w3 * Sd - 246. * dt * w * S2t
                   + 3. * dt3 * w3 * S2t + 246. * dt * w * S2e - 3. * dt3 * w3 * S2e) / (24. * dt4 * w4);
        K(5, 1) = -K(1, 5); K(5, 2) = -K(2, 5); K(5, 3) = -K(3, 5); K(5, 4) = -K(4, 5); K(5, 5) = 0.;
        K(5, 6) = -(-576. + 72. * dt2 * w2 + 4. * dt4 * w4 + (576. - 78. * dt2 * w2) * C2 * C2
                    - 3. * dt * w * S2t - 384. * dt * w * C1 * Sh + 48. * dt * w * (-16. + dt2 * w2) * C2 * Sh
                    - 1152. * S1 * Sh + 48. * dt2 * w2 * S1 * Sh + 480. * dt * w * C1 * S2
                    + 1152. * S1 * S2 - 48. * dt2 * w2 * S1 * S2 + 1152. * Sh * S2
                    - 240. * dt2 * w2 * Sh * S2 + 8. * dt4 * w4 * Sh * S2 - 576. * S2 * S2
                    + 78. * dt2 * w2 * S2 * S2 + 339. * dt * w * S2e - 6. * dt3 * w3 * S2e) / (12. * dt4 * w4);
        K(6, 1) = -K(1, 6); K(6, 2) = -K(2, 6); K(6, 3) = -K(3, 6); K(6, 4) = -K(4, 6); K(6, 5) = -K(5, 6);
        K(6, 6) = 0.;
        /* mass-type slab matrix, stcoeff.f90:179-258 */
        M(1, 1) = -4. / dt2; M(1, 2) = 8. / dt2; M(1, 3) = -4. / dt2;
        M(1, 4) = (3. * dt * w * C1 - S1 + S2) / dt2;
        M(1, 5) = (4. * (3. * S1 - 2. * Sh - S2)) / dt2;
        M(1, 6) = (dt * w * C2 - 3. * S1 + 3. * S2) / dt2;
        M(2, 1) = 0.; M(2, 2) = 0.; M(2, 3) = 0.;
        M(2, 4) = (-4. * (dt * w * C1 - S1 + S2)) / dt2;
        M(2, 5) = (16. * (-S1 + S2)) / dt2;
        M(2, 6) = (-4. * (dt * w * C2 - S1 + S2)) / dt2;
        M(3, 1) = 4. / dt2; M(3, 2) = -8. / dt2; M(3, 3) = 4. / dt2;
        M(3, 4) = (dt * w * C1 - 3. * S1 + 3. * S2) / dt2;
        M(3, 5) = (4. * (S1 + 2. * Sh - 3. * S2)) / dt2;
        M(3, 6) = (3. * dt * w * C2 - S1 + S2) / dt2;
        M(4, 1) = 0.; M(4, 2) = 0.; M(4, 3) = 0.;
        M(4, 4) = -(-2. + dt2 * w2 + 2. * Cd + (1. + dt2 * w2) * C2t + C2e - 2. * Cm) / (4. * dt2);
        M(4, 5) = (-8. * dt * w + 4. * dt3 * w3 + 3. * S2t
                   + 2. * dt * w * (4. * Cd + 5. * C2t + 2. * C2e - 4. * Cm - 3. * dt * w * S2t)
                   - 3. * S2e) / (4. * dt3 * w);
        M(4, 6) = (-16. * dt * w - 2. * dt3 * w3 + 8. * dt * w * C1 * C1 - 4. * dt * w * C2 * C2
                   - 8. * dt2 * w2 * C2 * S1 + 16. * dt * w * S1 * S1 - 6. * S2t + 3. * dt2 * w2 * S2t
                   + 8. * dt * w * S1 * S2 + 4. * dt * w * S2 * S2 + 6. * S2e + dt2 * w2 * S2e) / (8. * dt3 * w);
        M(5, 1) = 0.; M(5, 2) = 0.; M(5, 3) = 0.;
        M(5, 4) = -(4. * dt3 * w3 + 10. * dt * w * C1 * C1 - 4. * dt * w * C2 * C2 - 10. * dt * w * S1 * S1
                    + 3. * S2t + 2. * dt2 * w2 * S2t - 16. * dt2 * w2 * C1 * Sh + 16. * dt * w * S1 * Sh
                    - 16. * dt * w * Sh * S2 + 4. * dt * w * S2 * S2 - 3. * S2e) / (4. * dt3 * w);
        M(5, 5) = (4. * (C2t - C2e + 2. * cos(((3 * dt) / 2. + 2. * t1) * w)
                         - 2. * cos(((dt + 4. * t1) * w) / 2.))) / dt2;
        M(5, 6) = (4. * dt3 * w3 - 4. * dt * w * C1 * C1 + 10. * dt * w * C2 * C2 + 4. * dt * w * S1 * S1
                   + 3. * S2t + 16. * dt2 * w2 * C2 * Sh - 16. * dt * w * S1 * Sh + 16. * dt * w * Sh * S2
                   - 10. * dt * w * S2 * S2 - 3. * S2e - 2. * dt2 * w2 * S2e) / (4. * dt3 * w);
        M(6, 1) = 0.; M(6, 2) = 0.; M(6, 3) = 0.;
        M(6, 4) = (2. * dt * w + 2. * dt3 * w3 - 4. * dt * w * Cd + 4. * dt * w * C1 * C1
                   + 2. * dt * w * C2t + 4. * dt * w * C2e + 4. * dt * w * Cm + 6. * S2t
                   + dt2 * w2 * S2t - 8. * dt2 * w2 * C1 * S2 - 6. * S2e + 3. * dt2 * w2 * S2e) / (8. * dt3 * w);
        M(6, 5) = (8. * dt * w - 4. * dt3 * w3 - 3. * S2t + 3. * S2e
                   - 2. * dt * w * (4. * Cd + 2. * cos(2 * t1 * w) + 5. * C2e - 4. * Cm + 3. * dt * w * S2e))
                  / (4. * dt3 * w);
        M(6, 6) = (-2. + dt2 * w2 + 2. * Cd + C2t + (1. + dt2 * w2) * C2e - 2. * Cm) / (4. * dt2);
    }
    /* jump matrices at t_n (+), stcoeff.f90:260-304 */
    for (int e = 0; e < 36; ++e) ktn[e] = 0.0;
    KTN(1, 1) = 1.;
    MTN(1, 1) = 9. / dt2; MTN(1, 2) = -12. / dt2; MTN(1, 3) = 3. / dt2;
    MTN(1, 4) = (-3. * w * C1) / dt;
    MTN(1, 5) = (12. * (-S1 + Sh)) / dt2;
    MTN(1, 6) = (3. * (S1 - S2)) / dt2;
    MTN(2, 1) = MTN(1, 2); MTN(2, 2) = 16. / dt2; MTN(2, 3) = -4. / dt2;
    MTN(2, 4) = (4. * w * C1) / dt;
    MTN(2, 5) = (16. * (S1 - Sh)) / dt2;
    MTN(2, 6) = (4. * (-S1 + S2)) / dt2;
    MTN(3, 1) = MTN(1, 3); MTN(3, 2) = MTN(2, 3); MTN(3, 3) = 1. / dt2;
    MTN(3, 4) = -((w * C1) / dt);
    MTN(3, 5) = (4. * (-S1 + Sh)) / dt2;
    MTN(3, 6) = (S1 - S2) / dt2;
    MTN(4, 1) = MTN(1, 4); MTN(4, 2) = MTN(2, 4); MTN(4, 3) = MTN(3, 4);
    MTN(4, 4) = w2 * C1 * C1;
    MTN(4, 5) = (4. * w * C1 * (S1 - Sh)) / dt;
    MTN(4, 6) = (w * C1 * (-S1 + S2)) / dt;
    MTN(5, 1) = MTN(1, 5); MTN(5, 2) = MTN(2, 5); MTN(5, 3) = MTN(3, 5); MTN(5, 4) = MTN(4, 5);
    MTN(5, 5) = (16. * (S1 - Sh) * (S1 - Sh)) / dt2;
    MTN(5, 6) = (4. * (S1 - Sh) * (-S1 + S2)) / dt2;
    MTN(6, 1) = MTN(1, 6); MTN(6, 2) = MTN(2, 6); MTN(6, 3) = MTN(3, 6); MTN(6, 4) = MTN(4, 6);
    MTN(6, 5) = MTN(5, 6);
    MTN(6, 6) = (S1 - S2) * (S1 - S2) / dt2;
    /* jump matrices at t_n (-), stcoeff.f90:306-354 */
    for (int e = 0; e < 36; ++e) ktn1[e] = 0.0;
    KTN1(1, 3) = 1.;
    MTN1(1, 1) = -3. / dt2; MTN1(1, 2) = 12. / dt2; MTN1(1, 3) = -9 / dt2;
    MTN1(1, 4) = (-3. * (-S1 + S2)) / dt2;
    MTN1(1, 5) = (-12. * (Sh - S2)) / dt2;
    MTN1(1, 6) = (-3. * w * C2) / dt;
    MTN1(2, 1) = 4. / dt2; MTN1(2, 2) = -16. / dt2; MTN1(2, 3) = 12. / dt2;
    MTN1(2, 4) = (4. * (-S1 + S2)) / dt2;
    MTN1(2, 5) = (16. * (Sh - S2)) / dt2;
    MTN1(2, 6) = (4. * w * C2) / dt;
    MTN1(3, 1) = -1. / dt2; MTN1(3, 2) = 4. / dt2; MTN1(3, 3) = -3. / dt2;
    MTN1(3, 4) = -((-S1 + S2) / dt2);
    MTN1(3, 5) = (-4. * (Sh - S2)) / dt2;
    MTN1(3, 6) = -((w * C2) / dt);
    MTN1(4, 1) = (w * C1) / dt; MTN1(4, 2) = (-4. * w * C1) / dt; MTN1(4, 3) = (3. * w * C1) / dt;
    MTN1(4, 4) = (w * C1 * (-S1 + S2)) / dt;
    MTN1(4, 5) = (4. * w * C1 * (Sh - S2)) / dt;
    MTN1(4, 6) = w2 * C1 * C2;
    MTN1(5, 1) = (4. * (S1 - Sh)) / dt2;
    MTN1(5, 2) = (-16. * (S1 - Sh)) / dt2;
    MTN1(5, 3) = (12. * (S1 - Sh)) / dt2;
    MTN1(5, 4) = (4. * (S1 - Sh) * (-S1 + S2)) / dt2;
    MTN1(5, 5) = (16. * (S1 - Sh) * (Sh - S2)) / dt2;
    MTN1(5, 6) = (4. * w * C2 * (S1 - Sh)) / dt;
    MTN1(6, 1) = (-S1 + S2) / dt2;
    MTN1(6, 2) = (-4. * (-S1 + S2)) / dt2;
    MTN1(6, 3) = (3. * (-S1 + S2)) / dt2;
    MTN1(6, 4) = (-S1 + S2) * (-S1 + S2) / dt2;
    MTN1(6, 5) = (4. * (Sh - S2) * (-S1 + S2)) / dt2;
    MTN1(6, 6) = (w * C2 * (-S1 + S2)) / dt;
    (void)w3;
#undef K
#undef M
#undef KTN
#undef MTN
#undef KTN1
#undef MTN1
}

/* stcoeff_t: 3x3 temporal matrices of the thermal slab.  stcoeff.f90:1-60.   */
void or_stcoeff_t(double t1, double dt, double *k, double *m, double *ktn,
                  double *mtn, double *ktn1, double *mtn1)
{
    (void)t1;
#define A3(x, i, j) x[((i) - 1) + ((j) - 1) * 3]
    A3(k, 1, 1) = F32(2.0f / 15.0f) * dt; A3(k, 1, 2) = F32(1.0f / 15.0f) * dt; A3(k, 1, 3) = -F32(1.0f / 30.0f) * dt;
    A3(k, 2, 1) = F32(1.0f / 15.0f) * dt; A3(k, 2, 2) = F32(8.0f / 15.0f) * dt; A3(k, 2, 3) = F32(1.0f / 15.0f) * dt;
    A3(k, 3, 1) = -F32(1.0f / 30.0f) * dt; A3(k, 3, 2) = F32(1.0f / 15.0f) * dt; A3(k, 3, 3) = F32(2.0f / 15.0f) * dt;
    A3(m, 1, 1) = -.5; A3(m, 1, 2) = F32(2.0f / 3.0f); A3(m, 1, 3) = -F32(1.0f / 6.0f);
    A3(m, 2, 1) = -F32(2.0f / 3.0f); A3(m, 2, 2) = 0.; A3(m, 2, 3) = F32(2.0f / 3.0f);
    A3(m, 3, 1) = F32(1.0f / 6.0f); A3(m, 3, 2) = -F32(2.0f / 3.0f); A3(m, 3, 3) = .5;
    for (int e = 0; e < 9; ++e) { ktn[e] = 0.; mtn[e] = 0.; ktn1[e] = 0.; mtn1[e] = 0.; }
    A3(mtn, 1, 1) = 1.;
    A3(mtn1, 1, 3) = 1.;
#undef A3
}

/* getfextc_m: temporal load coefficients.  getfextc.f90:22-58.              */
void or_getfextc_m(double t1, double w, double dt, double *fextc, double *fint)
{
    const double dt2 = dt * dt, w2 = w * w;
    const double S1 = sin(t1 * w), C1 = cos(t1 * w);
    const double S2 = sin((dt + t1) * w), C2 = cos((dt + t1) * w);
    fextc[0] = -((3. * dt * w * C1 + dt * w * C2 + 4. * S1 - 4. * S2) / (dt2 * w2));
    fextc[1] = (4. * (dt * w * C1 + dt * w * C2 + 2. * S1 - 2. * S2)) / (dt2 * w2);
    fextc[2] = -((dt * w * C1 + 3. * dt * w * C2 + 4 * S1 - 4. * S2) / (dt2 * w2));
    fextc[3] = (16. - 4. * dt2 * w2 + 2. * (-7. + dt2 * w2) * cos(2. * t1 * w) - 2. * cos(2. * (dt + t1) * w)
                + 8. * dt * w * C2 * S1 + 9. * dt * w * sin(2. * t1 * w) - 32. * S1 * S2
                - dt * w * sin(2. * (dt + t1) * w)) / (8. * dt2 * w2);
    fextc[4] = (-cos(2. * t1 * w) + cos(2. * (dt + t1) * w) + 8. * cos(((dt + 4. * t1) * w) / 2.)
                - 8. * cos((3. * dt * w) / 2. + 2. * t1 * w)
                + dt * w * (sin(2. * t1 * w) + sin(2. * (dt + t1) * w)
                            - 4. * (sin(((dt + 4. * t1) * w) / 2.) + sin((3. * dt * w) / 2. + 2. * t1 * w))))
               / (2. * dt2 * w2);
    fextc[5] = (-16. + 4. * dt2 * w2 + 2. * cos(2. * t1 * w) - 2. * (-7. + dt2 * w2) * cos(2. * (dt + t1) * w)
                - dt * w * sin(2. * t1 * w) + 8. * (dt * w * C1 + 4. * S1) * S2
                + 9. * dt * w * sin(2. * (dt + t1) * w)) / (8. * dt2 * w2);
    fint[0] = -1.; fint[1] = 0.; fint[2] = 1.; fint[3] = 0.; fint[4] = 0.;
    fint[5] = -S1 + sin(w * (t1 + dt));
}

/* ------------------------------------------------------------------------ */
/* Temporal interpolation table Nt(6, nip) of one slab.                       */
/* kokkos_damage.f90:36-47: t_i = t_1 + (i-1) dt/nip (slab end excluded),     */
/* quadratic Lagrange N1..N3 and enrichments N_{3+k} = N_k (sin wt - sin wt_k)*/
/* Output Nt[i*6 + k] (time-point major, as gpu_Nt).                          */
/* ------------------------------------------------------------------------ */
void or_damage_nt(const double *tcoorde, double w, int nip, double *Nt)
{
    double dt = tcoorde[2] - tcoorde[0];
    for (int i = 1; i <= nip; ++i) {
        double t = tcoorde[0] + (i - 1) * dt / nip;
        double *o = Nt + (size_t)(i - 1) * 6;
        o[0] = 2.0 * (tcoorde[1] - t) * (tcoorde[2] - t) / (dt * dt);
        o[1] = -4.0 * (tcoorde[0] - t) * (tcoorde[2] - t) / (dt * dt);
        o[2] = 2.0 * (tcoorde[0] - t) * (tcoorde[1] - t) / (dt * dt);
        o[3] = o[0] * (sin(w * t) - sin(w * tcoorde[0]));
        o[4] = o[1] * (sin(w * t) - sin(w * tcoorde[1]));
        o[5] = o[2] * (sin(w * t) - sin(w * tcoorde[2]));
    }
}

/* ------------------------------------------------------------------------ */
/* Two-scale fatigue damage functor, one Gauss point over nip time points.    */
/* kokkos_kernel.h:44-344 with the constants of damage.h:8-17.  Same          */
/* arguments and flattened layouts as kokkos_host_ (kokkos_host.cpp:37-41):   */
/* nodes[8*g + a] (1-based), Nx/Ny/Nz[8*g + a], Nt[jj*6 + k], sol temporal-   */
/* major, per-GP state vectors [6*g + c] and scalars [g].                     */
/* ------------------------------------------------------------------------ */
/* The constants are variables so that a caller may set another material (the
   reference's are the defaults, damage.h:9-17): sig_f, sig_u, eps_D, Cy, h, S, s. */
static double g_dp[7] = {180.0, 577.0, 0.08, 1740.0, 0.2, 0.5, 0.5};
void or_damage_set_params(const double *params)
{
    static const double ref[7] = {180.0, 577.0, 0.08, 1740.0, 0.2, 0.5, 0.5};
    for (int i = 0; i < 7; ++i) g_dp[i] = params ? params[i] : ref[i];
}
#define D_SIGF (g_dp[0])
#define D_SIGU (g_dp[1])
#define D_EPSD (g_dp[2])
#define D_CY (g_dp[3])
#define D_H (g_dp[4])
#define D_S (g_dp[5])
#define D_SEXP (g_dp[6])
#define D_WD ((D_SIGU - D_SIGF) * D_EPSD)

/* One return-map step of the functor for a given strain (kokkos_kernel.h:85-341).
 * State: eo/ep/se/sb[6] = eps_old, eps_pl, sig_eff, sig_back; sc[5] = p, D, ws,
 * del_po, sig_d.  Same operation order as the reference. */
void or_damage_step(const double *strain, double ym, double nu, double *eo, double *ep,
                    double *se, double *sb, double *sc)
{
    double *p = &sc[0], *d = &sc[1], *ws = &sc[2], *del_po = &sc[3], *sig_d = &sc[4];
    double a = (1.0 + nu) / (3.0 * (1.0 - nu));
    double b = 2.0 * (4.0 - 5.0 * nu) / (15.0 * (1.0 - nu));
    double G = ym / (2.0 * (1.0 + nu));
    double Kb = ym / (3.0 * (1.0 - 2.0 * nu));
    double coef = ym / ((1. + nu) * (1. - 2. * nu));
    double del_p = 0.0;
    double dmat[6][6] = {{0.0}};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) dmat[i][j] = coef * nu;
    for (int i = 0; i < 3; ++i) {
        dmat[i][i] = coef * (1.0 - nu);
        dmat[i + 3][i + 3] = coef * (1.0 - 2.0 * nu) / 2.0;
    }
    double dd = *d;
    double c1 = 1.0 / (1.0 - b * dd);
    double c2 = (a - b) * dd / ((1.0 - b * dd) * (1.0 - a * dd));
    double c3 = b * (1.0 - dd) / (1.0 - b * dd);
    double tel[6], tee[6];
    for (int i = 0; i < 6; ++i) tel[i] = c1 * strain[i] + c3 * ep[i];
    double tr = 0.0;
    for (int i = 0; i < 3; ++i) tr += strain[i];
    tr = c2 * tr / 3.0;
    for (int i = 0; i < 3; ++i) tel[i] += tr;
    for (int i = 0; i < 6; ++i) tee[i] = tel[i] - ep[i];
    double tse[6] = {0.0};
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) tse[i] += dmat[i][j] * tee[j];
    double trs = 0.0;
    for (int i = 0; i < 3; ++i) trs += tse[i];
    trs = trs / 3.0;
    double tdev[6] = {tse[0] - trs, tse[1] - trs, tse[2] - trs, tse[3], tse[4], tse[5]};
    double tsig[6];
    for (int i = 0; i < 6; ++i) tsig[i] = tdev[i] - sb[i];
    double teq = 0.0;
    for (int i = 0; i < 3; ++i) teq += 1.5 * tsig[i] * tsig[i];
    for (int i = 3; i < 6; ++i) teq += 3.0 * tsig[i] * tsig[i];
    teq = sqrt(teq);
    if (teq <= D_SIGF) {                                        /* elastic :163 */
        for (int i = 0; i < 6; ++i) { se[i] = tse[i]; eo[i] = strain[i]; }
        *sig_d = teq;
        return;
    }
    /* plastic corrector :175 */
    double dstrain[6], dstress[6] = {0.0};
    for (int i = 0; i < 6; ++i) dstrain[i] = strain[i] - eo[i];
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) dstress[i] += dmat[i][j] * dstrain[j];
        dstress[i] = dstress[i] / (1.0 - b * dd);
    }
    double c4 = Kb * (a - b) * dd / ((1.0 - b * dd) * (1.0 - a * dd));
    double trd = 0.0;
    for (int i = 0; i < 3; ++i) trd += dstrain[i];
    for (int i = 0; i < 3; ++i) dstress[i] += c4 * trd;
    double Qs[6];
    for (int i = 0; i < 6; ++i) Qs[i] = sb[i] - se[i] - dstress[i];
    double gamma = D_CY * (1.0 - dd) + 3.0 * G * (1.0 - b) / (1.0 - b * dd);
    double trq = 0.0;
    for (int i = 0; i < 3; ++i) trq += Qs[i];
    trq = trq / 3.0;
    double dq[6] = {Qs[0] - trq, Qs[1] - trq, Qs[2] - trq, Qs[3], Qs[4], Qs[5]};
    double qeq = 0.0;
    for (int i = 0; i < 3; ++i) qeq += 1.5 * dq[i] * dq[i];
    for (int i = 3; i < 6; ++i) qeq += 3.0 * dq[i] * dq[i];
    qeq = sqrt(qeq);
    double tr_sigs = -trq;
    del_p = (qeq - D_SIGF) / gamma;
    *p += del_p;
    double c5 = 1.0 + gamma * del_p / D_SIGF;
    double ds[6];
    for (int i = 0; i < 6; ++i) ds[i] = -dq[i] / c5;
    double sigs[6] = {ds[0] + tr_sigs, ds[1] + tr_sigs, ds[2] + tr_sigs, ds[3], ds[4], ds[5]};
    double dpl[6];
    for (int i = 0; i < 6; ++i) {
        dpl[i] = 3.0 * ds[i] * del_p / (2.0 * D_SIGF);
        ep[i] += dpl[i];
    }
    for (int i = 0; i < 6; ++i) sb[i] += 2.0 * D_CY * (1.0 - dd) * dpl[i] / 3.0;
    for (int i = 0; i < 6; ++i) se[i] = sigs[i] + sb[i];
    double tse2 = 0.0;
    for (int i = 0; i < 3; ++i) tse2 += se[i];
    tse2 = tse2 / 3.0;
    double dse[6];
    for (int i = 0; i < 3; ++i) dse[i] = se[i] - tse2;
    for (int i = 3; i < 6; ++i) dse[i] = se[i];
    double seq = 0.0;
    for (int i = 0; i < 3; ++i) seq += 1.5 * dse[i] * dse[i];
    for (int i = 3; i < 6; ++i) seq += 3.0 * dse[i] * dse[i];
    seq = sqrt(seq);
    double sig_new = fabs(seq - D_SIGF);
    double sig_old = fabs(*sig_d - D_SIGF);
    *ws += (sig_new * del_p + sig_old * *del_po) / 2.0;
    if (*ws >= D_WD) {                                          /* :293 */
        double sp[3] = {0.0}, sn[3] = {0.0};
        for (int i = 0; i < 3; ++i) {
            if (se[i] < 0.0) sn[i] = se[i];
            else sp[i] = se[i];
        }
        double hr = (1.0 - dd) / (1.0 - D_H * dd);
        double Yp = (1.0 + nu) / (2.0 * ym) *
                    (sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2] +
                     D_H * hr * hr * (sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2]));
        double trt = 0.0;
        for (int i = 0; i < 3; ++i) trt += se[i];
        double Yn;
        if (trt >= 0.0) Yn = nu / (2.0 * ym) * trt * trt;
        else Yn = nu / (2.0 * ym) * D_H * hr * hr * trt * trt;
        double Y = Yp - Yn;
        /* (Y/S)^s with s = 0.5 (damage.h:13) evaluated correctly rounded:
         * libm pow differs from the exact power in the last bit (glibc
         * 0.52 ulp) and the return map amplifies last-bit differences at
         * load reversals, so the platform-independent value is used. */
        *d += (D_SEXP == 0.5 ? sqrt(Y / D_S) : pow(Y / D_S, D_SEXP)) * del_p;
    } else {
        *d = 0.0;
    }
    for (int i = 0; i < 6; ++i) eo[i] = strain[i];
    *sig_d = seq;
    *del_po = del_p;
}

/* One Gauss point over its nip time points: displacement interpolation and strain
 * (kokkos_kernel.h:58-84), then the return map. */
static void or_damage_gp(int ii, int nn, int ndofn, int nnse, int nnte, int ncmp, int nip,
                         const int *nodes, const double *Nx, const double *Ny, const double *Nz,
                         const double *Nt, const double *sol, const double *ym, const double *nu,
                         double *eps_old, double *eps_pl, double *sig_eff, double *sig_back,
                         double *p, double *d, double *ws, double *del_po, double *sig_d)
{
    double sc[5] = {p[ii], d[ii], ws[ii], del_po[ii], sig_d[ii]};
    for (int jj = 0; jj < nip; ++jj) {
        double disp[24] = {0.0};
        for (int i = 0; i < nnse; ++i) {
            int ndid = nodes[ii * nnse + i] - 1;
            for (int j = 0; j < ndofn; ++j)
                for (int k = 0; k < 2 * nnte; ++k) {
                    size_t dofid = (size_t)k * nn * ndofn + (size_t)ndid * ndofn + j;
                    disp[i * ndofn + j] += Nt[jj * 2 * nnte + k] * sol[dofid];
                }
        }
        double strain[6] = {0.0};
        for (int i = 0; i < nnse; ++i) {
            double nx = Nx[ii * nnse + i], ny = Ny[ii * nnse + i], nz = Nz[ii * nnse + i];
            strain[0] += nx * disp[i * ndofn + 0];
            strain[1] += ny * disp[i * ndofn + 1];
            strain[2] += nz * disp[i * ndofn + 2];
            strain[3] += ny * disp[i * ndofn + 0] + nx * disp[i * ndofn + 1];
            strain[4] += nz * disp[i * ndofn + 1] + ny * disp[i * ndofn + 2];
            strain[5] += nz * disp[i * ndofn + 0] + nx * disp[i * ndofn + 2];
        }
        or_damage_step(strain, ym[ii], nu[ii], eps_old + ii * ncmp, eps_pl + ii * ncmp,
                       sig_eff + ii * ncmp, sig_back + ii * ncmp, sc);
    }
    p[ii] = sc[0]; d[ii] = sc[1]; ws[ii] = sc[2]; del_po[ii] = sc[3]; sig_d[ii] = sc[4];
}

/* kokkos_host_ (kokkos_host.cpp:37-190): run the functor over all ngpe*ne    */
/* Gauss points.  rank is accepted for signature parity and unused.           */
void or_kokkos_host(int rank, int ne, int nn, int ndofn, int nnse, int nnte, int ngpe,
                    int ncmp, int nip, const int *nodes, const double *Nx, const double *Ny,
                    const double *Nz, const double *Nt, const double *sol, const double *ym,
                    const double *nu, double *eps_old, double *eps_pl, double *sig_eff,
                    double *sig_back, double *p, double *D, double *ws, double *del_po,
                    double *sig_d)
{
    (void)rank;
    for (int ii = 0; ii < ngpe * ne; ++ii)
        or_damage_gp(ii, nn, ndofn, nnse, nnte, ncmp, nip, nodes, Nx, Ny, Nz, Nt, sol, ym, nu,
                     eps_old, eps_pl, sig_eff, sig_back, p, D, ws, del_po, sig_d);
}

/* chkelesta: element e fails iff any of its ngpe Gauss points has D >= Dc;   */
/* its Gauss points and the element are marked failed (0).                    */
/* delelement.f90:109-136.  nefail receives 1-based element ids.             */
int or_chkelesta(int ne, int ngpe, double Dc, const double *D, int *elesta, int *gpsta,
                 int *nefail)
{
    int nef = 0;
    for (int i = 0; i < ne; ++i) {
        int any = 0;
        for (int g = 0; g < ngpe; ++g) if (D[ngpe * i + g] >= Dc) any = 1;
        if (any) {
            for (int g = 0; g < ngpe; ++g) gpsta[ngpe * i + g] = 0;
            elesta[i] = 0;
            nefail[nef++] = i + 1;
        }
    }
    return nef;
}
