// Host ILUT factorisation (Saad's dual-threshold incomplete LU), native C++.
//
// Same algorithm and output as SPARSKIT ilut (ilut.f:152-495), which the
// reference links for its Kronecker preconditioner (kronecker.f90:130-178):
//   * row norm tnorm = sum|a_ij| / nnz(row) (ilut.f:262-267);
//   * a multiplier of row jrow is dropped if |fact| <= droptol (absolute);
//   * U entries are kept if |w| > droptol*tnorm;
//   * the lfil largest L entries and the lfil-1 largest off-diagonal U entries
//     are kept, selected by a quick-split partition (qsplit, ilut.f:2487-2545)
//     so the stored order is the reference's;
//   * a zero pivot becomes (1e-4 + droptol)*tnorm with 1e-4 rounded to FP32;
//   * output: MSR alu/jlu (diagonal inverted in alu[0..n-1]) and ju, 1-based.
// The factorisation runs once per change of K; the triangular solves that use
// it run on the GPU (xt_prec.hip).
#include <cmath>
#include <vector>

#include "../../include/xtfem.h"

namespace {

// Partition a[0..n) (with ind) so that |a[i]| >= |a[ncut-1]| for i < ncut-1 and
// |a[i]| <= |a[ncut-1]| after it; ncut is a 1-based position as in the reference.
void quick_split(double *a, int *ind, int n, int ncut) {
    int first = 0, last = n - 1;
    const int cut = ncut - 1;
    if (cut < first || cut > last) return;
    for (;;) {
        int mid = first;
        const double key = std::fabs(a[mid]);
        for (int j = first + 1; j <= last; ++j) {
            if (std::fabs(a[j]) > key) {
                ++mid;
                std::swap(a[mid], a[j]);
                std::swap(ind[mid], ind[j]);
            }
        }
        std::swap(a[mid], a[first]);
        std::swap(ind[mid], ind[first]);
        if (mid == cut) return;
        if (mid > cut) last = mid - 1;
        else first = mid + 1;
    }
}

}  // namespace

extern "C" int xt_ilut(int n, const double *a, const int *ja, const int *ia, int lfil,
                       double droptol, double *alu, int *jlu, int *ju, int iwk, double *w,
                       int *jw, int *ierr) {
    if (!ierr || n < 0 || !a || !ja || !ia || !alu || !jlu || !ju || !w || !jw) return XT_E_ARG;
    if (lfil < 0) { *ierr = -4; return XT_OK; }
    // Working storage, 0-based: positions [0, lenl) hold the L part of the row,
    // position n0 = row index holds the diagonal and (row, row+lenu) the U part.
    // jr[col] = working position of col + 1, or 0 if col is not in the row.
    double *wv = w;
    int *wj = jw;          // column of each working entry
    int *jr = jw + n;      // nonzero indicator
    int next = n + 1;      // 0-based position of the next free MSR slot (1-based n+2)
    jlu[0] = n + 2;
    for (int j = 0; j < n; ++j) jr[j] = 0;
    for (int ii = 0; ii < n; ++ii) {
        const int j1 = ia[ii] - 1, j2 = ia[ii + 1] - 1;
        double tnorm = 0.0;
        for (int k = j1; k < j2; ++k) tnorm += std::fabs(a[k]);
        if (tnorm == 0.0) { *ierr = -5; return XT_OK; }
        tnorm = tnorm / (double)(float)(j2 - j1);
        int lenu = 1, lenl = 0;
        wj[ii] = ii;
        wv[ii] = 0.0;
        jr[ii] = ii + 1;
        for (int j = j1; j < j2; ++j) {
            const int k = ja[j] - 1;
            const double t = a[j];
            if (k < ii) {
                wj[lenl] = k; wv[lenl] = t; jr[k] = lenl + 1; ++lenl;
            } else if (k == ii) {
                wv[ii] = t;
            } else {
                const int pos = ii + lenu;
                ++lenu;
                wj[pos] = k; wv[pos] = t; jr[k] = pos + 1;
            }
        }
        // eliminate with previous rows, smallest column first
        int len = 0;
        for (int jj = 0; jj < lenl; ++jj) {
            int jrow = wj[jj], kmin = jj;
            for (int j = jj + 1; j < lenl; ++j)
                if (wj[j] < jrow) { jrow = wj[j]; kmin = j; }
            if (kmin != jj) {
                const int jold = wj[jj];
                wj[jj] = wj[kmin]; wj[kmin] = jold;
                jr[jrow] = jj + 1; jr[jold] = kmin + 1;
                std::swap(wv[jj], wv[kmin]);
            }
            jr[jrow] = 0;
            const double fact = wv[jj] * alu[jrow];
            if (std::fabs(fact) <= droptol) continue;
            for (int k = ju[jrow] - 1; k < jlu[jrow + 1] - 1; ++k) {
                const double s = fact * alu[k];
                const int j = jlu[k] - 1;
                const int jpos = jr[j];
                if (j >= ii) {
                    if (jpos == 0) {
                        if (lenu + 1 > n) { *ierr = -1; return XT_OK; }
                        const int pos = ii + lenu;
                        ++lenu;
                        wj[pos] = j; jr[j] = pos + 1; wv[pos] = -s;
                    } else {
                        wv[jpos - 1] -= s;
                    }
                } else {
                    if (jpos == 0) {
                        if (lenl + 1 > n) { *ierr = -1; return XT_OK; }
                        wj[lenl] = j; jr[j] = lenl + 1; wv[lenl] = -s;
                        ++lenl;
                    } else {
                        wv[jpos - 1] -= s;
                    }
                }
            }
            wv[len] = fact;
            wj[len] = jrow;
            ++len;
        }
        for (int k = 0; k < lenu; ++k) jr[wj[ii + k]] = 0;
        // L part: keep the lfil largest multipliers
        lenl = len;
        len = lenl < lfil ? lenl : lfil;
        quick_split(wv, wj, lenl, len);
        for (int k = 0; k < len; ++k) {
            if (next + 1 > iwk) { *ierr = -2; return XT_OK; }
            alu[next] = wv[k]; jlu[next] = wj[k] + 1; ++next;
        }
        ju[ii] = next + 1;
        // U part: threshold, then keep the lfil-1 largest off-diagonals
        len = 0;
        for (int k = 1; k < lenu; ++k) {
            if (std::fabs(wv[ii + k]) > droptol * tnorm) {
                ++len;
                wv[ii + len] = wv[ii + k];
                wj[ii + len] = wj[ii + k];
            }
        }
        lenu = len + 1;
        len = lenu < lfil ? lenu : lfil;
        quick_split(wv + ii + 1, wj + ii + 1, lenu - 1, len);
        if (len + next + 1 > iwk) { *ierr = -3; return XT_OK; }
        for (int k = ii + 1; k <= ii + len - 1; ++k) {
            jlu[next] = wj[k] + 1; alu[next] = wv[k]; ++next;
        }
        if (wv[ii] == 0.0) wv[ii] = ((double)0.0001f + droptol) * tnorm;
        alu[ii] = 1.0 / wv[ii];
        jlu[ii + 1] = next + 1;
    }
    *ierr = 0;
    return XT_OK;
}
