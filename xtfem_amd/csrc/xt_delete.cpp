// Element deletion: the sparse edits of K and M that follow a failed element
// (delelement.f90:1-59, 139-247, driver block main3d.f90:632-691), native C++.
//
// * xt_delelement: K := K - sum_e K_e over the failed elements, K_e from the
//   element stiffness routine elekmtx (elesubs.f90:207-251) with one 6x6
//   material matrix for every Gauss point of every failed element (the driver
//   passes the first Gauss point's, main3d.f90:650-652);
// * xt_chkcnsta / xt_chkndsta: surviving-element count per node and the nodes
//   left with none (orphans), in the reference's discovery order;
// * xt_adddiagonal: unit diagonal on every DOF of the orphan nodes;
// * xt_csrapplybc: the 1-0 method on fixed DOFs (spmtx.f90:268-307).
//
// Every matrix edit has the value semantics of the reference's CSparse-style
// addition cs_add (spmtx.f90:701-770): per row, the entries of A then those of
// B are scattered (x = alpha*a, then x += beta*b for a repeated column), the
// row is emitted in increasing column order and exact zeros are dropped.  The
// element matrices accumulate in the order of the failed-element list, so the
// result is bit-identical to the reference's.  The merge covers only the rows
// from the first to the last one a failed element touches (the others, already
// sorted and zero-free, are the result's own: kept in place or moved by the change
// in entry count); the checks and that move are the O(nnz(K)) part.  The
// refactorisation that follows runs on the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "xt_common.h"

#pragma clang fp contract(off)   // products and sums round as written

namespace {

struct Triplet {
    int row, col, seq;   // seq: position in the failed-element order
    double v;
};

// Whether row r of a 1-based CSR has strictly increasing columns and no exact zeros:
// such a row comes out of the scatter of C := 1*A + beta*B unchanged where B has
// nothing in it (1*a == a).
bool clean_row(const int *ia, const int *ja, const double *a, int r) {
    const int k0 = ia[r] - 1, k1 = ia[r + 1] - 1;
    for (int k = k0; k < k1; ++k)
        if (!(a[k] != 0.0 && (k == k0 || ja[k] > ja[k - 1]))) return false;
    return true;
}

// C := alpha*A + beta*B over rows rbeg .. rend-1 of an n-row 1-based CSR A.  B is given
// as per-row runs of (col, value) pairs sorted by row; the rows of C go to (ic, jc, c)
// with ic[0] = 1 at row rbeg (ic has rend - rbeg + 1 entries).
void csr_add_rows(int n, const int *ia, const int *ja, const double *a, double alpha,
                  const std::vector<int> &ib, const std::vector<int> &jb,
                  const std::vector<double> &b, double beta, std::vector<int> &ic,
                  std::vector<int> &jc, std::vector<double> &c, int rbeg, int rend) {
    std::vector<int> mark(n, -1);
    std::vector<double> x(n);
    std::vector<int> cols;
    cols.reserve(128);
    ic.assign(rend - rbeg + 1, 0);
    jc.clear();
    c.clear();
    jc.reserve((size_t)(ia[rend] - ia[rbeg]) + jb.size());
    c.reserve((size_t)(ia[rend] - ia[rbeg]) + jb.size());
    ic[0] = 1;
    for (int r = rbeg; r < rend; ++r) {
        const int k0 = ia[r] - 1, k1 = ia[r + 1] - 1;
        if (ib[r] == ib[r + 1] && alpha == 1.0 && clean_row(ia, ja, a, r)) {
            jc.insert(jc.end(), ja + k0, ja + k1);
            c.insert(c.end(), a + k0, a + k1);
            ic[r - rbeg + 1] = (int)jc.size() + 1;
            continue;
        }
        cols.clear();
        for (int k = k0; k < k1; ++k) {
            const int j = ja[k] - 1;
            if (mark[j] != r) { mark[j] = r; cols.push_back(j); x[j] = alpha * a[k]; }
            else x[j] = x[j] + alpha * a[k];
        }
        for (int k = ib[r]; k < ib[r + 1]; ++k) {
            const int j = jb[k];
            if (mark[j] != r) { mark[j] = r; cols.push_back(j); x[j] = beta * b[k]; }
            else x[j] = x[j] + beta * b[k];
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols)
            if (x[j] != 0.0) { jc.push_back(j + 1); c.push_back(x[j]); }
        ic[r - rbeg + 1] = (int)jc.size() + 1;
    }
}

// C := A + beta*B in A's own arrays (1-based, cap entries of storage), B as for
// csr_add_rows.  Only the rows from B's first to its last nonempty row are merged when
// every other row is clean (they are C's rows unchanged: the leading ones stay where
// they are, the trailing ones move by the change in entry count); otherwise every row
// is.  The result is the same either way.
int csr_add_in_place(int n, int *ia, int *ja, double *a, int *nz, int cap,
                     const std::vector<int> &ib, const std::vector<int> &jb,
                     const std::vector<double> &b, double beta, const char *what) {
    int r0 = 0, r1 = n;
    while (r0 < n && ib[r0] == ib[r0 + 1]) ++r0;
    while (r1 > r0 && ib[r1 - 1] == ib[r1]) --r1;   // B empty: r0 = r1 = n
    bool outside_clean = true;
    for (int r = 0; r < r0 && outside_clean; ++r) outside_clean = clean_row(ia, ja, a, r);
    for (int r = r1; r < n && outside_clean; ++r) outside_clean = clean_row(ia, ja, a, r);
    if (!outside_clean) { r0 = 0; r1 = n; }
    std::vector<int> ic, jc;
    std::vector<double> c;
    csr_add_rows(n, ia, ja, a, 1.0, ib, jb, b, beta, ic, jc, c, r0, r1);
    const long long base = ia[r0] - 1, old_len = ia[r1] - ia[r0];
    const long long delta = (long long)jc.size() - old_len;
    if ((long long)*nz + delta > (long long)cap) {
        xt::set_error("%s: result has %lld entries, storage holds %d", what,
                      (long long)*nz + delta, cap);
        return XT_E_ARG;
    }
    const long long tail = (long long)*nz - (ia[r1] - 1);   // entries of rows r1 ..
    if (delta != 0 && tail > 0) {
        std::memmove(ja + base + (long long)jc.size(), ja + ia[r1] - 1, sizeof(int) * tail);
        std::memmove(a + base + (long long)jc.size(), a + ia[r1] - 1, sizeof(double) * tail);
    }
    std::memcpy(ja + base, jc.data(), sizeof(int) * jc.size());
    std::memcpy(a + base, c.data(), sizeof(double) * c.size());
    for (int r = r0 + 1; r <= r1; ++r) ia[r] = (int)base + ic[r - r0];
    for (int r = r1 + 1; r <= n; ++r) ia[r] += (int)delta;
    *nz = (int)(*nz + delta);
    return XT_OK;
}

// Triplets -> per-row runs (0-based columns), duplicates summed in sequence
// order, exact zeros dropped (tricsr, spmtx.f90:359-407).
void triplets_to_rows(int n, std::vector<Triplet> &t, std::vector<int> &ib, std::vector<int> &jb,
                      std::vector<double> &b) {
    std::stable_sort(t.begin(), t.end(), [](const Triplet &x, const Triplet &y) {
        return x.row != y.row ? x.row < y.row : (x.col != y.col ? x.col < y.col : x.seq < y.seq);
    });
    ib.assign(n + 1, 0);
    jb.clear();
    b.clear();
    size_t i = 0;
    std::vector<int> cnt(n, 0);
    while (i < t.size()) {
        size_t k = i;
        double s = t[i].v;
        while (++k < t.size() && t[k].row == t[i].row && t[k].col == t[i].col) s = s + t[k].v;
        if (s != 0.0) { jb.push_back(t[i].col); b.push_back(s); ++cnt[t[i].row]; }
        i = k;
    }
    for (int r = 0; r < n; ++r) ib[r + 1] = ib[r] + cnt[r];
}

// 1-based CSR with nz entries, monotone row pointers and columns in [1, n].
bool valid_csr(int n, const int *ia, const int *ja, int nz) {
    if (ia[0] != 1 || ia[n] - 1 != nz) return false;
    for (int r = 0; r < n; ++r)
        if (ia[r + 1] < ia[r]) return false;
    for (int k = 0; k < nz; ++k)
        if (ja[k] < 1 || ja[k] > n) return false;
    return true;
}

// 3x3 determinant and cofactor inverse with the term order of procs.f90:300-310,
// 476-478 (column-major j[m + 3p] = J(m+1, p+1)).
double det3(const double *J) {
    auto a = [&](int i, int j) { return J[(i - 1) + 3 * (j - 1)]; };
    return a(1, 1) * a(2, 2) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) + a(1, 3) * a(2, 1) * a(3, 2) -
           a(1, 1) * a(2, 3) * a(3, 2) - a(1, 2) * a(2, 1) * a(3, 3) - a(1, 3) * a(2, 2) * a(3, 1);
}

bool inv3(const double *J, double *G) {
    auto a = [&](int i, int j) { return J[(i - 1) + 3 * (j - 1)]; };
    const double d = det3(J);
    if (!(std::fabs(d) > 10.0 * std::numeric_limits<double>::epsilon())) return false;
    double r[9];
    r[0 + 3 * 0] = a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2);
    r[0 + 3 * 1] = a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3);
    r[0 + 3 * 2] = a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2);
    r[1 + 3 * 0] = a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3);
    r[1 + 3 * 1] = a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1);
    r[1 + 3 * 2] = a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3);
    r[2 + 3 * 0] = a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1);
    r[2 + 3 * 1] = a(1, 2) * a(3, 1) - a(1, 1) * a(3, 2);
    r[2 + 3 * 2] = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
    for (int i = 0; i < 9; ++i) G[i] = r[i] / d;
    return true;
}

// Element stiffness of one hexahedron (elekmtx, elesubs.f90:207-251): for each
// Gauss point l, jac = dns_l X, bmat = (sd * blockdiag(gam, gam, gam)) dnxyz_l,
// ke += (bmat^T (stiff bmat)) * det * w_i * w_j * w_k.  Each matrix product is
// a sequential sum over its inner index.  Column-major operands; 3 DOFs per
// node, nnse nodes per element, ke is (3*nnse)^2.
bool elekmtx(int nnse, int nquads, int ngpe, const double *X /* nnse x 3 */, const double *sd,
             const double *stiff, const double *wqs, const double *dns, const double *dnxyz,
             double *ke) {
    const int nd = 3, nk = nd * nnse, ldd = nd * ngpe, ldx = nd * nd * ngpe;
    std::fill(ke, ke + (size_t)nk * nk, 0.0);
    std::vector<double> bmat(6 * nk), sb(6 * nk);
    int l = 0;
    for (int i = 0; i < nquads; ++i)
        for (int j = 0; j < nquads; ++j)
            for (int k = 0; k < nquads; ++k, ++l) {
                double jac[9], gam[9], gx[81] = {0}, t1[54];
                for (int p = 0; p < nd; ++p)
                    for (int m = 0; m < nd; ++m) {
                        double s = 0.0;
                        for (int a = 0; a < nnse; ++a)
                            s = s + dns[(l * nd + m) + (size_t)ldd * a] * X[a + nnse * p];
                        jac[m + 3 * p] = s;
                    }
                const double det = det3(jac);
                if (!inv3(jac, gam)) return false;
                for (int bl = 0; bl < 3; ++bl)
                    for (int p = 0; p < 3; ++p)
                        for (int m = 0; m < 3; ++m) gx[(3 * bl + m) + 9 * (3 * bl + p)] = gam[m + 3 * p];
                for (int c = 0; c < 9; ++c)
                    for (int r = 0; r < 6; ++r) {
                        double s = 0.0;
                        for (int q = 0; q < 9; ++q) s = s + sd[r + 6 * q] * gx[q + 9 * c];
                        t1[r + 6 * c] = s;
                    }
                for (int c = 0; c < nk; ++c)
                    for (int r = 0; r < 6; ++r) {
                        double s = 0.0;
                        for (int q = 0; q < 9; ++q)
                            s = s + t1[r + 6 * q] * dnxyz[(l * 9 + q) + (size_t)ldx * c];
                        bmat[r + 6 * c] = s;
                    }
                for (int c = 0; c < nk; ++c)
                    for (int r = 0; r < 6; ++r) {
                        double s = 0.0;
                        for (int q = 0; q < 6; ++q) s = s + stiff[r + 6 * q] * bmat[q + 6 * c];
                        sb[r + 6 * c] = s;
                    }
                for (int c = 0; c < nk; ++c)
                    for (int r = 0; r < nk; ++r) {
                        double s = 0.0;
                        for (int q = 0; q < 6; ++q) s = s + bmat[q + 6 * r] * sb[q + 6 * c];
                        double &e = ke[r + (size_t)nk * c];
                        e = e + s * det * wqs[i] * wqs[j] * wqs[k];
                    }
            }
    return true;
}

}  // namespace

extern "C" {

int xt_delelement(int nn, int ne, int ndofn, int nnse, int ngpe, int ncmp, int nquads,
                  const int *xconn, const double *xcoord, const double *dns, const double *dnxyz,
                  const double *sd, const double *stiff, const double *wqs, int nef,
                  const int *nefail, int *ki, int *kj, double *kv, int ndof, int *knz) {
    XT_ARG(nn > 0 && ne >= 0 && nnse > 0 && ndof > 0 && knz && *knz >= 0 && nef >= 0,
           "delelement sizes");
    if (ndofn != 3 || ncmp != 6 || nquads <= 0 || nquads * nquads * nquads != ngpe) {
        xt::set_error("delelement: hexahedra with 3 DOFs per node, 6 strain components and "
                      "nquads^3 Gauss points only (ndofn=%d ncmp=%d nquads=%d ngpe=%d)",
                      ndofn, ncmp, nquads, ngpe);
        return XT_E_UNSUPPORTED;
    }
    if (nef == 0) return XT_OK;
    XT_ARG(xconn && xcoord && dns && dnxyz && sd && stiff && wqs && nefail && ki && kj && kv,
           "delelement arrays");
    XT_ARG(valid_csr(ndof, ki, kj, *knz), "delelement: K is not a 1-based CSR of knz entries");
    const int nk = ndofn * nnse;
    std::vector<Triplet> t;
    t.reserve((size_t)nef * nk * nk);
    std::vector<double> X(nnse * 3), ke((size_t)nk * nk);
    std::vector<int> dofs(nk);
    for (int f = 0; f < nef; ++f) {
        const int e = nefail[f] - 1;
        XT_ARG(e >= 0 && e < ne, "delelement: failed element id out of range");
        for (int a = 0; a < nnse; ++a) {
            const int node = xconn[e + (size_t)ne * a];
            XT_ARG(node >= 1 && node <= nn, "delelement: node id out of range");
            for (int p = 0; p < 3; ++p) {
                X[a + nnse * p] = xcoord[(node - 1) + (size_t)nn * p];
                dofs[a * 3 + p] = (node - 1) * 3 + p;
            }
        }
        XT_ARG(dofs[nk - 1] < ndof, "delelement: DOF beyond ndof");
        if (!elekmtx(nnse, nquads, ngpe, X.data(), sd, stiff, wqs, dns, dnxyz, ke.data())) {
            xt::set_error("delelement: singular Jacobian in element %d (the reference STOPs)", e + 1);
            return XT_E_SINGULAR;
        }
        for (int r = 0; r < nk; ++r)
            for (int c = 0; c < nk; ++c) t.push_back({dofs[r], dofs[c], f, ke[r + (size_t)nk * c]});
    }
    std::vector<int> ib, jb;
    std::vector<double> b;
    triplets_to_rows(ndof, t, ib, jb, b);
    // kj/kv have room for the result by the caller's contract (at most
    // knz + nef*(3*nnse)^2 entries)
    return csr_add_in_place(ndof, ki, kj, kv, knz, *knz + nef * nk * nk, ib, jb, b, -1.0,
                            "delelement");
}

int xt_chkcnsta(int nn, int ne, int nnse, const int *xconn, int nef, const int *nefail,
                int *necnt) {
    XT_ARG(nn > 0 && ne >= 0 && nnse > 0 && nef >= 0, "chkcnsta sizes");
    if (nef == 0) return XT_OK;
    XT_ARG(xconn && nefail && necnt, "chkcnsta arrays");
    for (int f = 0; f < nef; ++f) {
        const int e = nefail[f] - 1;
        XT_ARG(e >= 0 && e < ne, "chkcnsta: failed element id out of range");
        for (int a = 0; a < nnse; ++a) {
            const int node = xconn[e + (size_t)ne * a];
            XT_ARG(node >= 1 && node <= nn, "chkcnsta: node id out of range");
            necnt[node - 1] -= 1;
        }
    }
    return XT_OK;
}

int xt_chkndsta(int nn, int ne, int nnse, const int *xconn, int nef, const int *nefail,
                const int *necnt, int *ndf, int *ndfail) {
    XT_ARG(nn > 0 && ne >= 0 && nnse > 0 && nef >= 0 && ndf && ndfail, "chkndsta arguments");
    *ndf = 0;
    std::fill(ndfail, ndfail + nn, 0);
    if (nef == 0) return XT_OK;
    XT_ARG(xconn && nefail && necnt, "chkndsta arrays");
    std::vector<char> seen(nn, 0);
    int cnt = 0;
    for (int f = 0; f < nef; ++f) {
        const int e = nefail[f] - 1;
        XT_ARG(e >= 0 && e < ne, "chkndsta: failed element id out of range");
        for (int a = 0; a < nnse; ++a) {
            const int node = xconn[e + (size_t)ne * a];
            XT_ARG(node >= 1 && node <= nn, "chkndsta: node id out of range");
            if (necnt[node - 1] == 0 && !seen[node - 1]) {
                seen[node - 1] = 1;
                ndfail[cnt++] = node;
            }
        }
    }
    *ndf = cnt;
    return XT_OK;
}

int xt_adddiagonal(int nnse, int ndofn, int ndf, const int *ndfail, int *ki, int *kj, double *kv,
                   int ndof, int *knz) {
    (void)nnse;
    XT_ARG(ndofn > 0 && ndf >= 0 && ndof > 0 && knz && *knz >= 0, "adddiagonal sizes");
    if (ndf == 0) return XT_OK;
    XT_ARG(ndfail && ki && kj && kv, "adddiagonal arrays");
    XT_ARG(valid_csr(ndof, ki, kj, *knz), "adddiagonal: K is not a 1-based CSR of knz entries");
    std::vector<Triplet> t;
    t.reserve((size_t)ndf * ndofn);
    for (int i = 0; i < ndf; ++i)
        for (int k = 0; k < ndofn; ++k) {
            const int d = (ndfail[i] - 1) * ndofn + k;
            XT_ARG(d >= 0 && d < ndof, "adddiagonal: DOF out of range");
            t.push_back({d, d, (int)t.size(), 1.0});
        }
    std::vector<int> ib, jb;
    std::vector<double> b;
    triplets_to_rows(ndof, t, ib, jb, b);
    // storage: the caller's arrays hold knz + ndf*ndofn entries (the reference
    // grows knz by that much before its addition)
    return csr_add_in_place(ndof, ki, kj, kv, knz, *knz + ndf * ndofn, ib, jb, b, 1.0,
                            "adddiagonal");
}

int xt_csrapplybc(int *ia, int *ja, double *a, int n, int *nz, const int *bc) {
    XT_ARG(n > 0 && nz && *nz >= 0 && ia && ja && a && bc, "csrapplybc arguments");
    XT_ARG(valid_csr(n, ia, ja, *nz), "csrapplybc: not a 1-based CSR of nz entries");
    // compaction in place: the write position never passes the read position
    int out = 0;
    for (int r = 0; r < n; ++r) {
        const int k0 = ia[r] - 1, k1 = ia[r + 1] - 1;
        ia[r] = out + 1;
        for (int k = k0; k < k1; ++k) {
            const int c = ja[k] - 1;
            double v = a[k];
            if (bc[r] != 0 || bc[c] != 0) v = (r == c) ? 1.0 : 0.0;
            if (v != 0.0) { ja[out] = c + 1; a[out] = v; ++out; }
        }
    }
    ia[n] = out + 1;
    *nz = out;
    return XT_OK;
}

int xt_csr_values_in_pattern(int n, const int *pia, const int *pja, const int *ia, const int *ja,
                             const double *a, double *out) {
    XT_ARG(n > 0 && pia && pja && ia && ja && a && out, "csr_values_in_pattern arguments");
    for (int r = 0; r < n; ++r) {   // row-wise merge of two column-sorted rows
        int k = pia[r] - 1;
        const int k1 = pia[r + 1] - 1;
        for (int m = ia[r] - 1; m < ia[r + 1] - 1; ++m) {
            while (k < k1 && pja[k] < ja[m]) out[k++] = 0.0;
            if (k == k1 || pja[k] != ja[m]) {
                xt::set_error("csr_values_in_pattern: row %d column %d is not in the pattern", r + 1,
                          ja[m]);
                return XT_E_ARG;
            }
            out[k++] = a[m];
        }
        while (k < k1) out[k++] = 0.0;
    }
    return XT_OK;
}

}  // extern "C"
