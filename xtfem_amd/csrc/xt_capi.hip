// Reference-signature entry points: host arrays in and out, work on the GPU.
//
// Each function keeps the argument list and meaning of the Fortran routine it
// replaces (cited per function) so a caller can swap the symbol.  Data are
// staged to the device per call; callers that keep matrices and state resident
// across slabs use the handle API (xt_op_*, xt_prec_*, xt_slab_solve,
// xt_damage_device) instead.
#include <cmath>
#include <vector>

#include <cstdlib>

#include "xt_common.h"

namespace xt {

int elestatus_device(int ne, int ngpe, double Dc, const double *D, int *elesta, int *gpsta,
                     int *failed, hipStream_t s);
}  // namespace xt

using namespace xt;

namespace {

struct OpGuard {
    xt_op op = nullptr;
    ~OpGuard() { if (op) xt_op_destroy(op); }
};
struct PrecGuard {
    xt_prec pr = nullptr;
    ~PrecGuard() { if (pr) xt_prec_destroy(pr); }
};

// column-major q x q (Fortran) -> row-major
void colmajor_to_rowmajor(int q, const double *a, double *out) {
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) out[i * q + j] = a[i + j * q];
}

// Y = alpha (A(x)K + B(x)M) X + beta Y on temporal-major host vectors
int kron_host(const double *a, const double *b, double alpha, double beta, int q,
              const int *ki, const int *kj, const double *kv, const int *mi, const int *mj,
              const double *mv, int p, const double *x, double *y) {
    XT_ARG(a && b && ki && kj && kv && mi && mj && mv && x && y && p > 0 && q > 0 && q <= 8,
           "kronspmv arguments");
    OpGuard g;
    const int *rp[2] = {ki, mi};
    const int *ci[2] = {kj, mj};
    const double *vv[2] = {kv, mv};
    XT_TRY(xt_op_create(p, q, 2, rp, ci, vv, &g.op));
    std::vector<double> tm(2 * q * q);
    colmajor_to_rowmajor(q, a, tm.data());
    colmajor_to_rowmajor(q, b, tm.data() + q * q);
    const size_t n = (size_t)p * q;
    DevBuf dtm, dx, dy, tx, ty;
    XT_TRY(dtm.upload(tm.data(), sizeof(double) * tm.size()));
    XT_TRY(tx.upload(x, sizeof(double) * n));
    XT_TRY(dx.reserve(sizeof(double) * n));
    XT_TRY(dy.reserve(sizeof(double) * n));
    XT_TRY(ty.reserve(sizeof(double) * n));
    XT_TRY(xt_to_interleaved(p, q, 1, tx.as<double>(), dx.as<double>(), nullptr));
    if (beta != 0.0) {
        XT_HIP(hipMemcpy(ty.ptr, y, sizeof(double) * n, hipMemcpyHostToDevice));
        XT_TRY(xt_to_interleaved(p, q, 1, ty.as<double>(), dy.as<double>(), nullptr));
    }
    XT_TRY(xt_op_apply(g.op, 1, dtm.ptr, alpha, beta, dx.ptr, dy.ptr, 0, nullptr));
    XT_TRY(xt_from_interleaved(p, q, 1, dy.as<double>(), ty.as<double>(), nullptr));
    XT_HIP(hipMemcpy(y, ty.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
    return XT_OK;
}

}  // namespace

extern "C" {

int xt_kronspmv(const double *a, const double *b, int q, const int *ki, const int *kj,
                const double *kv, const int *mi, const int *mj, const double *mv, int p,
                const double *x, double *y) {
    return kron_host(a, b, 1.0, 0.0, q, ki, kj, kv, mi, mj, mv, p, x, y);
}

int xt_kronspgemv(const double *a, const double *b, double alpha, double beta, int q,
                  const int *ki, const int *kj, const double *kv, const int *mi, const int *mj,
                  const double *mv, int p, const double *x, double *y) {
    return kron_host(a, b, alpha, beta, q, ki, kj, kv, mi, mj, mv, p, x, y);
}

int xt_amux(int n, const double *x, double *y, const double *a, const int *ja, const int *ia) {
    XT_ARG(n > 0 && x && y && a && ja && ia, "amux arguments");
    OpGuard g;
    const int *rp[1] = {ia};
    const int *ci[1] = {ja};
    const double *vv[1] = {a};
    XT_TRY(xt_op_create(n, 1, 1, rp, ci, vv, &g.op));
    const double one = 1.0;
    DevBuf dt, dx, dy;
    XT_TRY(dt.upload(&one, sizeof(double)));
    XT_TRY(dx.upload(x, sizeof(double) * n));
    XT_TRY(dy.reserve(sizeof(double) * n));
    XT_TRY(xt_op_apply(g.op, 1, dt.ptr, 1.0, 0.0, dx.ptr, dy.ptr, 0, nullptr));
    XT_HIP(hipMemcpy(y, dy.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
    return XT_OK;
}

int xt_lusol(int n, const double *y, double *x, const double *alu, const int *jlu,
             const int *ju) {
    XT_ARG(n > 0 && y && x && alu && jlu && ju, "lusol arguments");
    PrecGuard g;
    XT_TRY(xt_prec_create(n, alu, jlu, ju, &g.pr));
    DevBuf dy, dx;
    XT_TRY(dy.upload(y, sizeof(double) * n));
    XT_TRY(dx.reserve(sizeof(double) * n));
    XT_TRY(xt_prec_apply(g.pr, 1, 1, nullptr, dy.ptr, dx.ptr, 0, nullptr));
    XT_HIP(hipMemcpy(x, dx.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
    return XT_OK;
}

int xt_kronlusol(const double *a, int q, const double *alu, const int *jlu, const int *ju,
                 int p, const double *x, double *y) {
    XT_ARG(a && q > 0 && q <= 8 && alu && jlu && ju && p > 0 && x && y, "kronlusol arguments");
    std::vector<double> ar(q * q), inv(q * q);
    colmajor_to_rowmajor(q, a, ar.data());
    if (small_inverse_rowmajor(q, ar.data(), inv.data())) {
        set_error("Matrix is numerically singular!");
        return XT_E_SINGULAR;
    }
    PrecGuard g;
    XT_TRY(xt_prec_create(p, alu, jlu, ju, &g.pr));
    const size_t n = (size_t)p * q;
    DevBuf dm, tx, dx, dy;
    XT_TRY(dm.upload(inv.data(), sizeof(double) * q * q));
    XT_TRY(tx.upload(x, sizeof(double) * n));
    XT_TRY(dx.reserve(sizeof(double) * n));
    XT_TRY(dy.reserve(sizeof(double) * n));
    XT_TRY(xt_to_interleaved(p, q, 1, tx.as<double>(), dx.as<double>(), nullptr));
    XT_TRY(xt_prec_apply(g.pr, 1, q, dm.ptr, dx.ptr, dy.ptr, 0, nullptr));
    XT_TRY(xt_from_interleaved(p, q, 1, dy.as<double>(), tx.as<double>(), nullptr));
    XT_HIP(hipMemcpy(y, tx.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
    return XT_OK;
}

// One host-array GMRES solve of (A (x) K + B (x) M) sol = rhs with the preconditioner pr
// (kronpgmres and mpipgmres share it): the operator is built from the CSR arrays, rhs and
// sol staged into the interleaved device layout, the FP64 slab solve run with the
// reference's algorithm (mode 0), sol and the residual history copied back.
static int host_pgmres(const double *a, const double *b, int q, const int *ki, const int *kj,
                       const double *kv, const int *mi, const int *mj, const double *mv, int p,
                       xt_prec pr, double *rhs, double *sol, int im, double eps, int maxits,
                       double *hist, int hist_cap, int *nhist, int *its, int *ierr) {
    OpGuard og;
    const int *rp[2] = {ki, mi};
    const int *ci[2] = {kj, mj};
    const double *vv[2] = {kv, mv};
    XT_TRY(xt_op_create(p, q, 2, rp, ci, vv, &og.op));
    const size_t n = (size_t)p * q;
    std::vector<double> ab(2 * q * q);
    for (int k = 0; k < q * q; ++k) { ab[k] = a[k]; ab[q * q + k] = b[k]; }
    DevBuf t, db, dxs;
    XT_TRY(t.reserve(sizeof(double) * n));
    XT_TRY(db.reserve(sizeof(double) * n));
    XT_TRY(dxs.reserve(sizeof(double) * n));
    XT_HIP(hipMemcpy(t.ptr, rhs, sizeof(double) * n, hipMemcpyHostToDevice));
    XT_TRY(xt_to_interleaved(p, q, 1, t.as<double>(), db.as<double>(), nullptr));
    XT_HIP(hipMemcpy(t.ptr, sol, sizeof(double) * n, hipMemcpyHostToDevice));
    XT_TRY(xt_to_interleaved(p, q, 1, t.as<double>(), dxs.as<double>(), nullptr));
    std::vector<double> h;
    int itv = 0, ie = 0;
    double rr = 0.0;
    XT_TRY(slab_solve_impl(og.op, pr, 1, ab.data(), db.as<double>(), dxs.as<double>(), im,
                           eps, maxits, 0, &itv, &ie, &rr, nullptr, &h));
    XT_TRY(xt_from_interleaved(p, q, 1, dxs.as<double>(), t.as<double>(), nullptr));
    XT_HIP(hipMemcpy(sol, t.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
    *ierr = ie;
    if (its) *its = itv;
    if (nhist) {
        const int m = (int)h.size() < hist_cap ? (int)h.size() : hist_cap;
        for (int k = 0; k < m && hist; ++k) hist[k] = h[k];
        *nhist = m;
    }
    return XT_OK;
}

int xt_kronpgmres(const double *a, const double *b, int q, const int *ki, const int *kj,
                  const double *kv, const int *mi, const int *mj, const double *mv, int p,
                  const double *alu, const int *jlu, const int *ju, double *rhs, double *sol,
                  int im, double eps, int maxits, double *hist, int hist_cap, int *nhist,
                  int *its, int *ierr) {
    XT_ARG(a && b && q > 0 && q <= 8 && p > 0 && rhs && sol && ierr, "kronpgmres arguments");
    PrecGuard pg;
    XT_TRY(xt_prec_create(p, alu, jlu, ju, &pg.pr));
    return host_pgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, pg.pr, rhs, sol, im, eps, maxits,
                       hist, hist_cap, nhist, its, ierr);
}

int xt_mpipgmres(const double *a, const double *b, int q, const int *ki, const int *kj,
                 const double *kv, const int *mi, const int *mj, const double *mv, int p,
                 xt_prec pr, double *rhs, double *sol, int im, double eps, int maxits,
                 double *hist, int hist_cap, int *nhist, int *its, int *ierr) {
    XT_ARG(a && b && q > 0 && q <= 8 && p > 0 && pr && rhs && sol && ierr,
           "mpipgmres arguments");
    return host_pgmres(a, b, q, ki, kj, kv, mi, mj, mv, p, pr, rhs, sol, im, eps, maxits, hist,
                       hist_cap, nhist, its, ierr);
}

int xt_kokkos_host(int rank, int ne, int nn, int ndofn, int nnse, int nnte, int ngpe, int ncmp,
                   int nip, const int *nodes, const double *Nx, const double *Ny,
                   const double *Nz, const double *Nt, const double *sol, const double *ym,
                   const double *nu, double *eps_old, double *eps_pl, double *sig_eff,
                   double *sig_back, double *p, double *D, double *ws, double *del_po,
                   double *sig_d) {
    (void)rank;
    XT_ARG(ndofn == 3 && nnse == 8 && nnte == 3 && ngpe == 8 && ncmp == 6,
           "the damage kernel is for C3D8 elements (ndofn 3, nnse 8, nnte 3, ngpe 8, ncmp 6)");
    XT_ARG(ne > 0 && nn > 0 && nip >= 0, "damage sizes");
    const size_t ngp = (size_t)ngpe * ne, nsol = (size_t)2 * nnte * nn * ndofn;
    DevBuf dn, dx, dy, dz, dt, ds, dym, dnu, e0, e1, e2, e3, s0, s1, s2, s3, s4;
    XT_TRY(dn.upload(nodes, sizeof(int) * nnse * ngp));
    XT_TRY(dx.upload(Nx, sizeof(double) * nnse * ngp));
    XT_TRY(dy.upload(Ny, sizeof(double) * nnse * ngp));
    XT_TRY(dz.upload(Nz, sizeof(double) * nnse * ngp));
    XT_TRY(dt.upload(Nt, sizeof(double) * 2 * nnte * (size_t)(nip > 0 ? nip : 1)));
    XT_TRY(ds.upload(sol, sizeof(double) * nsol));
    XT_TRY(dym.upload(ym, sizeof(double) * ngp));
    XT_TRY(dnu.upload(nu, sizeof(double) * ngp));
    double *st6[4] = {eps_old, eps_pl, sig_eff, sig_back};
    DevBuf *d6[4] = {&e0, &e1, &e2, &e3};
    double *st1[5] = {p, D, ws, del_po, sig_d};
    DevBuf *d1[5] = {&s0, &s1, &s2, &s3, &s4};
    for (int k = 0; k < 4; ++k) XT_TRY(d6[k]->upload(st6[k], sizeof(double) * ncmp * ngp));
    for (int k = 0; k < 5; ++k) XT_TRY(d1[k]->upload(st1[k], sizeof(double) * ngp));
    XT_TRY(xt_damage_device(ne, nn, nip, dn.as<int>(), dx.as<double>(), dy.as<double>(),
                            dz.as<double>(), dt.as<double>(), ds.as<double>(), dym.as<double>(),
                            dnu.as<double>(), e0.as<double>(), e1.as<double>(), e2.as<double>(),
                            e3.as<double>(), s0.as<double>(), s1.as<double>(), s2.as<double>(),
                            s3.as<double>(), s4.as<double>(), nullptr));
    for (int k = 0; k < 4; ++k)
        XT_HIP(hipMemcpy(st6[k], d6[k]->ptr, sizeof(double) * ncmp * ngp, hipMemcpyDeviceToHost));
    for (int k = 0; k < 5; ++k)
        XT_HIP(hipMemcpy(st1[k], d1[k]->ptr, sizeof(double) * ngp, hipMemcpyDeviceToHost));
    return XT_OK;
}

// The reference's own C symbols of the damage boundary, so kokkos_damage.f90 and
// main3d.f90 link against this library unchanged.
void kokkos_init_(void) {}
void kokkos_finish_(void) { (void)hipDeviceSynchronize(); }
void kokkos_host_(int rank, int ne, int nn, int ndofn, int nnse, int nnte, int ngpe, int ncmp,
                  int nip, int *nodes, double *Nx, double *Ny, double *Nz, double *Nt,
                  double *sol, double *ym, double *nu, double *eps_old, double *eps_pl,
                  double *sig_eff, double *sig_back, double *p, double *D, double *ws,
                  double *del_po, double *sig_d) {
    const int rc = xt_kokkos_host(rank, ne, nn, ndofn, nnse, nnte, ngpe, ncmp, nip, nodes, Nx,
                                  Ny, Nz, Nt, sol, ym, nu, eps_old, eps_pl, sig_eff, sig_back, p,
                                  D, ws, del_po, sig_d);
    if (rc != XT_OK) {   // the reference has no error return here: fail loudly
        std::fprintf(stderr, "kokkos_host_: %s (code %d)\n", get_error(), rc);
        std::abort();
    }
}

int xt_chkelesta(int ne, int ngpe, double Dc, const double *D, int *elesta, int *gpsta, int *nef,
                 int *nefail) {
    XT_ARG(ne >= 0 && ngpe > 0 && D && elesta && gpsta && nef && nefail, "chkelesta arguments");
    *nef = 0;
    if (ne == 0) return XT_OK;
    const size_t ngp = (size_t)ne * ngpe;
    DevBuf dD, de, dg, df;
    XT_TRY(dD.upload(D, sizeof(double) * ngp));
    XT_TRY(de.upload(elesta, sizeof(int) * ne));
    XT_TRY(dg.upload(gpsta, sizeof(int) * ngp));
    XT_TRY(df.reserve(sizeof(int) * ne));
    XT_TRY(elestatus_device(ne, ngpe, Dc, dD.as<double>(), de.as<int>(), dg.as<int>(),
                            df.as<int>(), nullptr));
    std::vector<int> failed(ne);
    XT_HIP(hipMemcpy(failed.data(), df.ptr, sizeof(int) * ne, hipMemcpyDeviceToHost));
    XT_HIP(hipMemcpy(elesta, de.ptr, sizeof(int) * ne, hipMemcpyDeviceToHost));
    XT_HIP(hipMemcpy(gpsta, dg.ptr, sizeof(int) * ngp, hipMemcpyDeviceToHost));
    int cnt = 0;
    for (int e = 0; e < ne; ++e)
        if (failed[e]) nefail[cnt++] = e + 1;
    *nef = cnt;
    return XT_OK;
}

}  // extern "C"
