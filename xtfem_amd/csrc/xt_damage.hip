// Two-scale fatigue damage per Gauss point, FP64, device-resident state.
//
// Restates the functor Kokkos_Kernel::operator() (kokkos_kernel.h:44-344, constants
// damage.h:8-17) launched by kokkos_host_ (kokkos_host.cpp:37-190): one lane per
// Gauss point walks the nip temporal interpolation points of the slab in order,
// with the elastic predictor / plastic corrector / Lemaitre damage update of the
// reference.
//
// The return map branches on the trial stress at every time point; at load
// reversals the trial stress lies on the yield surface, so a last-bit difference
// flips branches and the history variables diverge.  The kernel therefore keeps
// the reference's exact operation sequence (same products, same summation order,
// sums started from 0.0) with floating-point contraction disabled in this file,
// and evaluates the damage power (Y/S)^0.5 as the correctly rounded sqrt, the
// same choice as the oracle.  Division and sqrt are IEEE correctly rounded on
// gfx950, so every lane reproduces the CPU restatement bit for bit.
//
// Work sharing: the 8 Gauss points of a C3D8 element interpolate the same 24
// nodal displacements disp = sum_k Nt_k u_k (kokkos_kernel.h:58-71).  The 8
// lanes of an element each interpolate one node (identical operations to the
// reference) and exchange the 24 values through LDS, so the interpolation costs
// 36 instead of 288 FP64 operations per lane and time point.  The per-GP state
// stays in registers across the nip steps and is written back once.
// Element status (chkelesta, delelement.f90:109-136) is a second small kernel.
#include "xt_common.h"

#pragma clang fp contract(off)

namespace xt {

// Material constants of the damage law.  Defaults: damage.h:9-17 (Cy = 1740, h = 0.2,
// S = 0.5, s = 0.5, sig_f = 180, sig_u = 577, eps_D = 0.08); the damage threshold is
// wd = (sig_u - sig_f) eps_D (kokkos_kernel.h:47).
struct DamageParams {
    double sigf = 180.0, sigu = 577.0, epsd = 0.08, cy = 1740.0, h = 0.2, S = 0.5, sexp = 0.5;
};
#define XT_CY dp.cy
#define XT_H dp.h
#define XT_S dp.S
#define XT_SIGF dp.sigf
#define XT_WD ((dp.sigu - dp.sigf) * dp.epsd)

// Block = one wave of 64 lanes = 8 elements x 8 Gauss points (GP ii = 8 e + g).
__global__ void __launch_bounds__(64)
damage_kernel(int ngp, int nn, int nip, const int *__restrict__ nodes,
              const double *__restrict__ Nx, const double *__restrict__ Ny,
              const double *__restrict__ Nz, const double *__restrict__ Nt,
              const double *__restrict__ sol, const double *__restrict__ ym,
              const double *__restrict__ nu, double *__restrict__ eps_old,
              double *__restrict__ eps_pl, double *__restrict__ sig_eff,
              double *__restrict__ sig_back, double *__restrict__ pp, double *__restrict__ dd,
              double *__restrict__ wsv, double *__restrict__ del_po_v,
              double *__restrict__ sig_d_v, const DamageParams dp) {
    __shared__ double xch[2][64][4];  // [parity][lane][disp of the lane's node, padded]
    const int lane = threadIdx.x;
    const int ii = blockIdx.x * 64 + lane;
    const bool act = ii < ngp;
    const int g = lane & 7;       // node of the element this lane interpolates
    const int e0 = lane & ~7;     // first lane of this lane's element
    // nodal temporal coefficients of node g: sl[k][j] = sol(k, node, j)
    double sl[6][3];
    double nx[8], ny[8], nz[8];
    double nui = 0.3, ymi = 1.0;
    double eo[6], ep[6], se[6], sb[6];
    double p = 0.0, d = 0.0, ws = 0.0, del_po = 0.0, sig_d = 0.0;
    if (act) {
        const size_t plane = (size_t)nn * 3;
        const size_t nd = (size_t)(nodes[(size_t)ii * 8 + g] - 1) * 3;
        for (int k = 0; k < 6; ++k)
            for (int j = 0; j < 3; ++j) sl[k][j] = sol[k * plane + nd + j];
        for (int a = 0; a < 8; ++a) {
            nx[a] = Nx[(size_t)ii * 8 + a];
            ny[a] = Ny[(size_t)ii * 8 + a];
            nz[a] = Nz[(size_t)ii * 8 + a];
        }
        nui = nu[ii];
        ymi = ym[ii];
        for (int c = 0; c < 6; ++c) {
            eo[c] = eps_old[(size_t)ii * 6 + c];
            ep[c] = eps_pl[(size_t)ii * 6 + c];
            se[c] = sig_eff[(size_t)ii * 6 + c];
            sb[c] = sig_back[(size_t)ii * 6 + c];
        }
        p = pp[ii]; d = dd[ii]; ws = wsv[ii]; del_po = del_po_v[ii]; sig_d = sig_d_v[ii];
    } else {
        for (int k = 0; k < 6; ++k) for (int j = 0; j < 3; ++j) sl[k][j] = 0.0;
        for (int a = 0; a < 8; ++a) { nx[a] = 0.0; ny[a] = 0.0; nz[a] = 0.0; }
        for (int c = 0; c < 6; ++c) { eo[c] = 0.0; ep[c] = 0.0; se[c] = 0.0; sb[c] = 0.0; }
    }
    // constants of kokkos_kernel.h:49-53 and the elastic matrix of :88-100
    const double a = (1.0 + nui) / (3.0 * (1.0 - nui));
    const double b = 2.0 * (4.0 - 5.0 * nui) / (15.0 * (1.0 - nui));
    const double G = ymi / (2.0 * (1.0 + nui));
    const double Kb = ymi / (3.0 * (1.0 - 2.0 * nui));
    const double coef = ymi / ((1. + nui) * (1. - 2. * nui));
    const double dl = coef * nui;
    const double dg = coef * (1.0 - nui);
    const double dsh = coef * (1.0 - 2.0 * nui) / 2.0;

    for (int jj = 0; jj < nip; ++jj) {
        const double *nt = Nt + (size_t)jj * 6;
        // disp of node g (kokkos_kernel.h:60-71: disp += Nt_k * sol, k ascending)
        double dsp[3];
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += nt[k] * sl[k][j];
            dsp[j] = s;
        }
        double (*buf)[4] = xch[jj & 1];
        buf[lane][0] = dsp[0];
        buf[lane][1] = dsp[1];
        buf[lane][2] = dsp[2];
        __syncthreads();
        if (!act) continue;
        // strain (kokkos_kernel.h:74-84)
        double strain[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < 8; ++i) {
            const double d0 = buf[e0 + i][0], d1 = buf[e0 + i][1], d2 = buf[e0 + i][2];
            strain[0] += nx[i] * d0;
            strain[1] += ny[i] * d1;
            strain[2] += nz[i] * d2;
            strain[3] += ny[i] * d0 + nx[i] * d1;
            strain[4] += nz[i] * d1 + ny[i] * d2;
            strain[5] += nz[i] * d0 + nx[i] * d2;
        }
        double del_p = 0.0;
        // elastic predictor (kokkos_kernel.h:104-171)
        const double c1 = 1.0 / (1.0 - b * d);
        const double c2 = (a - b) * d / ((1.0 - b * d) * (1.0 - a * d));
        const double c3 = b * (1.0 - d) / (1.0 - b * d);
        double tel[6], tee[6];
        for (int i = 0; i < 6; ++i) tel[i] = c1 * strain[i] + c3 * ep[i];
        double tr = 0.0;
        for (int i = 0; i < 3; ++i) tr += strain[i];
        tr = c2 * tr / 3.0;
        for (int i = 0; i < 3; ++i) tel[i] += tr;
        for (int i = 0; i < 6; ++i) tee[i] = tel[i] - ep[i];
        // D_el * tee with the reference's row sums (zero entries contribute exact zeros)
        double tse[6];
        tse[0] = ((0.0 + dg * tee[0]) + dl * tee[1]) + dl * tee[2];
        tse[1] = ((0.0 + dl * tee[0]) + dg * tee[1]) + dl * tee[2];
        tse[2] = ((0.0 + dl * tee[0]) + dl * tee[1]) + dg * tee[2];
        tse[3] = 0.0 + dsh * tee[3];
        tse[4] = 0.0 + dsh * tee[4];
        tse[5] = 0.0 + dsh * tee[5];
        double trs = 0.0;
        for (int i = 0; i < 3; ++i) trs += tse[i];
        trs = trs / 3.0;
        double tsig[6];
        tsig[0] = (tse[0] - trs) - sb[0];
        tsig[1] = (tse[1] - trs) - sb[1];
        tsig[2] = (tse[2] - trs) - sb[2];
        tsig[3] = tse[3] - sb[3];
        tsig[4] = tse[4] - sb[4];
        tsig[5] = tse[5] - sb[5];
        double teq = 0.0;
        for (int i = 0; i < 3; ++i) teq += 1.5 * tsig[i] * tsig[i];
        for (int i = 3; i < 6; ++i) teq += 3.0 * tsig[i] * tsig[i];
        teq = sqrt(teq);
        if (teq <= XT_SIGF) {
            for (int i = 0; i < 6; ++i) { se[i] = tse[i]; eo[i] = strain[i]; }
            sig_d = teq;
            continue;
        }
        // plastic corrector (kokkos_kernel.h:175-263)
        double ds_[6], dsig[6];
        for (int i = 0; i < 6; ++i) ds_[i] = strain[i] - eo[i];
        dsig[0] = (((0.0 + dg * ds_[0]) + dl * ds_[1]) + dl * ds_[2]) / (1.0 - b * d);
        dsig[1] = (((0.0 + dl * ds_[0]) + dg * ds_[1]) + dl * ds_[2]) / (1.0 - b * d);
        dsig[2] = (((0.0 + dl * ds_[0]) + dl * ds_[1]) + dg * ds_[2]) / (1.0 - b * d);
        dsig[3] = (0.0 + dsh * ds_[3]) / (1.0 - b * d);
        dsig[4] = (0.0 + dsh * ds_[4]) / (1.0 - b * d);
        dsig[5] = (0.0 + dsh * ds_[5]) / (1.0 - b * d);
        const double c4 = Kb * (a - b) * d / ((1.0 - b * d) * (1.0 - a * d));
        double trd = 0.0;
        for (int i = 0; i < 3; ++i) trd += ds_[i];
        for (int i = 0; i < 3; ++i) dsig[i] += c4 * trd;
        double Qs[6];
        for (int i = 0; i < 6; ++i) Qs[i] = sb[i] - se[i] - dsig[i];
        const double gamma = XT_CY * (1.0 - d) + 3.0 * G * (1.0 - b) / (1.0 - b * d);
        double trq = 0.0;
        for (int i = 0; i < 3; ++i) trq += Qs[i];
        trq = trq / 3.0;
        double dq[6] = {Qs[0] - trq, Qs[1] - trq, Qs[2] - trq, Qs[3], Qs[4], Qs[5]};
        double qeq = 0.0;
        for (int i = 0; i < 3; ++i) qeq += 1.5 * dq[i] * dq[i];
        for (int i = 3; i < 6; ++i) qeq += 3.0 * dq[i] * dq[i];
        qeq = sqrt(qeq);
        const double tr_sigs = -trq;
        del_p = (qeq - XT_SIGF) / gamma;
        p += del_p;
        const double c5 = 1.0 + gamma * del_p / XT_SIGF;
        double dvs[6];
        for (int i = 0; i < 6; ++i) dvs[i] = -dq[i] / c5;
        double sigs[6] = {dvs[0] + tr_sigs, dvs[1] + tr_sigs, dvs[2] + tr_sigs, dvs[3], dvs[4], dvs[5]};
        double dpl[6];
        for (int i = 0; i < 6; ++i) {
            dpl[i] = 3.0 * dvs[i] * del_p / (2.0 * XT_SIGF);
            ep[i] += dpl[i];
        }
        for (int i = 0; i < 6; ++i) sb[i] += 2.0 * XT_CY * (1.0 - d) * dpl[i] / 3.0;
        for (int i = 0; i < 6; ++i) se[i] = sigs[i] + sb[i];
        double tse2 = 0.0;
        for (int i = 0; i < 3; ++i) tse2 += se[i];
        tse2 = tse2 / 3.0;
        double dse[6] = {se[0] - tse2, se[1] - tse2, se[2] - tse2, se[3], se[4], se[5]};
        double seq = 0.0;
        for (int i = 0; i < 3; ++i) seq += 1.5 * dse[i] * dse[i];
        for (int i = 3; i < 6; ++i) seq += 3.0 * dse[i] * dse[i];
        seq = sqrt(seq);
        const double sig_new = fabs(seq - XT_SIGF);
        const double sig_old = fabs(sig_d - XT_SIGF);
        ws += (sig_new * del_p + sig_old * del_po) / 2.0;
        if (ws >= XT_WD) {
            // Lemaitre damage increment (kokkos_kernel.h:293-330)
            double sp[3] = {0.0, 0.0, 0.0}, sn[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < 3; ++i) {
                if (se[i] < 0.0) sn[i] = se[i];
                else sp[i] = se[i];
            }
            const double hr = (1.0 - d) / (1.0 - XT_H * d);
            const double Yp = (1.0 + nui) / (2.0 * ymi) *
                              (sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2] +
                               XT_H * hr * hr * (sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2]));
            double trt = 0.0;
            for (int i = 0; i < 3; ++i) trt += se[i];
            double Yn;
            if (trt >= 0.0) Yn = nui / (2.0 * ymi) * trt * trt;
            else Yn = nui / (2.0 * ymi) * XT_H * hr * hr * trt * trt;
            // (Y/S)^s: s = 0.5 as the correctly rounded sqrt (the reference's value)
            d += (dp.sexp == 0.5 ? sqrt((Yp - Yn) / XT_S) : pow((Yp - Yn) / XT_S, dp.sexp)) * del_p;
        } else {
            d = 0.0;
        }
        for (int i = 0; i < 6; ++i) eo[i] = strain[i];
        sig_d = seq;
        del_po = del_p;
    }
    if (!act) return;
    for (int c = 0; c < 6; ++c) {
        eps_old[(size_t)ii * 6 + c] = eo[c];
        eps_pl[(size_t)ii * 6 + c] = ep[c];
        sig_eff[(size_t)ii * 6 + c] = se[c];
        sig_back[(size_t)ii * 6 + c] = sb[c];
    }
    pp[ii] = p;
    dd[ii] = d;
    wsv[ii] = ws;
    del_po_v[ii] = del_po;
    sig_d_v[ii] = sig_d;
}

// element e failed iff any of its ngpe Gauss points has D >= Dc
__global__ void __launch_bounds__(256)
elestatus_kernel(int ne, int ngpe, double Dc, const double *__restrict__ D, int *elesta,
                 int *gpsta, int *failed) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int any = 0;
    for (int g = 0; g < ngpe; ++g) any |= (D[(size_t)e * ngpe + g] >= Dc);
    failed[e] = any;
    if (any) {
        elesta[e] = 0;
        for (int g = 0; g < ngpe; ++g) gpsta[(size_t)e * ngpe + g] = 0;
    }
}

// kokkos_damage.f90:102-107: a failed Gauss point carries D = ws = 0
__global__ void __launch_bounds__(256)
zero_failed_kernel(int ngp, const int *__restrict__ gpsta, double *D, double *ws) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngp && gpsta[g] == 0) {
        D[g] = 0.0;
        ws[g] = 0.0;
    }
}

}  // namespace xt

using namespace xt;

extern "C" int xt_damage_device_params(int ne, int nn, int nip, const int *nodes,
                                       const double *Nx, const double *Ny, const double *Nz,
                                       const double *Nt, const double *sol, const double *ym,
                                       const double *nu, double *eps_old, double *eps_pl,
                                       double *sig_eff, double *sig_back, double *p, double *D,
                                       double *ws, double *del_po, double *sig_d,
                                       const double *params, void *stream) {
    XT_ARG(ne >= 0 && nn > 0 && nip >= 0, "damage sizes");
    DamageParams dp;
    if (params) {
        dp.sigf = params[0]; dp.sigu = params[1]; dp.epsd = params[2]; dp.cy = params[3];
        dp.h = params[4]; dp.S = params[5]; dp.sexp = params[6];
        XT_ARG(dp.sigf > 0.0 && dp.sigu > dp.sigf && dp.S > 0.0, "damage parameters");
    }
    XT_ARG(nodes && Nx && Ny && Nz && (nip == 0 || Nt) && sol && ym && nu && eps_old && eps_pl &&
               sig_eff && sig_back && p && D && ws && del_po && sig_d,
           "damage arrays");
    const int ngp = 8 * ne;
    if (ngp == 0 || nip == 0) return XT_OK;
    const bool prof = prof_on(XT_PROF_DAMAGE);
    if (prof) prof_begin(XT_PROF_DAMAGE, (hipStream_t)stream);
    hipLaunchKernelGGL(damage_kernel, dim3(ceil_div(ngp, 64)), dim3(64), 0,
                       (hipStream_t)stream, ngp, nn, nip, nodes, Nx, Ny, Nz, Nt, sol, ym, nu,
                       eps_old, eps_pl, sig_eff, sig_back, p, D, ws, del_po, sig_d, dp);
    if (prof)   // state 31 FP64 read + written, 24 derivatives + 8 node ids, Nt, 24 disp gathers per step
        prof_end(XT_PROF_DAMAGE, (hipStream_t)stream,
                 (double)ngp * (31 * 16 + 24 * 8 + 8 * 4) + 8.0 * 6 * nip + 8.0 * 18 * nn, 0.0);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

extern "C" int xt_damage_device(int ne, int nn, int nip, const int *nodes, const double *Nx,
                                const double *Ny, const double *Nz, const double *Nt,
                                const double *sol, const double *ym, const double *nu,
                                double *eps_old, double *eps_pl, double *sig_eff,
                                double *sig_back, double *p, double *D, double *ws,
                                double *del_po, double *sig_d, void *stream) {
    return xt_damage_device_params(ne, nn, nip, nodes, Nx, Ny, Nz, Nt, sol, ym, nu, eps_old,
                                   eps_pl, sig_eff, sig_back, p, D, ws, del_po, sig_d, nullptr,
                                   stream);
}

namespace xt {
int elestatus_device(int ne, int ngpe, double Dc, const double *D, int *elesta, int *gpsta,
                     int *failed, hipStream_t s) {
    if (ne == 0) return XT_OK;
    hipLaunchKernelGGL(elestatus_kernel, dim3(ceil_div(ne, 256)), dim3(256), 0, s, ne, ngpe, Dc,
                       D, elesta, gpsta, failed);
    XT_HIP(hipGetLastError());
    return XT_OK;
}
}  // namespace xt

extern "C" int xt_elestatus_device(int ne, int ngpe, double Dc, const double *D, int *elesta,
                                   int *gpsta, int *failed, void *stream) {
    XT_ARG(ne >= 0 && ngpe > 0, "elestatus sizes");
    if (ne == 0) return XT_OK;
    XT_ARG(D && elesta && gpsta && failed, "elestatus arrays");
    return elestatus_device(ne, ngpe, Dc, D, elesta, gpsta, failed, (hipStream_t)stream);
}

extern "C" int xt_damage_zero_failed(int ngp, const int *gpsta, double *D, double *ws,
                                     void *stream) {
    XT_ARG(ngp >= 0, "zero_failed size");
    if (ngp == 0) return XT_OK;
    XT_ARG(gpsta && D && ws, "zero_failed arrays");
    hipLaunchKernelGGL(zero_failed_kernel, dim3(ceil_div(ngp, 256)), dim3(256), 0,
                       (hipStream_t)stream, ngp, gpsta, D, ws);
    XT_HIP(hipGetLastError());
    return XT_OK;
}
