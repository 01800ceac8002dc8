// Spatial assembly on the GPU: element matrices of every C3D8 element, summed into the
// global K and M in the reference's order, boundary conditions applied, all in HBM.
//
// Restates mpiglbmtx_m (mpisubs.f90:69-139) with elemtx_m (elesubs.f90:2-58) and
// setsdstiff (elesubs.f90:374-401) per Gauss point, tricsr (spmtx.f90:359-407) and
// csrapplybc (spmtx.f90:268-307).  The driver repeats this every slab because E and nu
// follow the temperature (main3d.f90:752-783, thermoelastic.f90:1-46); here only the
// arithmetic repeats: the symbolic part (global pattern, and for each global entry the
// element contributions in element order) is built once per mesh on the host.
//
// * element kernel: one 64-lane workgroup per element.  Lanes 0-7 each take one Gauss
//   point: Jacobian, determinant, cofactor inverse, shape derivatives, the B matrix and
//   D B with D from that point's E and nu; all 64 lanes then form 9 of the 576 entries
//   of ke and me each, summing the points in order.  Every product and sum is rounded
//   as written (no contraction), with the loop orders of the reference's MATMULs, so ke
//   and me are bit-identical to the reference's.
// * gather kernel: one lane per global entry sums its element contributions in element
//   order (tricsr adds duplicates in that order) and applies the 1-0 method on fixed
//   DOFs.  The pattern keeps entries that sum to zero (tricsr would drop them).
// * thermal side (1 DOF a node): conductivity and capacity per element as elemtx_t
//   (elesubs.f90:60-116), gathered like K and M (mpiglbmtx_t, mpisubs.f90:141-188); and
//   per slab the temperature-dependent material (material_update, thermoelastic.f90:1-46)
//   with the thermal load vv of mpiglbmtx_m (calc_thermo_stress thermoelastic.f90:48-78,
//   elevec_m elesubs.f90:119-167, summed per DOF in element order, mpisubs.f90:129-131).
//   elemtx_t and elevec_m form the Jacobian transposed with respect to elemtx_m
//   (jac(m,p) = sum_a dns(a,p) x(a,m)), so their determinant and inverse round
//   differently; both forms are kept.
#include <algorithm>
#include <new>
#include <vector>

#include "xt_common.h"

#pragma clang fp contract(off)

namespace xt {
namespace {

constexpr double kGauss2 = 0.577350269189626;   // procs.f90:537
// local (xi, eta, zeta) sign of each C3D8 node (shape3x, procs.f90:1320-1327)
constexpr int kSign[8][3] = {{-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1},
                             {1, -1, 1},  {1, -1, -1},  {1, 1, -1},  {1, 1, 1}};

// N[g][a] and dN[g][m][a] at the 8 Gauss points (g = 4i + 2j + k), shape3x/shape3dx.
void shape_tables(double *N, double *dN) {
    const double gp[2] = {-kGauss2, kGauss2};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const int g = 4 * i + 2 * j + k;
                const double pt[3] = {gp[i], gp[j], gp[k]};
                for (int a = 0; a < 8; ++a) {
                    double f[3];
                    for (int d = 0; d < 3; ++d) f[d] = 1.0 + kSign[a][d] * pt[d];
                    N[g * 8 + a] = 0.125 * f[0] * f[1] * f[2];
                    dN[(g * 3 + 0) * 8 + a] = (0.125 * kSign[a][0]) * f[1] * f[2];
                    dN[(g * 3 + 1) * 8 + a] = (0.125 * kSign[a][1]) * f[0] * f[2];
                    dN[(g * 3 + 2) * 8 + a] = (0.125 * kSign[a][2]) * f[0] * f[1];
                }
            }
}

// 3x3 determinant with the term order of procs.f90:476-478 (J row-major, J[m][p]).
__host__ __device__ inline double det3(const double (&J)[3][3]) {
    return J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
           J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2] - J[0][2] * J[1][1] * J[2][0];
}

// jac[m][p] = sum_a dN[m][a] X[a][p], a in order (MATMUL(dns, ncoorde))
__host__ __device__ inline void jacobian(const double *dN, const double (*X)[3],
                                         double (&J)[3][3]) {
    for (int m = 0; m < 3; ++m)
        for (int p = 0; p < 3; ++p) {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s = s + dN[m * 8 + a] * X[a][p];
            J[m][p] = s;
        }
}

// cofactor inverse (procs.f90:300-310), then / det
__host__ __device__ inline void inverse3(const double (&J)[3][3], double det, double (&gam)[3][3]) {
    gam[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    gam[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    gam[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    gam[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    gam[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    gam[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    gam[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    gam[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    gam[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) gam[i][j] = gam[i][j] / det;
}

// Geometry of elemtx_t / elevec_m at one Gauss point (elesubs.f90:86-95, 143-152):
// jac(m,p) = sum_a dns(a,p) x(a,m) (the transpose of elemtx_m's), its determinant and
// inverse, and G[m][a] = sum_p gam(p,m) dns(a,p) = dN_a/dx_m.
__device__ inline void thermal_geometry(const double *dN, const double (*X)[3],
                                        double (&G)[3][8], double &det) {
    double J[3][3], T[3][3], gam[3][3];
    jacobian(dN, X, J);
    for (int m = 0; m < 3; ++m)
        for (int p = 0; p < 3; ++p) T[m][p] = J[p][m];
    det = det3(T);
    inverse3(T, det, gam);
    for (int m = 0; m < 3; ++m)
        for (int a = 0; a < 8; ++a) {
            double s = 0.0;
            for (int p = 0; p < 3; ++p) s = s + gam[p][m] * dN[p * 8 + a];
            G[m][a] = s;
        }
}

__global__ void __launch_bounds__(64)
element_kernel(int ne, const double *__restrict__ xcoord, const int *__restrict__ xconn,
               const double *__restrict__ ym, const double *__restrict__ nu, double rho,
               const double *__restrict__ Ntab, const double *__restrict__ dNtab,
               double *__restrict__ ke, double *__restrict__ me, double *__restrict__ Nx,
               double *__restrict__ Ny, double *__restrict__ Nz) {
    __shared__ double sB[8][6][24], sDB[8][6][24], sD[8][6][6], sX[8][3], sdet[8], sN[8][8];
    const int e = blockIdx.x, lane = threadIdx.x;
    for (int q = lane; q < 8 * 6 * 24; q += 64) (&sB[0][0][0])[q] = 0.0;
    for (int q = lane; q < 8 * 36; q += 64) (&sD[0][0][0])[q] = 0.0;
    if (lane < 24) {
        const int a = lane / 3, d = lane - 3 * (lane / 3);
        sX[a][d] = xcoord[(size_t)(xconn[(size_t)e * 8 + a] - 1) * 3 + d];
    }
    __syncthreads();
    if (lane < 8) {   // one Gauss point per lane
        const int l = lane;
        const double *dN = dNtab + l * 24;
        double J[3][3];
        jacobian(dN, sX, J);
        const double det = det3(J);
        double gam[3][3];
        inverse3(J, det, gam);
        const size_t row = (size_t)e * 8 + l;
        // shape derivatives dN_a/dx_p = sum_m gam[p][m] dN[m][a]; B of elemtx_m
        for (int a = 0; a < 8; ++a) {
            double g[3];
            for (int p = 0; p < 3; ++p) {
                double s = 0.0;
                for (int m = 0; m < 3; ++m) s = s + gam[p][m] * dN[m * 8 + a];
                g[p] = s;
            }
            Nx[row * 8 + a] = g[0];
            Ny[row * 8 + a] = g[1];
            Nz[row * 8 + a] = g[2];
            sB[l][0][3 * a + 0] = g[0];
            sB[l][1][3 * a + 1] = g[1];
            sB[l][2][3 * a + 2] = g[2];
            sB[l][3][3 * a + 0] = g[1];
            sB[l][3][3 * a + 1] = g[0];
            sB[l][4][3 * a + 1] = g[2];
            sB[l][4][3 * a + 2] = g[1];
            sB[l][5][3 * a + 0] = g[2];
            sB[l][5][3 * a + 2] = g[0];
        }
        // setsdstiff with this point's E and nu
        const double E = ym[row], v = nu[row];
        const double cf = E / ((1.0 + v) * (1.0 - 2.0 * v));
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) sD[l][i][j] = cf * v;
            sD[l][i][i] = cf * (1.0 - v);
            sD[l][i + 3][i + 3] = cf * (0.5 - v);
        }
        sdet[l] = det;
        for (int a = 0; a < 8; ++a) sN[l][a] = Ntab[l * 8 + a];
    }
    __syncthreads();
    // D B per point, sum over j in order (MATMUL(stiff, bmat))
    for (int q = lane; q < 8 * 6 * 24; q += 64) {
        const int l = q / 144, rem = q - 144 * l, i = rem / 24, c = rem - 24 * i;
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s = s + sD[l][i][j] * sB[l][j][c];
        sDB[l][i][c] = s;
    }
    __syncthreads();
    // ke = sum_l (B^T D B) det, me = sum_l (N^T rho N per component) det, points in order
    for (int q = lane; q < 576; q += 64) {
        const int r = q / 24, c = q - r * 24;
        double sk = 0.0, sm = 0.0;
        for (int l = 0; l < 8; ++l) {
            double ket = 0.0;
            for (int i = 0; i < 6; ++i) ket = ket + sB[l][i][r] * sDB[l][i][c];
            sk = sk + ket * sdet[l];
            const double met = (r % 3 == c % 3) ? sN[l][r / 3] * (rho * sN[l][c / 3]) : 0.0;
            sm = sm + met * sdet[l];
        }
        ke[(size_t)e * 576 + q] = sk;
        me[(size_t)e * 576 + q] = sm;
    }
}

// val[k] = sum of the contributions of entry k in element order, then the 1-0 method
__global__ void __launch_bounds__(256)
gather_kernel(long long nnz, const int *__restrict__ row, const int *__restrict__ col,
              const long long *__restrict__ cptr, const int *__restrict__ cidx,
              const double *__restrict__ src, const int *__restrict__ fixed,
              double *__restrict__ val) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const long long j0 = cptr[k], j1 = cptr[k + 1];
    double s = src[cidx[j0]];
    for (long long j = j0 + 1; j < j1; ++j) s = s + src[cidx[j]];
    const int i = row[k], jc = col[k] - 1;
    if (fixed && (fixed[i] || fixed[jc])) s = (i == jc) ? 1.0 : 0.0;
    val[k] = s;
}

// elemtx_t for one element per 64-lane workgroup: lanes 0-7 form one Gauss point's
// geometry each, then lane a*8+b sums entry (a, b) over the points in order,
// me += ((rho cp N_a) N_b) det and ke += ((kappa G_ma) G_mb) det for m = x, y, z.
__global__ void __launch_bounds__(64)
thermal_element_kernel(int ne, const double *__restrict__ xcoord, const int *__restrict__ xconn,
                       double kappa, double rhocp, const double *__restrict__ Ntab,
                       const double *__restrict__ dNtab, double *__restrict__ kt,
                       double *__restrict__ mt) {
    __shared__ double sG[8][3][8], sX[8][3], sdet[8], sN[8][8];
    const int e = blockIdx.x, lane = threadIdx.x;
    if (lane < 24) {
        const int a = lane / 3, d = lane - 3 * (lane / 3);
        sX[a][d] = xcoord[(size_t)(xconn[(size_t)e * 8 + a] - 1) * 3 + d];
    }
    sN[lane >> 3][lane & 7] = Ntab[lane];
    __syncthreads();
    if (lane < 8) {
        double G[3][8], det;
        thermal_geometry(dNtab + lane * 24, sX, G, det);
        for (int m = 0; m < 3; ++m)
            for (int a = 0; a < 8; ++a) sG[lane][m][a] = G[m][a];
        sdet[lane] = det;
    }
    __syncthreads();
    const int a = lane >> 3, b = lane & 7;
    double sk = 0.0, sm = 0.0;
    for (int l = 0; l < 8; ++l) {
        sm = sm + ((rhocp * sN[l][a]) * sN[l][b]) * sdet[l];
        for (int m = 0; m < 3; ++m) sk = sk + ((kappa * sG[l][m][a]) * sG[l][m][b]) * sdet[l];
    }
    kt[(size_t)e * 64 + lane] = sk;
    mt[(size_t)e * 64 + lane] = sm;
}

// Default-real (FP32) literals of material_update, thermoelastic.f90:34-35
constexpr double kYmDt1 = (double)1.6f, kYmDt2 = (double)0.1f, kNuDt1 = (double)5.0e-5f;

// One slab's material and thermal load, one element per 64-lane workgroup.  Lanes 0-7
// (one Gauss point each): T at the point (sum_a N_a T_a, a in order), E and nu of
// material_update, the thermal stress of calc_thermo_stress ((D11 + D12 + D13) alpha
// (T - Tref) with D of setsdstiff) and the geometry of elevec_m; lanes 0-23 then sum
// ve(3a + m) = sum_l (G_ma sigma_m) det over the points in order.
__global__ void __launch_bounds__(64)
thermo_load_kernel(int ne, const double *__restrict__ xcoord, const int *__restrict__ xconn,
                   const double *__restrict__ T, long long tstride, long long toff, double Tref,
                   double alpha, double ym0, double nu0, const double *__restrict__ Ntab,
                   const double *__restrict__ dNtab, double *__restrict__ ym,
                   double *__restrict__ nu, double *__restrict__ ve) {
    __shared__ double sG[8][3][8], sX[8][3], sdet[8], sS[8][3], sT[8];
    const int e = blockIdx.x, lane = threadIdx.x;
    if (lane < 24) {
        const int a = lane / 3, d = lane - 3 * (lane / 3);
        sX[a][d] = xcoord[(size_t)(xconn[(size_t)e * 8 + a] - 1) * 3 + d];
    } else if (lane < 32) {
        const int a = lane - 24;
        sT[a] = T[(long long)(xconn[(size_t)e * 8 + a] - 1) * tstride + toff];
    }
    __syncthreads();
    if (lane < 8) {
        const int l = lane;
        double temp = 0.0;
        for (int a = 0; a < 8; ++a) temp = temp + Ntab[l * 8 + a] * sT[a];
        const double d = temp - Tref;
        double E = ym0, v = nu0;
        if (temp >= 0.0) {
            E = (ym0 - kYmDt1 * d) - (kYmDt2 * d) * d;
            v = nu0 + kNuDt1 * d;
        }
        const size_t g = (size_t)e * 8 + l;
        ym[g] = E;
        nu[g] = v;
        const double cf = E / ((1.0 + v) * (1.0 - 2.0 * v));
        const double dd = cf * (1.0 - v), od = cf * v;
        const double strain = alpha * (temp - Tref);
        sS[l][0] = ((dd + od) + od) * strain;
        sS[l][1] = ((od + dd) + od) * strain;
        sS[l][2] = ((od + od) + dd) * strain;
        double G[3][8], det;
        thermal_geometry(dNtab + l * 24, sX, G, det);
        for (int m = 0; m < 3; ++m)
            for (int a = 0; a < 8; ++a) sG[l][m][a] = G[m][a];
        sdet[l] = det;
    }
    __syncthreads();
    if (lane < 24) {
        const int a = lane / 3, m = lane - 3 * (lane / 3);
        double v = 0.0;
        for (int l = 0; l < 8; ++l) v = v + (sG[l][m][a] * sS[l][m]) * sdet[l];
        ve[(size_t)e * 24 + lane] = v;
    }
}

// vv[d] = sum of the element vectors' entries for DOF d, elements in order, from 0
__global__ void __launch_bounds__(256)
dof_gather_kernel(int n, const int *__restrict__ ptr, const int *__restrict__ idx,
                  const double *__restrict__ src, double *__restrict__ dst) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    double s = 0.0;
    for (int j = ptr[d]; j < ptr[d + 1]; ++j) s = s + src[idx[j]];
    dst[d] = s;
}

struct Pattern {                 // one global matrix: 1-based CSR plus contribution lists
    long long nnz = 0;
    std::vector<int> rp, ci;     // host copies (1-based)
    DevBuf row, col, cptr, cidx, val;
};

}  // namespace
}  // namespace xt

using namespace xt;

struct xt_asm_s {
    int nn = 0, ne = 0, p = 0;
    DevBuf xcoord, xconn, fixed, Ntab, dNtab, ke, me, Nx, Ny, Nz;
    Pattern K, M;
    // thermal side: one node-level pattern for K_t and M_t; values with the 1-0 method
    // (tval[0] K_t, tval[1] M_t) and without (tval[2], tval[3]); per-DOF lists of the
    // element-vector entries of vv
    bool thermal = false;
    Pattern TP;
    DevBuf tfixed, kt, mt, tval[4], vptr, vidx, ve;
};

namespace {

// Symbolic assembly: for each DOF row, the contributions of every element touching it
// (element order), stably sorted by column; equal columns form one entry whose
// contribution list keeps element order.  same_comp: only couplings of equal
// displacement components (the consistent mass, elesubs.f90:51-53).
// nd: DOFs per element (24 displacement, 8 temperature); element matrices are stored
// nd x nd row-major per element.
int build_pattern(int p, int ne, int nd, const std::vector<int> &dofs, bool same_comp,
                  Pattern &P) {
    std::vector<long long> cnt(p + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int r = 0; r < nd; ++r) cnt[dofs[(size_t)e * nd + r] + 1] += same_comp ? nd / 3 : nd;
    for (int i = 0; i < p; ++i) cnt[i + 1] += cnt[i];
    std::vector<std::pair<int, int>> con((size_t)cnt[p]);   // (column, source index)
    std::vector<long long> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < ne; ++e)
        for (int r = 0; r < nd; ++r) {
            const int i = dofs[(size_t)e * nd + r];
            for (int c = 0; c < nd; ++c) {
                if (same_comp && r % 3 != c % 3) continue;
                con[fill[i]++] = {dofs[(size_t)e * nd + c], e * nd * nd + r * nd + c};
            }
        }
    P.rp.assign(p + 1, 1);
    P.ci.clear();
    std::vector<int> row, cidx;
    std::vector<long long> cptr;
    cidx.reserve(con.size());
    for (int i = 0; i < p; ++i) {
        auto b = con.begin() + cnt[i], en = con.begin() + cnt[i + 1];
        std::stable_sort(b, en, [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
            return x.first < y.first;
        });
        for (auto it = b; it != en; ++it) {
            if (it == b || it->first != (it - 1)->first) {
                P.ci.push_back(it->first + 1);
                row.push_back(i);
                cptr.push_back((long long)cidx.size());
            }
            cidx.push_back(it->second);
        }
        P.rp[i + 1] = (int)P.ci.size() + 1;
    }
    cptr.push_back((long long)cidx.size());
    P.nnz = (long long)P.ci.size();
    XT_TRY(P.row.upload(row.data(), sizeof(int) * row.size()));
    XT_TRY(P.col.upload(P.ci.data(), sizeof(int) * P.ci.size()));
    XT_TRY(P.cptr.upload(cptr.data(), sizeof(long long) * cptr.size()));
    XT_TRY(P.cidx.upload(cidx.data(), sizeof(int) * cidx.size()));
    XT_TRY(P.val.reserve(sizeof(double) * P.nnz));
    return XT_OK;
}

}  // namespace

extern "C" {

int xt_asm_create(int nn, int ne, const double *xcoord, const int *xconn, const int *fixed,
                  xt_asm *out) {
    XT_ARG(out && nn > 0 && ne > 0 && xcoord && xconn && fixed, "asm arguments");
    XT_ARG((long long)ne * 576 < (1LL << 31), "too many elements for 32-bit contribution indices");
    const int p = 3 * nn;
    std::vector<int> dofs((size_t)ne * 24);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a) {
            const int node = xconn[(size_t)e * 8 + a];
            XT_ARG(node >= 1 && node <= nn, "asm: node id out of range");
            for (int c = 0; c < 3; ++c) dofs[(size_t)e * 24 + 3 * a + c] = 3 * (node - 1) + c;
        }
    double N[64], dN[192];
    shape_tables(N, dN);
    // the geometry is fixed: every element's Jacobian determinant is checked once here
    for (int e = 0; e < ne; ++e) {
        double X[8][3];
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) X[a][d] = xcoord[(size_t)(xconn[(size_t)e * 8 + a] - 1) * 3 + d];
        for (int l = 0; l < 8; ++l) {
            double J[3][3];
            jacobian(dN + l * 24, X, J);
            if (!(det3(J) > 0.0)) {
                set_error("asm: element %d has a non-positive Jacobian determinant (node order "
                          "must follow C3D8)", e + 1);
                return XT_E_ARG;
            }
        }
    }
    xt_asm h = new (std::nothrow) xt_asm_s;
    if (!h) return XT_E_NOMEM;
    h->nn = nn;
    h->ne = ne;
    h->p = p;
    std::vector<int> fx(p);
    for (int i = 0; i < p; ++i) fx[i] = fixed[i] != 0;
    int rc = XT_OK;
    if (rc == XT_OK) rc = h->xcoord.upload(xcoord, sizeof(double) * 3 * (size_t)nn);
    if (rc == XT_OK) rc = h->xconn.upload(xconn, sizeof(int) * 8 * (size_t)ne);
    if (rc == XT_OK) rc = h->fixed.upload(fx.data(), sizeof(int) * p);
    if (rc == XT_OK) rc = h->Ntab.upload(N, sizeof(N));
    if (rc == XT_OK) rc = h->dNtab.upload(dN, sizeof(dN));
    if (rc == XT_OK) rc = h->ke.reserve(sizeof(double) * 576 * (size_t)ne);
    if (rc == XT_OK) rc = h->me.reserve(sizeof(double) * 576 * (size_t)ne);
    for (DevBuf *t : {&h->Nx, &h->Ny, &h->Nz})
        if (rc == XT_OK) rc = t->reserve(sizeof(double) * 64 * (size_t)ne);
    if (rc == XT_OK) rc = build_pattern(p, ne, 24, dofs, false, h->K);
    if (rc == XT_OK) rc = build_pattern(p, ne, 24, dofs, true, h->M);
    if (rc == XT_OK) {   // per-DOF element-vector entries for vv, elements in order
        std::vector<int> vptr(p + 1, 0), vidx((size_t)ne * 24);
        for (size_t k = 0; k < (size_t)ne * 24; ++k) vptr[dofs[k] + 1]++;
        for (int i = 0; i < p; ++i) vptr[i + 1] += vptr[i];
        std::vector<int> fill(vptr.begin(), vptr.end() - 1);
        for (size_t k = 0; k < (size_t)ne * 24; ++k) vidx[fill[dofs[k]]++] = (int)k;
        rc = h->vptr.upload(vptr.data(), sizeof(int) * vptr.size());
        if (rc == XT_OK) rc = h->vidx.upload(vidx.data(), sizeof(int) * vidx.size());
        if (rc == XT_OK) rc = h->ve.reserve(sizeof(double) * 24 * (size_t)ne);
    }
    if (rc == XT_OK && hipDeviceSynchronize() != hipSuccess) rc = XT_E_HIP;
    if (rc != XT_OK) { delete h; return rc; }
    *out = h;
    return XT_OK;
}

void xt_asm_destroy(xt_asm h) { delete h; }

int xt_asm_pattern(xt_asm h, int which, long long *nnz, int *rowptr, int *colidx) {
    XT_ARG(h && (which >= 0 && which <= 2) && nnz, "asm_pattern arguments");
    XT_ARG(which < 2 || h->thermal, "asm_pattern: thermal matrices not built (xt_asm_thermal)");
    const Pattern &P = which == 0 ? h->K : which == 1 ? h->M : h->TP;
    *nnz = P.nnz;
    if (rowptr) std::copy(P.rp.begin(), P.rp.end(), rowptr);
    if (colidx) std::copy(P.ci.begin(), P.ci.end(), colidx);
    return XT_OK;
}

int xt_asm_run(xt_asm h, const double *ym, const double *nu, double rho, void *stream) {
    XT_ARG(h && ym && nu, "asm_run arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(element_kernel, dim3(h->ne), dim3(64), 0, s, h->ne, h->xcoord.as<double>(),
                       h->xconn.as<int>(), ym, nu, rho, h->Ntab.as<double>(),
                       h->dNtab.as<double>(), h->ke.as<double>(), h->me.as<double>(),
                       h->Nx.as<double>(), h->Ny.as<double>(), h->Nz.as<double>());
    XT_HIP(hipGetLastError());
    for (int t = 0; t < 2; ++t) {
        Pattern &P = t == 0 ? h->K : h->M;
        if (!P.nnz) continue;
        hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(P.nnz, 256)), dim3(256), 0, s, P.nnz,
                           P.row.as<int>(), P.col.as<int>(), P.cptr.as<long long>(),
                           P.cidx.as<int>(), t == 0 ? h->ke.as<double>() : h->me.as<double>(),
                           h->fixed.as<int>(), P.val.as<double>());
        XT_HIP(hipGetLastError());
    }
    return XT_OK;
}

int xt_asm_values(xt_asm h, int which, const double **vals) {
    XT_ARG(h && vals && (which >= 0 && which <= 8), "asm_values arguments");
    XT_ARG(which < 5 || h->thermal, "asm_values: thermal matrices not built (xt_asm_thermal)");
    const DevBuf *b = which == 0 ? &h->K.val : which == 1 ? &h->M.val
                    : which == 2 ? &h->Nx : which == 3 ? &h->Ny : which == 4 ? &h->Nz
                    : &h->tval[which - 5];
    *vals = b->as<double>();
    return XT_OK;
}

int xt_asm_copy(xt_asm h, int which, double *dst, void *stream) {
    const double *src = nullptr;
    XT_TRY(xt_asm_values(h, which, &src));
    XT_ARG(dst, "asm_copy destination");
    const size_t n = which == 0 ? (size_t)h->K.nnz : which == 1 ? (size_t)h->M.nnz
                   : which >= 5 ? (size_t)h->TP.nnz : (size_t)h->ne * 64;
    XT_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice,
                          (hipStream_t)stream));
    return XT_OK;
}

int xt_asm_device_pattern(xt_asm h, int which, const int **col, const int **row) {
    XT_ARG(h && (which >= 0 && which <= 2), "asm_device_pattern arguments");
    XT_ARG(which < 2 || h->thermal, "asm_device_pattern: thermal matrices not built");
    const Pattern &P = which == 0 ? h->K : which == 1 ? h->M : h->TP;
    if (col) *col = P.col.as<int>();
    if (row) *row = P.row.as<int>();
    return XT_OK;
}

int xt_asm_thermal(xt_asm h, const int *tfixed, double kappa, double cp, double rho,
                   void *stream) {
    XT_ARG(h && tfixed, "asm_thermal arguments");
    XT_ARG((long long)h->ne * 64 < (1LL << 31), "too many elements for 32-bit contribution indices");
    hipStream_t s = (hipStream_t)stream;
    if (!h->thermal) {
        std::vector<int> xc((size_t)h->ne * 8), nodes((size_t)h->ne * 8);
        XT_HIP(hipMemcpy(xc.data(), h->xconn.as<int>(), sizeof(int) * xc.size(),
                         hipMemcpyDeviceToHost));
        for (size_t k = 0; k < xc.size(); ++k) nodes[k] = xc[k] - 1;
        XT_TRY(build_pattern(h->nn, h->ne, 8, nodes, false, h->TP));
        XT_TRY(h->kt.reserve(sizeof(double) * 64 * (size_t)h->ne));
        XT_TRY(h->mt.reserve(sizeof(double) * 64 * (size_t)h->ne));
        XT_TRY(h->tfixed.reserve(sizeof(int) * (size_t)h->nn));
        for (DevBuf &v : h->tval) XT_TRY(v.reserve(sizeof(double) * (size_t)h->TP.nnz));
        h->thermal = true;
    }
    std::vector<int> fx(h->nn);
    for (int i = 0; i < h->nn; ++i) fx[i] = tfixed[i] != 0;
    XT_HIP(hipMemcpyAsync(h->tfixed.as<int>(), fx.data(), sizeof(int) * fx.size(),
                          hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(thermal_element_kernel, dim3(h->ne), dim3(64), 0, s, h->ne,
                       h->xcoord.as<double>(), h->xconn.as<int>(), kappa, rho * cp,
                       h->Ntab.as<double>(), h->dNtab.as<double>(), h->kt.as<double>(),
                       h->mt.as<double>());
    XT_HIP(hipGetLastError());
    const Pattern &P = h->TP;
    for (int t = 0; t < 4; ++t) {
        hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(P.nnz, 256)), dim3(256), 0, s, P.nnz,
                           P.row.as<int>(), P.col.as<int>(), P.cptr.as<long long>(),
                           P.cidx.as<int>(), (t & 1) ? h->mt.as<double>() : h->kt.as<double>(),
                           t < 2 ? h->tfixed.as<int>() : nullptr, h->tval[t].as<double>());
        XT_HIP(hipGetLastError());
    }
    // the host copy of the mask must outlive the asynchronous upload
    XT_HIP(hipStreamSynchronize(s));
    return XT_OK;
}

int xt_asm_thermo_load(xt_asm h, const double *T, long long tstride, long long toff, double Tref,
                       double alpha, double ym0, double nu0, double *ym, double *nu, double *vv,
                       void *stream) {
    XT_ARG(h && T && ym && nu && vv && tstride > 0 && toff >= 0 && toff < tstride,
           "asm_thermo_load arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(thermo_load_kernel, dim3(h->ne), dim3(64), 0, s, h->ne,
                       h->xcoord.as<double>(), h->xconn.as<int>(), T, tstride, toff, Tref, alpha,
                       ym0, nu0, h->Ntab.as<double>(), h->dNtab.as<double>(), ym, nu,
                       h->ve.as<double>());
    XT_HIP(hipGetLastError());
    hipLaunchKernelGGL(dof_gather_kernel, dim3(ceil_div(h->p, 256)), dim3(256), 0, s, h->p,
                       h->vptr.as<int>(), h->vidx.as<int>(), h->ve.as<double>(), vv);
    XT_HIP(hipGetLastError());
    return XT_OK;
}

}  // extern "C"
