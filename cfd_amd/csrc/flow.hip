// Flow kernels of the explicit SUPG step.
//
//  k_deltat / k_deltat_clip  subrutinas.f90:155-218   element dt, global min, 10x clip
//  k_estab, k_estab_patch    subrutinas.f90:332-446   tau_SUGN1..3 and shock delta
//  k_element_rhs             calcRHS.f90:36-151 (+ FUENTE subrutinas.f90:1060-1091)
//  k_gather_rhs              the atomic scatters of calcRHS/FUENTE as a gather
//  k_stage                   gather + RK stage update + primitives + BCs
//                            (subrutinas.f90:679-826)
//
// Element work is written per slot (esup order) instead of scattered with
// atomics; each node then sums its own contiguous slot segment, calcRHS
// contributions first and FUENTE contributions after, in ascending element
// order -- the order of the reference's sequential scatter.  Everything is
// deterministic run to run.
//
// Floating point follows the reference expression by expression: this file
// is compiled with -ffp-contract=off, x**2 is x*x, x**1.5d0 / x**(-.5d0) are
// correctly rounded pow_3_2 / pow_m1_2 (crmath.hpp), x**.5d0 is sqrt().
#include "cfd_internal.hpp"
#include "crmath.hpp"
#include "reduce.hpp"

namespace cfd {
namespace {

// ---------------------------------------------------------------------------
// deltat
// ---------------------------------------------------------------------------
// One element's time step (subrutinas.f90:178-207) from its nodes' T and
// the magnitudes of their velocity components relative to the mesh
__device__ __forceinline__ double deltat_elem(const double (&Tn)[3], const double (&vu)[3],
                                              const double (&vv)[3], double area, double t_inf,
                                              double fsafe) {
    double vumax = 0.0, vvmax = 0.0;
    const double t_iel = (Tn[0] + Tn[1] + Tn[2]) / 3.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (vu[i] > vumax) vumax = vu[i];
        if (vv[i] > vvmax) vvmax = vv[i];
    }
    const double hh = sqrt(2.0 * area);
    const double vel = sqrt(vumax * vumax + vvmax * vvmax);
    const double smu = 110.0;
    const double fmu = 0.017 * pow_3_2(t_iel / t_inf) * (t_inf + smu) / (t_iel + smu);
    const double et = fmu;
    const double pe = (vel * hh) / (2.0 * et);
    const double alpha = pe / 3.0 < 1.0 ? pe / 3.0 : 1.0;
    const double deltatu = 1.0 / (4.0 * et / (hh * hh) + alpha * vel / hh);
    const double deltatc = 1.0 / (4.0 * et / (hh * hh));
    return fsafe / (1.0 / deltatc + 1.0 / deltatu);
}

__global__ void __launch_bounds__(kBlock)
k_deltat(const int* __restrict__ inpoel, const double* __restrict__ area, const double* __restrict__ T,
         const double* __restrict__ VX, const double* __restrict__ VY, const double* __restrict__ WX,
         const double* __restrict__ WY, double* __restrict__ DT, int nelem, double t_inf,
         double fsafe, double* __restrict__ partials, const int* __restrict__ chunk,
         const int* __restrict__ chunk_case, const Consts* __restrict__ cases) {
    // grid-stride over all elements, or (chunk) block b over [chunk[b], chunk[b+1])
    // with (cases) the constants of the chunk's case
    if (chunk && cases) {
        const Consts& q = cases[chunk_case[blockIdx.x]];
        t_inf = q.t_inf;
        fsafe = q.fsafe;
    }
    // grid-stride: logical blocks in XCD order (neighbouring elements' nodes
    // mostly in one L2); the min partial goes to the logical block
    const int lb = chunk ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
    const int beg = chunk ? chunk[blockIdx.x] + (int)threadIdx.x
                          : (int)(lb * blockDim.x + threadIdx.x);
    const int end = chunk ? chunk[blockIdx.x + 1] : nelem;
    const int step = chunk ? (int)blockDim.x : (int)(gridDim.x * blockDim.x);
    double dmin = 1.0e20;
    for (int e = beg; e < end; e += step) {
        const int* ip = inpoel + 3 * e;
        const int pp[3] = {ip[0] - 1, ip[1] - 1, ip[2] - 1};
        double Tn[3], vu[3], vv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Tn[i] = T[pp[i]];
            vu[i] = fabs(VX[pp[i]] - WX[pp[i]]);
            vv[i] = fabs(VY[pp[i]] - WY[pp[i]]);
        }
        const double dtelem = deltat_elem(Tn, vu, vv, area[e], t_inf, fsafe);
        DT[e] = dtelem;
        if (dtelem < dmin) dmin = dtelem;
    }
    block_partial(dmin, OpMin(), partials, lb);
}

// deltat in patch order: one block per patch, the patch's node values (T,
// |VX - WX|, |VY - WY|) staged in LDS, each element's DT from deltat_elem
// (the same values as k_deltat), one min partial per patch
struct DtNode {
    double T, vu, vv;
};
__global__ void __launch_bounds__(512)
k_deltat_patch(const int* __restrict__ node_base, const int* __restrict__ elem_base,
               const int* __restrict__ pnode, const uint32_t* __restrict__ loc,
               const int* __restrict__ perm, int npatch, const double* __restrict__ area,
               const double* __restrict__ T, const double* __restrict__ VX,
               const double* __restrict__ VY, const double* __restrict__ WX,
               const double* __restrict__ WY, double* __restrict__ DT, double t_inf, double fsafe,
               double* __restrict__ partials) {
    extern __shared__ double4 smem4[];
    DtNode* sN = reinterpret_cast<DtNode*>(smem4);
    const int p = xcd_block(blockIdx.x, npatch);  // grid = npatch: logical patch
    const int nb = node_base[p], nl = node_base[p + 1] - nb;
    for (int l = threadIdx.x; l < nl; l += blockDim.x) {
        const int g = pnode[nb + l] - 1;
        // null WX/WY: a mesh at rest (W = 0, and v - 0 = v exactly)
        sN[l] = DtNode{T[g], fabs(VX[g] - (WX ? WX[g] : 0.0)), fabs(VY[g] - (WY ? WY[g] : 0.0))};
    }
    __syncthreads();
    const long long e0 = elem_base[p];
    const int nep = elem_base[p + 1] - elem_base[p];
    double dmin = 1.0e20;
    for (int q = threadIdx.x; q < nep; q += blockDim.x) {
        const long long e = e0 + q;
        const uint64_t w = (uint64_t)loc[2 * e] | ((uint64_t)loc[2 * e + 1] << 32);
        const int ln[3] = {(int)(w & 1023u), (int)((w >> 10) & 1023u), (int)((w >> 20) & 1023u)};
        double Tn[3], vu[3], vv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const DtNode& nd = sN[ln[i]];
            Tn[i] = nd.T;
            vu[i] = nd.vu;
            vv[i] = nd.vv;
        }
        const int eo = perm[e];
        const double dtelem = deltat_elem(Tn, vu, vv, area[eo], t_inf, fsafe);
        DT[eo] = dtelem;
        if (dtelem < dmin) dmin = dtelem;
    }
    block_partial(dmin, OpMin(), partials, p);
}

__global__ void k_deltat_clip(double* __restrict__ DT, const double* __restrict__ dtmin, int nelem) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    double cota = 10.0 * (*dtmin);
    if (DT[e] > cota) DT[e] = cota;
}

// ---------------------------------------------------------------------------
// ESTAB: k_estab in element order (the FP64 outputs SHOC, T_SUGN1..3 per
// element); k_estab_patch (below, with the patch structures) in patch order
// for the fused paths.  Both run estab_elem, so both give the same values.
// ---------------------------------------------------------------------------
// The element's SUPG parameters from its three nodes' values (node k of the
// element in its connectivity order): shock capturing sh and tau1..3.
__device__ inline void estab_elem(const double DNX[3], const double DNY[3], const double rho[3],
                                  const double Tn[3], const double vxn[3], const double vyn[3],
                                  const double wxn[3], const double wyn[3], const double gmn[3],
                                  double fr, double tinf, double rinf, double dtmin, double& sh,
                                  double& rrr, double& s2, double& s3) {
    double gm = (gmn[0] + gmn[1] + gmn[2]) / 3.0;
    double tau = 0.0, h_rgne = 0.0, h_rgn = 0.0, h_jgn = 0.0;
    double r1 = rho[0], r2 = rho[1], r3 = rho[2];
    double rho_elem = (r1 + r2 + r3) / 3.0;
    double vx = (vxn[0] + vxn[1] + vxn[2]) / 3.0;
    double vy = (vyn[0] + vyn[1] + vyn[2]) / 3.0;
    double wx = (wxn[0] + wxn[1] + wxn[2]) / 3.0;
    double wy = (wyn[0] + wyn[1] + wyn[2]) / 3.0;
    vx = vx - wx;
    vy = vy - wy;
    double vel2 = sqrt(vx * vx + vy * vy);
    double drx = r1 * DNX[0] + r2 * DNX[1] + r3 * DNX[2];
    double dry = r1 * DNY[0] + r2 * DNY[1] + r3 * DNY[2];
    double dr2 = sqrt(drx * drx + dry * dry) + 1.0e-20;
    double T1 = Tn[0], T2 = Tn[1], T3 = Tn[2];
    double dtx = T1 * DNX[0] + T2 * DNX[1] + T3 * DNX[2];
    double dty = T1 * DNY[0] + T2 * DNY[1] + T3 * DNY[2];
    double dt2 = sqrt(dtx * dtx + dty * dty) + 1.0e-20;
    // "velocity gradient" of an element constant: the rounding residue of
    // sum(dN) sets the direction (reproduced as written)
    double dux = vel2 * DNX[0] + vel2 * DNX[1] + vel2 * DNX[2];
    double duy = vel2 * DNY[0] + vel2 * DNY[1] + vel2 * DNY[2];
    double du2 = sqrt(dux * dux + duy * duy) + 1.0e-20;
    double rtx = dtx / dt2, rty = dty / dt2;
    double rjx = drx / dr2, rjy = dry / dr2;
    double rux = dux / du2, ruy = duy / du2;
    double temp = (T1 + T2 + T3) / 3.0;
    double cc = sqrt(gm * fr * temp);
    double smu = 110.0;
    double fmu = 0.017 * pow_3_2(temp / tinf) * (tinf + smu) / (temp + smu);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double term_1 = fabs(vx * DNX[i] + vy * DNY[i]);
        double term_2 = fabs(rjx * DNX[i] + rjy * DNY[i]);
        double h_rgn1 = fabs(rtx * DNX[i] + rty * DNY[i]);
        double h_rgn2 = fabs(rux * DNX[i] + ruy * DNY[i]);
        tau = tau + term_1 + term_2 * cc;
        h_rgne = h_rgne + h_rgn1;
        h_rgn = h_rgn + h_rgn2;
        h_jgn = h_jgn + term_2;
    }
    tau = 1.0 / tau;
    h_rgne = 2.0 / h_rgne;
    h_rgn = 2.0 / h_rgn;
    if (h_rgn > 10.0) h_rgn = 0.0;
    h_jgn = 2.0 / h_jgn;
    if (h_jgn > 10.0) h_jgn = 0.0;
    double tr1 = dr2 * h_jgn / rho_elem;
    double zzz = h_jgn / (2.0 * cc);
    sh = (tr1 + tr1 * tr1) * 0.5 * (cc * cc) * zzz;
    double resumen = 1.0 / (tau * tau) + (2.0 / dtmin) * (2.0 / dtmin);
    rrr = pow_m1_2(resumen);
    s2 = rrr;
    s3 = rrr;
    if (fmu != 0.0) {
        double ts3 = (h_rgn * h_rgn) / (4.0 * fmu / rinf);
        double ts3e = (h_rgne * h_rgne) / (4.0 * fmu / rinf);
        s2 = pow_m1_2(resumen + 1.0 / (ts3 * ts3));
        s3 = pow_m1_2(resumen + 1.0 / (ts3e * ts3e));
    }
}

// Element order: the FP64 outputs SHOC, T_SUGN1..3 per element
__global__ void __launch_bounds__(kBlock)
k_estab(const int* __restrict__ inpoel, const double* __restrict__ U, const double* __restrict__ T,
        const double* __restrict__ GAMM, const double* __restrict__ VEL_X,
        const double* __restrict__ VEL_Y, const double* __restrict__ W_X,
        const double* __restrict__ W_Y, const double* __restrict__ dNx,
        const double* __restrict__ dNy, int nelem, double fr, double tinf, double rinf,
        const double* __restrict__ dtminp, double* __restrict__ shoc, double* __restrict__ t1,
        double* __restrict__ t2, double* __restrict__ t3) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    const int n[3] = {inpoel[3 * e] - 1, inpoel[3 * e + 1] - 1, inpoel[3 * e + 2] - 1};
    const double DNX[3] = {dNx[3 * e], dNx[3 * e + 1], dNx[3 * e + 2]};
    const double DNY[3] = {dNy[3 * e], dNy[3 * e + 1], dNy[3 * e + 2]};
    double rho[3], Tn[3], vxn[3], vyn[3], wxn[3], wyn[3], gmn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rho[k] = U[4 * n[k]];
        Tn[k] = T[n[k]];
        vxn[k] = VEL_X[n[k]];
        vyn[k] = VEL_Y[n[k]];
        wxn[k] = W_X[n[k]];
        wyn[k] = W_Y[n[k]];
        gmn[k] = GAMM[n[k]];
    }
    double sh, rrr, s2, s3;
    estab_elem(DNX, DNY, rho, Tn, vxn, vyn, wxn, wyn, gmn, fr, tinf, rinf, *dtminp, sh, rrr, s2,
               s3);
    shoc[e] = sh;
    t1[e] = rrr;
    t2[e] = s2;
    t3[e] = s3;
}

// SHOC, T_SUGN1..3 from the patch-order copies (fused paths): element e's
// values sit at patch position inv_perm[e]
template <class R>
__global__ void k_estab_widen(const R* __restrict__ nt, const int* __restrict__ inv_perm,
                              double inv_cte, int nelem, double* __restrict__ shoc,
                              double* __restrict__ t1, double* __restrict__ t2,
                              double* __restrict__ t3) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    const long long q = inv_perm[e];
    shoc[e] = (double)nt[4 * q] * inv_cte;
    t1[e] = nt[4 * q + 1];
    t2[e] = nt[4 * q + 2];
    t3[e] = nt[4 * q + 3];
}

// ---------------------------------------------------------------------------
// SUPG element residual
// ---------------------------------------------------------------------------
struct ElemConsts {
    double g, Cv, lambda_ref, mu_ref, T_inf, cte;
    int viscous;
};

__device__ inline double4 ld4(const double* p) { return *reinterpret_cast<const double4*>(p); }
__device__ inline void st4(double* p, double a, double b, double c, double d) {
    *reinterpret_cast<double4*>(p) = make_double4(a, b, c, d);
}

// One element: calcRHS.f90:37-141 into rt[node][eq], unscaled order kept.
template <bool kTheta>
__device__ inline void supg_element(const ElemConsts& k, const double (&Uv)[3][4],
                                   const double (&Th)[3][4], const double (&Tn)[3],
                                   const double (&Nx)[3], const double (&Ny)[3], double tau1,
                                   double tau2, double tau3, double shoc, double area, double dtl,
                                   double (&rt)[3][4]) {
    const double g = k.g;
    double Ux[4], Uy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        Ux[q] = Uv[0][q] * Nx[0] + Uv[1][q] * Nx[1] + Uv[2][q] * Nx[2];
        Uy[q] = Uv[0][q] * Ny[0] + Uv[1][q] * Ny[1] + Uv[2][q] * Ny[2];
    }
    const double tau[3] = {tau1, tau2, tau3};
    double nu = shoc * k.cte;
    double T_avg = (Tn[0] + Tn[1] + Tn[2]) / 3.0;
    double pw = pow_3_2(T_avg / k.T_inf);
    double mu = k.mu_ref * pw * (k.T_inf + 110.0) / (T_avg + 110.0);
    double lambda = k.lambda_ref * pw * (k.T_inf + 194.0) / (T_avg + 194.0);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) rt[j][q] = 0.0;
    // quadrature points in a rolled loop: bounds the live registers of the
    // FP64 element (occupancy) without changing any operation
#pragma unroll 1
    for (int kq = 0; kq < 3; ++kq) {
        // N(j,k) = 0 for j == k, 1/2 otherwise (midpoint of the edge opposite node k)
        double Nw[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) Nw[j] = (j == kq) ? 0.0 : 0.5;
        double Uk[4], th[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Uk[q] = Nw[0] * Uv[0][q] + Nw[1] * Uv[1][q] + Nw[2] * Uv[2][q];
            if (kTheta)
                th[q] = Nw[0] * Th[0][q] + Nw[1] * Th[1][q] + Nw[2] * Th[2][q];
            else
                th[q] = 0.0;
        }
        double rho = Uk[0];
        double v1 = Uk[1] / rho;
        double v2 = Uk[2] / rho;
        double en = Uk[3] / rho;
        double V_sq = v1 * v1 + v2 * v2;
        double A[4];
        A[0] = Ux[1] + Uy[2];
        A[1] = 0.5 * Ux[0] * (V_sq * (g - 1.0) - 2.0 * (v1 * v1)) - Ux[1] * v1 * (g - 3.0) -
               Ux[2] * v2 * (g - 1.0) + Ux[3] * (g - 1.0) - Uy[0] * v1 * v2 + Uy[1] * v2 + Uy[2] * v1;
        A[2] = -(Ux[0] * v1 * v2) + Ux[1] * v2 + Ux[2] * v1 +
               0.5 * Uy[0] * (V_sq * (g - 1.0) - 2.0 * (v2 * v2)) - Uy[1] * v1 * (g - 1.0) -
               Uy[2] * v2 * (g - 3.0) + Uy[3] * (g - 1.0);
        A[3] = Ux[0] * v1 * (V_sq * (g - 1.0) - en * g) -
               0.5 * Ux[1] * (V_sq * (g - 1.0) - 2.0 * en * g + 2.0 * (v1 * v1) * (g - 1.0)) -
               Ux[2] * v1 * v2 * (g - 1.0) + Ux[3] * g * v1 + Uy[0] * v2 * (V_sq * (g - 1.0) - en * g) -
               Uy[1] * v1 * v2 * (g - 1.0) -
               0.5 * Uy[2] * (V_sq * (g - 1.0) - 2.0 * en * g + 2.0 * (v2 * v2) * (g - 1.0)) +
               Uy[3] * g * v2;
        double a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = th[q] + A[q];
        double B1[4], B2[4];
        B1[0] = a[1];
        B1[1] = v1 * (-g + 3.0) * a[1] - v2 * (g - 1.0) * a[2] + (g - 1.0) * a[3] +
                (0.5 * V_sq * (g - 1.0) - (v1 * v1)) * a[0];
        B1[2] = -(v1 * v2 * a[0]) + v1 * a[2] + v2 * a[1];
        B1[3] = g * v1 * a[3] - v1 * v2 * (g - 1.0) * a[2] + v1 * (V_sq * (g - 1.0) - en * g) * a[0] +
                (-(0.5 * V_sq * (g - 1.0)) + en * g - (v1 * v1) * (g - 1.0)) * a[1];
        B2[0] = a[2];
        B2[1] = -(v1 * v2 * a[0]) + v1 * a[2] + v2 * a[1];
        B2[2] = -(v1 * (g - 1.0) * a[1]) + v2 * (-g + 3.0) * a[2] + (g - 1.0) * a[3] +
                (0.5 * V_sq * (g - 1.0) - (v2 * v2)) * a[0];
        B2[3] = g * v2 * a[3] - v1 * v2 * (g - 1.0) * a[1] + v2 * (V_sq * (g - 1.0) - en * g) * a[0] +
                (-(0.5 * V_sq * (g - 1.0)) + en * g - (v2 * v2) * (g - 1.0)) * a[2];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                rt[j][q] = rt[j][q] + Nw[j] * A[q] + tau[j] * (Nx[j] * B1[q] + Ny[j] * B2[q]) +
                           nu * (Nx[j] * Ux[q] + Ny[j] * Uy[q]);
        if (k.viscous) {
            const double Cv = k.Cv;
            double K1[4], K2[4];
            K1[1] = (2.0 / 3.0) * mu * (-(2.0 * Ux[0] * v1) + 2.0 * Ux[1] + Uy[0] * v2 - Uy[2]) / rho;
            K1[2] = mu * (-(Ux[0] * v2) + Ux[2] - Uy[0] * v1 + Uy[1]) / rho;
            K1[3] = (1.0 / 3.0) *
                    (Cv * mu * (-(Uy[0] * v1 * v2) + 3.0 * Uy[1] * v2 - 2.0 * Uy[2] * v1) -
                     Ux[0] * (Cv * mu * (3.0 * V_sq + (v1 * v1)) - 3.0 * lambda * (V_sq - en)) +
                     Ux[1] * v1 * (4.0 * Cv * mu - 3.0 * lambda) + 3.0 * Ux[2] * v2 * (Cv * mu - lambda) +
                     3.0 * Ux[3] * lambda) /
                    (Cv * rho);
            K2[1] = mu * (-(Ux[0] * v2) + Ux[2] - Uy[0] * v1 + Uy[1]) / rho;
            K2[2] = (2.0 / 3.0) * mu * (Ux[0] * v1 - Ux[1] - 2.0 * Uy[0] * v2 + 2.0 * Uy[2]) / rho;
            K2[3] = (1.0 / 3.0) *
                    (Cv * mu * (-(Ux[0] * v1 * v2) - 2.0 * Ux[1] * v2 + 3.0 * Ux[2] * v1) -
                     Uy[0] * (Cv * mu * (3.0 * V_sq + (v2 * v2)) - 3.0 * lambda * (V_sq - en)) +
                     3.0 * Uy[1] * v1 * (Cv * mu - lambda) + Uy[2] * v2 * (4.0 * Cv * mu - 3.0 * lambda) +
                     3.0 * Uy[3] * lambda) /
                    (Cv * rho);
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int q = 1; q < 4; ++q) rt[j][q] = rt[j][q] + (Nx[j] * K1[q] + Ny[j] * K2[q]);
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) rt[j][q] = rt[j][q] * area * dtl / 3.0;
}

// FUENTE for one element (subrutinas.f90:1064-1078)
__device__ inline void fuente_element(const double (&Uv)[3][4], const double (&Nx)[3],
                                      const double (&Ny)[3], const double (&w_x)[3],
                                      const double (&w_y)[3], double area, double dtl,
                                      double (&rf)[3][4]) {
    // sp(j,k): weight of local node j at quadrature point k
    constexpr double sp[3][3] = {{0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}, {0.0, 0.5, 0.5}};
    double Ux[4], Uy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        Ux[q] = Uv[0][q] * Nx[0] + Uv[1][q] * Nx[1] + Uv[2][q] * Nx[2];
        Uy[q] = Uv[0][q] * Ny[0] + Uv[1][q] * Ny[1] + Uv[2][q] * Ny[2];
    }
    double AR = area * dtl / 3.0;
    double wx[3], wy[3];
#pragma unroll
    for (int kq = 0; kq < 3; ++kq) {
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            sx = sx + sp[j][kq] * w_x[j];
            sy = sy + sp[j][kq] * w_y[j];
        }
        wx[kq] = sx;
        wy[kq] = sy;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            rf[j][q] = -(AR * (sp[j][0] * (Ux[q] * wx[0] + Uy[q] * wy[0]) +
                               sp[j][1] * (Ux[q] * wx[1] + Uy[q] * wy[1]) +
                               sp[j][2] * (Ux[q] * wx[2] + Uy[q] * wy[2])));
}

// One thread per element.  Connectivity, slot positions and the shape
// gradients of the block's elements are staged through LDS with coalesced
// loads; U/T of the three nodes are gathered (32 B + 8 B per node).
template <bool kCalc, bool kTheta, bool kFuente>
__global__ void __launch_bounds__(kBlock)
k_element_rhs(ElemConsts kc, const int* __restrict__ inpoel, const int* __restrict__ epos,
              const double* __restrict__ U, const double* __restrict__ theta,
              const double* __restrict__ T, const double* __restrict__ dNx,
              const double* __restrict__ dNy, const double* __restrict__ area,
              const double* __restrict__ shoc, const double* __restrict__ dtl,
              const double* __restrict__ t1, const double* __restrict__ t2,
              const double* __restrict__ t3, const double* __restrict__ WX,
              const double* __restrict__ WY, double* __restrict__ c_calc,
              double* __restrict__ c_fuente, int nelem) {
    __shared__ int s_in[3 * kBlock];
    __shared__ int s_pos[3 * kBlock];
    __shared__ double s_nx[3 * kBlock];
    __shared__ double s_ny[3 * kBlock];
    const int e0 = blockIdx.x * kBlock;
    const int nloc = min(kBlock, nelem - e0);
    for (int q = threadIdx.x; q < 3 * nloc; q += kBlock) {
        s_in[q] = inpoel[3 * (long long)e0 + q];
        s_pos[q] = epos[3 * (long long)e0 + q];
        s_nx[q] = dNx[3 * (long long)e0 + q];
        s_ny[q] = dNy[3 * (long long)e0 + q];
    }
    __syncthreads();
    if ((int)threadIdx.x >= nloc) return;
    const int e = e0 + threadIdx.x;
    const int l = 3 * threadIdx.x;
    int p[3] = {s_in[l] - 1, s_in[l + 1] - 1, s_in[l + 2] - 1};
    double Nx[3] = {s_nx[l], s_nx[l + 1], s_nx[l + 2]};
    double Ny[3] = {s_ny[l], s_ny[l + 1], s_ny[l + 2]};
    double Uv[3][4], Th[3][4], Tn[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double4 u = ld4(U + 4 * (long long)p[j]);
        Uv[j][0] = u.x; Uv[j][1] = u.y; Uv[j][2] = u.z; Uv[j][3] = u.w;
        Tn[j] = kCalc ? T[p[j]] : 0.0;
        if (kTheta) {
            double4 t = ld4(theta + 4 * (long long)p[j]);
            Th[j][0] = t.x; Th[j][1] = t.y; Th[j][2] = t.z; Th[j][3] = t.w;
        } else {
            Th[j][0] = Th[j][1] = Th[j][2] = Th[j][3] = 0.0;
        }
    }
    const double ar = area[e], dt = dtl[e];
    if (kCalc) {
        double rt[3][4];
        supg_element<kTheta>(kc, Uv, Th, Tn, Nx, Ny, t1[e], t2[e], t3[e], shoc[e], ar, dt, rt);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            st4(c_calc + 4 * (long long)s_pos[l + j], rt[j][0], rt[j][1], rt[j][2], rt[j][3]);
    }
    if (kFuente) {
        double w_x[3] = {WX[p[0]], WX[p[1]], WX[p[2]]};
        double w_y[3] = {WY[p[0]], WY[p[1]], WY[p[2]]};
        double rf[3][4];
        fuente_element(Uv, Nx, Ny, w_x, w_y, ar, dt, rf);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            st4(c_fuente + 4 * (long long)s_pos[l + j], rf[j][0], rf[j][1], rf[j][2], rf[j][3]);
    }
}

// Slot segments of a block of consecutive nodes are one contiguous range of
// the slot buffer.  It is staged through LDS in chunks with coalesced loads
// (consecutive lanes, consecutive slots); each node then adds its own slots
// from LDS in ascending order.  (Strided per-node loads straight from HBM
// re-fetched every line about twice: each wave's segment span was evicted
// from L2 between its iterations.)
constexpr int kChunk = 1024;

template <class CT>  // double4 (FP64 contributions) or float4 (FP32, summed in FP64)
__device__ inline void staged_sum(const CT* __restrict__ src, int s0, int s1, int b, int e, CT* sbuf,
                                  double& r0, double& r1, double& r2, double& r3) {
    for (int c0 = s0; c0 < s1; c0 += kChunk) {
        const int c1 = min(c0 + kChunk, s1);
        __syncthreads();
        for (int q = c0 + (int)threadIdx.x; q < c1; q += (int)blockDim.x) sbuf[q - c0] = src[q];
        __syncthreads();
        const int lo = max(b, c0), hi = min(e, c1);
        for (int s = lo; s < hi; ++s) {
            CT c = sbuf[s - c0];
            r0 = r0 + (double)c.x;
            r1 = r1 + (double)c.y;
            r2 = r2 + (double)c.z;
            r3 = r3 + (double)c.w;
        }
    }
}

// rhs_out(i) = rhs_in(i) + calc slots in order (+ FUENTE slots in order)
template <class CT>
__global__ void __launch_bounds__(kBlock)
k_gather_rhs(const int* __restrict__ esup2, const double* rhs_in, const CT* __restrict__ c_calc,
             const CT* __restrict__ c_fuente, double* rhs_out,
             int npoin) {  // rhs_in may alias rhs_out (same index only)
    __shared__ CT sbuf[kChunk];
    const int i0 = blockIdx.x * kBlock, i1 = min(i0 + kBlock, npoin);
    const int i = i0 + threadIdx.x;
    const bool act = i < npoin;
    const int s0 = esup2[i0], s1 = esup2[i1];
    const int b = act ? esup2[i] : 0, e = act ? esup2[i + 1] : 0;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    if (act && rhs_in) {
        double4 v = ld4(rhs_in + 4 * (long long)i);
        r0 = v.x; r1 = v.y; r2 = v.z; r3 = v.w;
    }
    staged_sum(c_calc, s0, s1, b, e, sbuf, r0, r1, r2, r3);
    if (c_fuente) staged_sum(c_fuente, s0, s1, b, e, sbuf, r0, r1, r2, r3);
    if (act) st4(rhs_out + 4 * (long long)i, r0, r1, r2, r3);
}

// One RK stage for one node, given its summed residual (subrutinas.f90:679-826,
// NGAS = 0).  kFinal = false: only T is stored (the next stage's only nodal
// input besides U); the final stage stores everything.
struct StageDev {
    double rk_fact, fr;
    const double *M, *GAMM, *WX, *WY;
    const int *bcidx, *bcflags;
    const double* bcvals;
    double *U1, *RHS, *T, *VX, *VY, *RHO, *E, *P, *RMACH;
};

template <bool kFinal>
__device__ inline void stage_node(const StageDev& a, int i, double4 u, double r0, double r1,
                                  double r2, double r3) {
    if (kFinal && a.RHS) st4(a.RHS + 4 * (long long)i, r0, r1, r2, r3);
    double f = a.rk_fact / a.M[i];
    double u0 = u.x - f * r0, u1 = u.y - f * r1, u2 = u.z - f * r2, u3 = u.w - f * r3;
    double gamm = a.GAMM[i];
    double rho = u0;
    double vx = u1 / rho;
    double vy = u2 / rho;
    double en = u3 / rho;
    double vel2 = vx * vx + vy * vy;
    double pr = rho * (gamm - 1.0) * (en - 0.5 * vel2);
    double tt = pr / (rho * a.fr);
    int bi = a.bcidx[i];
    if (!kFinal) {  // T after the BCs: only FIX T changes it
        if (bi >= 0 && (a.bcflags[bi] & 8)) tt = a.bcvals[8 * bi + 5];
        a.T[i] = tt;
        return;
    }
    a.P[i] = pr;
    a.RMACH[i] = sqrt(vel2 / (tt * gamm * a.fr));
    if (bi >= 0) {
        int fl = a.bcflags[bi];
        const double* v = a.bcvals + 8 * bi;
        if (fl & 1) { vx = v[0]; vy = v[1]; }                 // FIXVEL
        if (fl & 2) {                                          // NORMALVEL
            double wx = a.WX[i], wy = a.WY[i], nx = v[2], ny = v[3];
            double pp = -ny * (vx - wx) + nx * (vy - wy);
            vx = -ny * pp + wx;
            vy = nx * pp + wy;
        }
        if (fl & 4) rho = v[4];                                // FIX rho
        if (fl & 8) {                                          // FIX T
            double gm = gamm - 1.0;
            tt = v[5];
            en = tt * a.fr / gm + 0.5 * (vx * vx + vy * vy);
        }
    }
    a.RHO[i] = rho;
    a.VX[i] = vx;
    a.VY[i] = vy;
    a.E[i] = en;
    a.T[i] = tt;
    st4(a.U1 + 4 * (long long)i, rho, vx * rho, vy * rho, en * rho);
}

// Gather (calcRHS slots, then FUENTE slots, in esup order) + one RK stage
template <bool kFuente, bool kFinal>
__global__ void __launch_bounds__(kBlock)
k_stage(const int* __restrict__ esup2, const double4* __restrict__ c_calc,
        const double4* __restrict__ c_fuente, const double* __restrict__ U, StageDev a,
        int npoin) {
    __shared__ double4 sbuf[kChunk];
    const int i0 = blockIdx.x * kBlock, i1 = min(i0 + kBlock, npoin);
    const int i = i0 + threadIdx.x;
    const bool act = i < npoin;
    const int s0 = esup2[i0], s1 = esup2[i1];
    const int b = act ? esup2[i] : 0, e = act ? esup2[i + 1] : 0;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    staged_sum(c_calc, s0, s1, b, e, sbuf, r0, r1, r2, r3);
    if (kFuente) staged_sum(c_fuente, s0, s1, b, e, sbuf, r0, r1, r2, r3);
    if (!act) return;
    stage_node<kFinal>(a, i, ld4(U + 4 * (long long)i), r0, r1, r2, r3);
}

// ---------------------------------------------------------------------------
// Patch element residual (same equations as calcRHS.f90:37-141 with theta =
// 0, i.e. as RK calls it), in binary32 (FP32 path) or FP64 (fused FP64 path)
// with contraction; the sums over the three quadrature points are taken once
// per element:
//   Galerkin  sum_k N(j,k) A_k         = (sum_k A_k - A_j) / 2
//   SUPG      tau_j (Nx_j sum_k A1 A_k + Ny_j sum_k A2 A_k)
//   shock     3 nu (Nx_j Ux + Ny_j Uy)
//   viscous   Nx_j sum_k K1_k + Ny_j sum_k K2_k
// The reference's expanded Jacobian entries (calcRHS.f90:58-117) are
// evaluated as Jacobian-vector products of the Euler fluxes at the
// quadrature state (rho, u, v, E):
//   dF1/dU w = (w1, 2u w1 - u^2 w0 + dp, v w1 + u w2 - uv w0, H (w1 - u w0) + u (w3 + dp))
//   dF2/dU w = (w2, v w1 + u w2 - uv w0, 2v w2 - v^2 w0 + dp, H (w2 - v w0) + v (w3 + dp))
//   dp = (g-1)(w3 - u w1 - v w2 + q^2/2 w0),  H = g E - (g-1) q^2/2
// and the viscous fluxes (calcRHS.f90:119-137) through the velocity
// gradients: K1 = (0, txx, txy, u txx + v txy + (lambda/Cv) de/dx), K2 the
// same with y, e = E - q^2/2; both are the reference's polynomials regrouped.
// Gradients and quadrature-point states come from FP64 node values: the
// gradient uses differences to node 1, (U_j - U_1) N_j, so a uniform state
// gives exactly zero whatever the rounding of the shape gradients.
// Contributions are float4 or double4 slots summed in FP64 per node.
// ---------------------------------------------------------------------------
template <class R>
struct ElemConstsP {
    R g, Cv, lambda_ref, mu_ref, T_inf;
    int viscous;
};

template <class R> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

__device__ __forceinline__ float4 make4(float a, float b, float c, float d) {
    return make_float4(a, b, c, d);
}
__device__ __forceinline__ double4 make4(double a, double b, double c, double d) {
    return make_double4(a, b, c, d);
}

// reciprocal: the hardware approximation in binary32, a rounded division in FP64
__device__ __forceinline__ float rcpR(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpR(double x) { return 1.0 / x; }

// One element's residual at its three nodes, calcRHS plus (kFuente) FUENTE,
// into out[j] (node j of the element), in the arithmetic of R.  u0..u2,
// T, WX, WY: the FP64 node values; Nx, Ny: shape gradients; ar =
// area*dtl/3; nu, tau from ESTAB.
template <class R, bool kFuente>
__device__ inline void elem_residual(const ElemConstsP<R>& kc, const double4 u0, const double4 u1,
                                     const double4 u2, const double (&T)[3],
                                     const double (&WX)[3], const double (&WY)[3],
                                     const R (&Nx)[3], const R (&Ny)[3], const R ar, const R nu,
                                     const R (&tau)[3], typename Vec4<R>::type (&out)[3]) {
#pragma clang fp contract(fast)
    const double a0[4] = {u0.x, u0.y, u0.z, u0.w};
    const double a1[4] = {u1.x, u1.y, u1.z, u1.w};
    const double a2[4] = {u2.x, u2.y, u2.z, u2.w};
    // gradients; Q[k] = twice the state at the midpoint opposite node k
    // (velocities and energy are ratios, so the factor 2 cancels there)
    R Ux[4], Uy[4], Q[3][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const R d1 = (R)(a1[q] - a0[q]), d2 = (R)(a2[q] - a0[q]);
        Ux[q] = d1 * Nx[1] + d2 * Nx[2];
        Uy[q] = d1 * Ny[1] + d2 * Ny[2];
        Q[0][q] = (R)(a1[q] + a2[q]);
        Q[1][q] = (R)(a0[q] + a2[q]);
        Q[2][q] = (R)(a0[q] + a1[q]);
    }
    const R g = kc.g, gm1 = g - (R)1;
    R mu = 0, kl = 0;
    if (kc.viscous) {  // Sutherland laws at the element-average temperature
        if constexpr (sizeof(R) == 8) {
            const double Tavg = (T[0] + T[1] + T[2]) / 3.0;
            const double pw = pow_3_2(Tavg / kc.T_inf);
            mu = kc.mu_ref * pw * (kc.T_inf + 110.0) / (Tavg + 110.0);
            kl = kc.lambda_ref * pw * (kc.T_inf + 194.0) / (Tavg + 194.0) / kc.Cv;
        } else {
            const float Tavg = (float)((T[0] + T[1] + T[2]) * (1.0 / 3.0));
            const float x = Tavg / kc.T_inf;
            const float pw = x * __builtin_sqrtf(x);
            mu = kc.mu_ref * pw * (kc.T_inf + 110.0f) * rcpR(Tavg + 110.0f);
            kl = kc.lambda_ref * pw * (kc.T_inf + 194.0f) * rcpR(Tavg + 194.0f) * rcpR(kc.Cv);
        }
    }
    const R mu23 = (R)(2.0 / 3.0) * mu;
    R A[3][4], S1[4] = {0, 0, 0, 0}, S2[4] = {0, 0, 0, 0}, V1[4] = {0, 0, 0, 0},
               V2[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R ir2 = rcpR(Q[k][0]);  // 1 / (2 rho)
        const R u = Q[k][1] * ir2, v = Q[k][2] * ir2, E = Q[k][3] * ir2;
        const R uu = u * u, vv = v * v, uv = u * v, hq = (R)0.5 * (uu + vv);
        const R H = g * E - gm1 * hq;
        // A = dF1/dU Ux + dF2/dU Uy
        const R dpx = gm1 * (Ux[3] - u * Ux[1] - v * Ux[2] + hq * Ux[0]);
        const R dpy = gm1 * (Uy[3] - u * Uy[1] - v * Uy[2] + hq * Uy[0]);
        R* a = A[k];
        a[0] = Ux[1] + Uy[2];
        a[1] = (R)2 * u * Ux[1] - uu * Ux[0] + dpx + v * Uy[1] + u * Uy[2] - uv * Uy[0];
        a[2] = v * Ux[1] + u * Ux[2] - uv * Ux[0] + (R)2 * v * Uy[2] - vv * Uy[0] + dpy;
        a[3] = H * (Ux[1] - u * Ux[0]) + u * (Ux[3] + dpx) + H * (Uy[2] - v * Uy[0]) +
               v * (Uy[3] + dpy);
        // SUPG: dF1/dU a and dF2/dU a
        const R dpa = gm1 * (a[3] - u * a[1] - v * a[2] + hq * a[0]);
        const R c12 = v * a[1] + u * a[2] - uv * a[0], e3 = a[3] + dpa;
        S1[0] += a[1];
        S1[1] += (R)2 * u * a[1] - uu * a[0] + dpa;
        S1[2] += c12;
        S1[3] += H * (a[1] - u * a[0]) + u * e3;
        S2[0] += a[2];
        S2[1] += c12;
        S2[2] += (R)2 * v * a[2] - vv * a[0] + dpa;
        S2[3] += H * (a[2] - v * a[0]) + v * e3;
        if (kc.viscous) {
            const R ir = (R)2 * ir2;
            const R ux = (Ux[1] - u * Ux[0]) * ir, vx = (Ux[2] - v * Ux[0]) * ir;
            const R uy = (Uy[1] - u * Uy[0]) * ir, vy = (Uy[2] - v * Uy[0]) * ir;
            const R txx = mu23 * ((R)2 * ux - vy), tyy = mu23 * ((R)2 * vy - ux);
            const R txy = mu * (uy + vx);
            const R ex = (Ux[3] - E * Ux[0]) * ir - (u * ux + v * vx);
            const R ey = (Uy[3] - E * Uy[0]) * ir - (u * uy + v * vy);
            V1[1] += txx;
            V1[2] += txy;
            V1[3] += u * txx + v * txy + kl * ex;
            V2[1] += txy;
            V2[2] += tyy;
            V2[3] += u * txy + v * tyy + kl * ey;
        }
    }
    const R nu3 = (R)3 * nu, arh = (R)0.5 * ar;
    R SA[4], P1[4], P2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        SA[q] = A[0][q] + A[1][q] + A[2][q];
        P1[q] = nu3 * Ux[q] + V1[q];
        P2[q] = nu3 * Uy[q] + V2[q];
    }
    R r[3][4];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const R x = Nx[j] * (tau[j] * S1[q] + P1[q]) + Ny[j] * (tau[j] * S2[q] + P2[q]);
            r[j][q] = ar * x + arh * (SA[q] - A[j][q]);
        }
    }
    if (kFuente) {
        // subrutinas.f90:1064-1078: -AR sum_k sp(j,k) (Ux wx_k + Uy wy_k)
        const R w_x[3] = {(R)WX[0], (R)WX[1], (R)WX[2]};
        const R w_y[3] = {(R)WY[0], (R)WY[1], (R)WY[2]};
        // sp(:,1) = (1/2,1/2,0), sp(:,2) = (0,1/2,1/2), sp(:,3) = (1/2,0,1/2)
        const R h = (R)0.5;
        const R wx[3] = {h * (w_x[0] + w_x[1]), h * (w_x[1] + w_x[2]), h * (w_x[0] + w_x[2])};
        const R wy[3] = {h * (w_y[0] + w_y[1]), h * (w_y[1] + w_y[2]), h * (w_y[0] + w_y[2])};
        const R sp[3][3] = {{h, 0, h}, {h, h, 0}, {0, h, h}};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                R acc = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc += sp[j][k] * (Ux[q] * wx[k] + Uy[q] * wy[k]);
                r[j][q] += -ar * acc;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) out[j] = make4(r[j][0], r[j][1], r[j][2], r[j][3]);
}

// Fused residual + stage, one patch per block: node values to LDS; each
// thread computes elements of the patch into LDS slots (float4 or double4);
// each node sums its slots in FP64, then interior nodes take the RK stage and
// halo nodes store their partial sum.
struct PatchDev {
    const uint32_t* loc;
    const int *elem_base, *node_base, *n_int, *halo_base, *pnode;
    const uint16_t* lptr;
    double* hpart;
    int nelem, maxn, pe, npatch;
    const int* pcase;      // batches: case of each patch (null: kc and a as given)
    const Consts* cases;
};

// At most 168 VGPRs, so that 3 waves per SIMD stay resident (the LDS of the
// default patch shapes allows 3 blocks of 256 threads per CU): the FP64
// fixed-mesh stages fit without spilling (174 VGPRs would leave 2 waves),
// the FP64 moving-mesh stages spill 24 B per thread and still run faster
// than at 188 VGPRs and 2 waves
template <class R, bool kFuente, bool kFinal>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3)))
k_patch_stage(PatchDev pd, ElemConstsP<R> kc, PatchIn<R> in, StageDev a) {
    using V4 = typename Vec4<R>::type;
    extern __shared__ double4 smem4[];
    const int maxn = pd.maxn, pe = pd.pe;
    double4* sU = smem4;                                          // maxn
    double2* sW = reinterpret_cast<double2*>(sU + maxn);          // maxn (kFuente)
    V4* sC = reinterpret_cast<V4*>(sW + (kFuente ? maxn : 0));    // 3*pe
    double* sT = reinterpret_cast<double*>(sC + 3 * pe);          // maxn
    uint16_t* sP = reinterpret_cast<uint16_t*>(sT + maxn);        // maxn+1
    // one patch per block, or (resident grid) patches blockIdx.x + k gridDim.x
    for (int p0 = blockIdx.x; p0 < pd.npatch; p0 += gridDim.x) {
        // measured: the XCD order helps the FP32 stages (149 -> 143 us per
        // launch at 4.2M triangles) and slows the FP64 ones (204 -> 208 us)
        const int p = sizeof(R) == sizeof(float) && (int)gridDim.x == pd.npatch
                          ? xcd_block(p0, pd.npatch)
                          : p0;
        ElemConstsP<R> kl = kc;
        StageDev al = a;
        if (pd.pcase) {  // the constants of this patch's case, rounded as the host rounds them
            const Consts& q = pd.cases[pd.pcase[p]];
            kl.g = (R)q.gama;
            kl.Cv = (R)q.fcv;
            kl.lambda_ref = (R)q.fk;
            kl.mu_ref = (R)q.fmu;
            kl.T_inf = (R)q.t_inf;
            kl.viscous = q.fmu > 2.2250738585072014e-308 ? 1 : 0;
            al.fr = q.fr;
        }
        const int nb = pd.node_base[p], nl = pd.node_base[p + 1] - nb, nint = pd.n_int[p];
        const long long e0 = pd.elem_base[p];
        const int nep = pd.elem_base[p + 1] - pd.elem_base[p];
        for (int l = threadIdx.x; l < nl; l += blockDim.x) {
            const int g = pd.pnode[nb + l] - 1;
            sU[l] = ld4(in.U + 4 * (long long)g);
            sT[l] = in.T[g];
            if (kFuente) sW[l] = make_double2(in.WX[g], in.WY[g]);
        }
        for (int l = threadIdx.x; l <= nl; l += blockDim.x) sP[l] = pd.lptr[nb + p + l];
        __syncthreads();
        for (int q = threadIdx.x; q < nep; q += blockDim.x) {
            const long long e = e0 + q;
            const uint64_t w = (uint64_t)pd.loc[2 * e] | ((uint64_t)pd.loc[2 * e + 1] << 32);
            const int ln[3] = {(int)(w & 1023u), (int)((w >> 10) & 1023u), (int)((w >> 20) & 1023u)};
            const int ls[3] = {(int)((w >> 30) & 2047u), (int)((w >> 41) & 2047u),
                               (int)((w >> 52) & 2047u)};
            // P1 gradients sum to zero: the first from the other two
            const V4 gN = *reinterpret_cast<const V4*>(in.dN + 4 * e);
            const R Nx[3] = {-(gN.x + gN.y), gN.x, gN.y};
            const R Ny[3] = {-(gN.z + gN.w), gN.z, gN.w};
            const V4 t4 = *reinterpret_cast<const V4*>(in.nt + 4 * e);  // nu, tau1..3
            const R tau[3] = {t4.y, t4.z, t4.w};
            const double T3[3] = {sT[ln[0]], sT[ln[1]], sT[ln[2]]};
            double wx3[3] = {0.0, 0.0, 0.0}, wy3[3] = {0.0, 0.0, 0.0};
            if (kFuente) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    wx3[k] = sW[ln[k]].x;
                    wy3[k] = sW[ln[k]].y;
                }
            }
            V4 out[3];
            elem_residual<R, kFuente>(kl, sU[ln[0]], sU[ln[1]], sU[ln[2]], T3, wx3, wy3, Nx, Ny,
                                      in.fac[e], t4.x, tau, out);
#pragma unroll
            for (int k = 0; k < 3; ++k) sC[ls[k]] = out[k];
        }
        __syncthreads();
        for (int l = threadIdx.x; l < nl; l += blockDim.x) {
            double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
            for (int s = sP[l]; s < sP[l + 1]; ++s) {
                const V4 c = sC[s];
                r0 = r0 + (double)c.x;
                r1 = r1 + (double)c.y;
                r2 = r2 + (double)c.z;
                r3 = r3 + (double)c.w;
            }
            if (l < nint)
                stage_node<kFinal>(al, pd.pnode[nb + l] - 1, sU[l], r0, r1, r2, r3);
            else
                st4(pd.hpart + 4 * (long long)(pd.halo_base[p] + l - nint), r0, r1, r2, r3);
        }
        __syncthreads();  // the next patch overwrites the LDS arrays
    }
}

// Halo nodes: partial sums in ascending patch order, then the stage
template <bool kFinal>
__global__ void __launch_bounds__(kBlock)
k_halo_stage(const int* __restrict__ hnode, const int* __restrict__ hptr,
             const int* __restrict__ hidx, const double* __restrict__ hpart,
             const double* __restrict__ U, StageDev a, int nhalo, const int* __restrict__ hcase,
             const Consts* __restrict__ cases) {
    const int h = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (h >= nhalo) return;
    if (hcase) a.fr = cases[hcase[h]].fr;  // batches: the node's case
    const int g = hnode[h] - 1;
    const double4 u = ld4(U + 4 * (long long)g);
    const int j0 = hptr[h], j1 = hptr[h + 1];
    // the first kHaloBatch entries' loads all in flight at once (a node lies
    // in 2-4 patches as a rule), then the sums in entry order
    constexpr int kHaloBatch = 4;
    double4 v[kHaloBatch];
#pragma unroll
    for (int q = 0; q < kHaloBatch; ++q) {
        const int j = j0 + q;
        v[q] = j < j1 ? ld4(hpart + 4 * (long long)hidx[j]) : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
#pragma unroll
    for (int q = 0; q < kHaloBatch; ++q) {
        if (j0 + q < j1) {
            r0 = r0 + v[q].x;
            r1 = r1 + v[q].y;
            r2 = r2 + v[q].z;
            r3 = r3 + v[q].w;
        }
    }
    for (int j = j0 + kHaloBatch; j < j1; ++j) {
        const double4 w = ld4(hpart + 4 * (long long)hidx[j]);
        r0 = r0 + w.x;
        r1 = r1 + w.y;
        r2 = r2 + w.z;
        r3 = r3 + w.w;
    }
    stage_node<kFinal>(a, g, u, r0, r1, r2, r3);
}

// ESTAB in patch order (fused paths): one patch per block (or, resident
// grid, patches blockIdx.x + k gridDim.x); the patch's node values in LDS;
// element q's parameters, rounded to R, written at q (coalesced).  FP64
// gradients: the patch-order copy when the element type is double (an exact
// copy), else the element-order arrays through perm.
struct EstabNode {
    double rho, T, vx, vy, wx, wy, gm;
};
struct EstabIn {
    const double *U, *T, *GAMM, *VX, *VY, *WX, *WY;  // WX, WY null: W = 0
    const double *dNx, *dNy;      // element order
    const double* dN6;            // patch order, 6 per element (null: dNx/dNy through perm)
    const int* perm;
    const double* dtmin;          // + case * dt_stride
    double fr, tinf, rinf;
};

template <class R>
__global__ void __launch_bounds__(512)
k_estab_patch(PatchDev pd, EstabIn ei, ElemP<R> o) {
    extern __shared__ double4 smem4[];
    EstabNode* sN = reinterpret_cast<EstabNode*>(smem4);  // maxn
    for (int p0 = blockIdx.x; p0 < pd.npatch; p0 += gridDim.x) {
        const int p = (int)gridDim.x == pd.npatch ? xcd_block(p0, pd.npatch) : p0;
        const int cs = pd.pcase ? pd.pcase[p] : 0;
        const Consts* cq = pd.cases ? pd.cases + cs : nullptr;
        const double fr = cq ? cq->fr : ei.fr, tinf = cq ? cq->t_inf : ei.tinf;
        const double rinf = cq ? cq->rho_inf : ei.rinf, cte = cq ? cq->cte : o.cte;
        const double dtmin = ei.dtmin[(size_t)cs * o.dt_stride];
        const int nb = pd.node_base[p], nl = pd.node_base[p + 1] - nb;
        const long long e0 = pd.elem_base[p];
        const int nep = pd.elem_base[p + 1] - pd.elem_base[p];
        for (int l = threadIdx.x; l < nl; l += blockDim.x) {
            const int g = pd.pnode[nb + l] - 1;
            sN[l] = EstabNode{ei.U[4 * (long long)g], ei.T[g], ei.VX[g], ei.VY[g],
                              ei.WX ? ei.WX[g] : 0.0, ei.WY ? ei.WY[g] : 0.0, ei.GAMM[g]};
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nep; q += blockDim.x) {
            const long long e = e0 + q;
            const uint64_t w = (uint64_t)pd.loc[2 * e] | ((uint64_t)pd.loc[2 * e + 1] << 32);
            const int ln[3] = {(int)(w & 1023u), (int)((w >> 10) & 1023u), (int)((w >> 20) & 1023u)};
            const int eo = ei.perm[e];
            double DNX[3], DNY[3], rho[3], Tn[3], vxn[3], vyn[3], wxn[3], wyn[3], gmn[3];
            if (ei.dN6) {  // one 48-byte record: three 16-byte loads
                const double2* r = reinterpret_cast<const double2*>(ei.dN6 + 6 * e);
                const double2 a = r[0], b = r[1], c = r[2];
                DNX[0] = a.x;
                DNX[1] = a.y;
                DNX[2] = b.x;
                DNY[0] = b.y;
                DNY[1] = c.x;
                DNY[2] = c.y;
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    DNX[k] = ei.dNx[3 * (long long)eo + k];
                    DNY[k] = ei.dNy[3 * (long long)eo + k];
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const EstabNode& nd = sN[ln[k]];
                rho[k] = nd.rho;
                Tn[k] = nd.T;
                vxn[k] = nd.vx;
                vyn[k] = nd.vy;
                wxn[k] = nd.wx;
                wyn[k] = nd.wy;
                gmn[k] = nd.gm;
            }
            double sh, rrr, s2, s3;
            estab_elem(DNX, DNY, rho, Tn, vxn, vyn, wxn, wyn, gmn, fr, tinf, rinf, dtmin, sh, rrr,
                       s2, s3);
            *reinterpret_cast<typename Vec4<R>::type*>(o.nt + 4 * e) =
                make4((R)(sh * cte), (R)rrr, (R)s2, (R)s3);
            // DTL = DTMIN in every element unless local time steps blend in DT
            o.fac[e] = (R)(o.area[eo] * (o.dtl ? o.dtl[eo] : dtmin) / 3.0);
        }
        __syncthreads();  // the next patch overwrites sN
    }
}

// One gather of the element's six gradients, two outputs in patch order:
// b6[6q+k] = nx[3 perm[q] + k], b6[6q+3+k] = ny[3 perm[q] + k] (ESTAB's record,
// three 16-byte stores) and b4[4q..4q+3] = Nx2, Nx3, Ny2, Ny3 (the stage kernel's)
__global__ void k_to_patch64(const double* __restrict__ nx, const double* __restrict__ ny,
                             double* __restrict__ b6, double* __restrict__ b4,
                             const int* __restrict__ perm, int nelem) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nelem) return;
    const long long e = perm[q];
    const double x0 = nx[3 * e], x1 = nx[3 * e + 1], x2 = nx[3 * e + 2];
    const double y0 = ny[3 * e], y1 = ny[3 * e + 1], y2 = ny[3 * e + 2];
    double2* o6 = reinterpret_cast<double2*>(b6 + 6 * (long long)q);
    o6[0] = make_double2(x0, x1);
    o6[1] = make_double2(x2, y0);
    o6[2] = make_double2(y1, y2);
    *reinterpret_cast<typename Vec4<double>::type*>(b4 + 4 * (long long)q) = make4(x1, x2, y1, y2);
}

template <class R>
__global__ void k_to_patch_grad(const double* __restrict__ nx, const double* __restrict__ ny,
                                R* __restrict__ dN, const int* __restrict__ perm, int nelem) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nelem) return;
    const long long e = perm[q];
    *reinterpret_cast<typename Vec4<R>::type*>(dN + 4 * (long long)q) =
        make4((R)nx[3 * e + 1], (R)nx[3 * e + 2], (R)ny[3 * e + 1], (R)ny[3 * e + 2]);
}

}  // namespace

// ---------------------------------------------------------------------------
int launch_deltat_partials(const Tables& t, const Consts& c, const double* area, const double* T,
                           const double* VX, const double* VY, const double* WX, const double* WY,
                           double* DT, double* partials, hipStream_t s) {
    const unsigned g = std::min<unsigned>(grid_for(t.nelem), kRedBlocks);
    k_deltat<<<g, kBlock, 0, s>>>(t.inpoel.p, area, T, VX, VY, WX, WY, DT, t.nelem, c.t_inf, c.fsafe,
                                  partials, nullptr, nullptr, nullptr);
    CFD_HIP(hipGetLastError());
    return (int)g;
}

// Threads per block of the patch-order element passes (deltat, ESTAB): one
// block per patch, every thread forming the same number of elements
// (kPassElems) when the patch shape allows it.  Two per thread: a 384-element
// FP64 patch on 192 threads instead of 256 (where a third of the threads
// formed one element and the rest two) took ESTAB from 171 to 164 us at 4.2M
// triangles; one per thread (384 threads) took it to 228 us.
constexpr int kPassElems = 2;
constexpr int kPassMaxBlock = 512;
inline int patch_pass_block(int pe) {
    const int b = pe / kPassElems;
    return pe % (64 * kPassElems) == 0 && b <= kPassMaxBlock ? b : kBlock;
}

int launch_deltat_patch(const PatchSet& ps, const Consts& c, const double* area, const double* T,
                        const double* VX, const double* VY, const double* WX, const double* WY,
                        double* DT, double* partials, hipStream_t s) {
    CFD_CHECK(ps.npatch > 0 && ps.max_nodes <= kPatchMaxNodes, "internal: deltat patches");
    const size_t lds = (size_t)ps.max_nodes * sizeof(DtNode);
    k_deltat_patch<<<ps.npatch, patch_pass_block(ps.pe), lds, s>>>(ps.node_base.p, ps.elem_base.p, ps.pnode.p,
                                                  ps.loc.p, ps.perm.p, ps.npatch, area, T, VX, VY,
                                                  WX, WY, DT, c.t_inf, c.fsafe, partials);
    CFD_HIP(hipGetLastError());
    return ps.npatch;
}

void launch_deltat_chunks(const Tables& t, const Consts& c, const double* area, const double* T,
                          const double* VX, const double* VY, const double* WX, const double* WY,
                          double* DT, const int* chunk, int nchunks, double* partials,
                          const int* chunk_case, const Consts* cases, hipStream_t s) {
    k_deltat<<<nchunks, kBlock, 0, s>>>(t.inpoel.p, area, T, VX, VY, WX, WY, DT, t.nelem, c.t_inf,
                                        c.fsafe, partials, chunk, chunk_case, cases);
    CFD_HIP(hipGetLastError());
}

void launch_deltat(const Tables& t, const Consts& c, const double* area, const double* T,
                   const double* VX, const double* VY, const double* WX, const double* WY,
                   double* DT, double* dtmin_out, double* partials, hipStream_t s) {
    const int g = launch_deltat_partials(t, c, area, T, VX, VY, WX, WY, DT, partials, s);
    k_min_finalize<<<1, 256, 0, s>>>(partials, g, 1.0e20, __builtin_huge_val(), dtmin_out);
    k_deltat_clip<<<grid_for(t.nelem), kBlock, 0, s>>>(DT, dtmin_out, t.nelem);
    CFD_HIP(hipGetLastError());
}

void launch_estab(const Tables& t, const Consts& c, const double* U, const double* T,
                  const double* GAMM, const double* VX, const double* VY, const double* WX,
                  const double* WY, const double* dNx, const double* dNy, const double* dtmin,
                  double* shoc, double* t1, double* t2, double* t3, hipStream_t s) {
    k_estab<<<grid_for(t.nelem), kBlock, 0, s>>>(t.inpoel.p, U, T, GAMM, VX, VY, WX, WY, dNx, dNy,
                                                 t.nelem, c.fr, c.t_inf, c.rho_inf, dtmin, shoc,
                                                 t1, t2, t3);
    CFD_HIP(hipGetLastError());
}

template <class R>
void launch_estab_patch(const Tables& t, const PatchSet& ps, const Consts& c, const double* U,
                        const double* T, const double* GAMM, const double* VX, const double* VY,
                        const double* WX, const double* WY, const double* dNx, const double* dNy,
                        const double* dN6_patch, const double* dtmin,
                        const ElemP<R>& o, hipStream_t s) {
    CFD_CHECK(o.nt && ps.perm.p, "internal: patch outputs missing");
    CFD_CHECK(ps.max_nodes <= kPatchMaxNodes, "internal: patch beyond the node budget");
    PatchDev pd{};
    pd.loc = ps.loc.p;
    pd.elem_base = ps.elem_base.p;
    pd.node_base = ps.node_base.p;
    pd.pnode = ps.pnode.p;
    pd.nelem = t.nelem;
    pd.maxn = ps.max_nodes;
    pd.pe = ps.pe;
    pd.npatch = ps.npatch;
    pd.pcase = ps.cases ? ps.pcase.p : nullptr;
    pd.cases = ps.cases;
    EstabIn ei{U, T, GAMM, VX, VY, WX, WY, dNx, dNy, dN6_patch, ps.perm.p, dtmin,
               c.fr, c.t_inf, c.rho_inf};
    const size_t lds = (size_t)ps.max_nodes * sizeof(EstabNode);
    k_estab_patch<R><<<ps.npatch, patch_pass_block(ps.pe), lds, s>>>(pd, ei, o);
    CFD_HIP(hipGetLastError());
}

template <class R>
void launch_estab_widen(const ElemP<R>& o, const int* inv_perm, int nelem, double* shoc,
                        double* t1, double* t2, double* t3, hipStream_t s) {
    k_estab_widen<R><<<grid_for(nelem), kBlock, 0, s>>>(o.nt, inv_perm, 1.0 / o.cte, nelem,
                                                        shoc, t1, t2, t3);
    CFD_HIP(hipGetLastError());
}

template void launch_estab_patch<float>(const Tables&, const PatchSet&, const Consts&,
                                         const double*, const double*, const double*,
                                         const double*, const double*, const double*,
                                         const double*, const double*, const double*,
                                         const double*, const double*, const ElemP<float>&,
                                         hipStream_t);
template void launch_estab_patch<double>(const Tables&, const PatchSet&, const Consts&,
                                          const double*, const double*, const double*,
                                          const double*, const double*, const double*,
                                          const double*, const double*, const double*,
                                          const double*, const double*, const ElemP<double>&,
                                          hipStream_t);
template void launch_estab_widen<float>(const ElemP<float>&, const int*, int, double*, double*,
                                        double*, double*, hipStream_t);
template void launch_estab_widen<double>(const ElemP<double>&, const int*, int, double*, double*,
                                         double*, double*, hipStream_t);

static ElemConsts elem_consts(const Consts& c) {
    ElemConsts k;
    k.g = c.gama;
    k.Cv = c.fcv;
    k.lambda_ref = c.fk;
    k.mu_ref = c.fmu;
    k.T_inf = c.t_inf;
    k.cte = c.cte;
    k.viscous = c.fmu > 2.2250738585072014e-308 ? 1 : 0;  // mu_ref > tiny(0d0)
    return k;
}

void launch_element_rhs(const Tables& t, const Consts& c, const ElemIn& in, double* c_calc,
                        double* c_fuente, hipStream_t s) {
    ElemConsts k = elem_consts(c);
    dim3 g(grid_for(t.nelem)), b(kBlock);
    const bool th = in.theta != nullptr, fu = in.WX != nullptr;
#define CFD_ELEM(TH, FU)                                                                      \
    k_element_rhs<true, TH, FU><<<g, b, 0, s>>>(k, t.inpoel.p, t.epos.p, in.U, in.theta, in.T,       \
                                          in.dNx, in.dNy, in.area, in.shoc, in.dtl, in.t1,     \
                                          in.t2, in.t3, in.WX, in.WY, c_calc, c_fuente, t.nelem)
    if (th && fu) CFD_ELEM(true, true);
    else if (th) CFD_ELEM(true, false);
    else if (fu) CFD_ELEM(false, true);
    else CFD_ELEM(false, false);
#undef CFD_ELEM
    CFD_HIP(hipGetLastError());
}

void launch_fuente(const Tables& t, const ElemIn& in, double* c_fuente, hipStream_t s) {
    ElemConsts k{};
    k_element_rhs<false, false, true><<<grid_for(t.nelem), kBlock, 0, s>>>(
        k, t.inpoel.p, t.epos.p, in.U, nullptr, nullptr, in.dNx, in.dNy, in.area, nullptr, in.dtl,
        nullptr, nullptr, nullptr, in.WX, in.WY, nullptr, c_fuente, t.nelem);
    CFD_HIP(hipGetLastError());
}

void launch_gather_rhs(const Tables& t, const double* rhs_in, const double* c_calc,
                       const double* c_fuente, double* rhs_out, hipStream_t s) {
    k_gather_rhs<double4><<<grid_for(t.npoin), kBlock, 0, s>>>(
        t.esup2.p, rhs_in, reinterpret_cast<const double4*>(c_calc),
        reinterpret_cast<const double4*>(c_fuente), rhs_out, t.npoin);
    CFD_HIP(hipGetLastError());
}

static StageDev stage_dev(const Consts& c, const BcTable& bc, double rk_fact,
                          const StageNodes& nd) {
    StageDev a;
    a.rk_fact = rk_fact;
    a.fr = c.fr;
    a.M = nd.M;
    a.GAMM = nd.GAMM;
    a.WX = nd.WX;
    a.WY = nd.WY;
    a.bcidx = bc.bcidx.p;
    a.bcflags = bc.flags.p;
    a.bcvals = bc.vals.p;
    a.U1 = nd.U1;
    a.RHS = nd.RHS;
    a.T = nd.T;
    a.VX = nd.VX;
    a.VY = nd.VY;
    a.RHO = nd.RHO;
    a.E = nd.E;
    a.P = nd.P;
    a.RMACH = nd.RMACH;
    return a;
}

void launch_stage_update(const Tables& t, const Consts& c, const BcTable& bc, double rk_fact,
                         const double* c_calc, const double* c_fuente, const StageNodes& nd,
                         bool final_stage, hipStream_t s) {
    const StageDev a = stage_dev(c, bc, rk_fact, nd);
    const double4* cc = reinterpret_cast<const double4*>(c_calc);
    const double4* cf = reinterpret_cast<const double4*>(c_fuente);
    const unsigned g = grid_for(t.npoin);
    if (cf && final_stage)
        k_stage<true, true><<<g, kBlock, 0, s>>>(t.esup2.p, cc, cf, nd.U, a, t.npoin);
    else if (cf)
        k_stage<true, false><<<g, kBlock, 0, s>>>(t.esup2.p, cc, cf, nd.U, a, t.npoin);
    else if (final_stage)
        k_stage<false, true><<<g, kBlock, 0, s>>>(t.esup2.p, cc, cf, nd.U, a, t.npoin);
    else
        k_stage<false, false><<<g, kBlock, 0, s>>>(t.esup2.p, cc, cf, nd.U, a, t.npoin);
    CFD_HIP(hipGetLastError());
}

static size_t patch_lds_bytes(int maxn, bool fuente, int pe, size_t slot) {
    return (size_t)maxn * (sizeof(double4) + sizeof(double) + (fuente ? sizeof(double2) : 0)) +
           3 * (size_t)pe * slot + (size_t)(maxn + 1) * sizeof(uint16_t);
}

bool patch_lds_fits(int max_nodes, int pe, size_t slot_bytes) {
    return patch_lds_bytes(max_nodes, true, pe, slot_bytes) <= 65536;
}

template <class R, bool kFuente, bool kFinal>
static void patch_launch_t(const PatchSet& ps, const PatchDev& pd, const ElemConstsP<R>& kc,
                           const PatchIn<R>& in, const StageDev& a, hipStream_t s) {
    const size_t lds = patch_lds_bytes(pd.maxn, kFuente, pd.pe, sizeof(typename Vec4<R>::type));
    const int grid = ps.persistent && ps.grid[kFinal] > 0 ? ps.grid[kFinal] : ps.npatch;
    k_patch_stage<R, kFuente, kFinal><<<grid, ps.block, lds, s>>>(pd, kc, in, a);
    if (ps.nhalo > 0)
        k_halo_stage<kFinal><<<grid_for(ps.nhalo), kBlock, 0, s>>>(
            ps.hnode.p, ps.hptr.p, ps.hidx.p, ps.hpart.p, in.U, a, ps.nhalo,
            ps.cases ? ps.hcase.p : nullptr, ps.cases);
}

template <class R>
void launch_patch_stage(const Tables& t, const PatchSet& ps, const Consts& c, const BcTable& bc,
                        double rk_fact, const PatchIn<R>& in, const StageNodes& nd,
                        bool final_stage, hipStream_t s) {
    ElemConstsP<R> kc;
    kc.g = (R)c.gama;
    kc.Cv = (R)c.fcv;
    kc.lambda_ref = (R)c.fk;
    kc.mu_ref = (R)c.fmu;
    kc.T_inf = (R)c.t_inf;
    kc.viscous = c.fmu > 2.2250738585072014e-308 ? 1 : 0;
    PatchDev pd;
    pd.loc = ps.loc.p;
    pd.elem_base = ps.elem_base.p;
    pd.node_base = ps.node_base.p;
    pd.n_int = ps.n_int.p;
    pd.halo_base = ps.halo_base.p;
    pd.pnode = ps.pnode.p;
    pd.lptr = ps.lptr.p;
    pd.hpart = ps.hpart.p;
    pd.nelem = t.nelem;
    pd.maxn = ps.max_nodes;
    pd.pe = ps.pe;
    pd.npatch = ps.npatch;
    pd.pcase = ps.cases ? ps.pcase.p : nullptr;
    pd.cases = ps.cases;
    CFD_CHECK(ps.block >= 64 && ps.block <= 512 && ps.block % 64 == 0, "internal: patch block size");
    CFD_CHECK(patch_lds_fits(pd.maxn, ps.pe, sizeof(typename Vec4<R>::type)),
              "mesh: element patch exceeds the LDS budget");
    const StageDev a = stage_dev(c, bc, rk_fact, nd);
    const bool fu = in.WX != nullptr;
    if (fu && final_stage) patch_launch_t<R, true, true>(ps, pd, kc, in, a, s);
    else if (fu) patch_launch_t<R, true, false>(ps, pd, kc, in, a, s);
    else if (final_stage) patch_launch_t<R, false, true>(ps, pd, kc, in, a, s);
    else patch_launch_t<R, false, false>(ps, pd, kc, in, a, s);
    CFD_HIP(hipGetLastError());
}

template <class R, bool kFuente, bool kFinal>
static int resident_blocks(const PatchSet& ps) {
    int dev = 0, ncu = 0, per_cu = 0;
    CFD_HIP(hipGetDevice(&dev));
    CFD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds = patch_lds_bytes(ps.max_nodes, kFuente, ps.pe, sizeof(typename Vec4<R>::type));
    CFD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_patch_stage<R, kFuente, kFinal>), ps.block, lds));
    return std::max(1, std::min(ps.npatch, ncu * std::max(1, per_cu)));
}

template <class R>
void patch_make_persistent(PatchSet& ps, bool fuente) {
    ps.persistent = true;
    ps.grid[0] = fuente ? resident_blocks<R, true, false>(ps) : resident_blocks<R, false, false>(ps);
    ps.grid[1] = fuente ? resident_blocks<R, true, true>(ps) : resident_blocks<R, false, true>(ps);
}

template void patch_make_persistent<float>(PatchSet&, bool);
template void patch_make_persistent<double>(PatchSet&, bool);

void launch_to_patch64(const double* dNx, const double* dNy, double* dN6, double* dN4,
                       const int* perm, int nelem, hipStream_t s) {
    if (nelem <= 0) return;
    k_to_patch64<<<grid_for(nelem), kBlock, 0, s>>>(dNx, dNy, dN6, dN4, perm, nelem);
    CFD_HIP(hipGetLastError());
}

template void launch_patch_stage<float>(const Tables&, const PatchSet&, const Consts&,
                                        const BcTable&, double, const PatchIn<float>&,
                                        const StageNodes&, bool, hipStream_t);
template void launch_patch_stage<double>(const Tables&, const PatchSet&, const Consts&,
                                         const BcTable&, double, const PatchIn<double>&,
                                         const StageNodes&, bool, hipStream_t);

template <class R>
void launch_to_patch_grad(const double* dNx, const double* dNy, R* dN, const int* perm, int nelem,
                          hipStream_t s) {
    if (nelem <= 0) return;
    k_to_patch_grad<R><<<grid_for(nelem), kBlock, 0, s>>>(dNx, dNy, dN, perm, nelem);
    CFD_HIP(hipGetLastError());
}
template void launch_to_patch_grad<float>(const double*, const double*, float*, const int*, int,
                                          hipStream_t);

}  // namespace cfd
