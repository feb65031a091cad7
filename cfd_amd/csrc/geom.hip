// Geometry kernels of the path: P1 gradients and areas (deriv), lumped mass
// (MASAS), mesh-Laplace values (mLaplace), the GCL mass update (gcl.f90),
// wall normals (NORMALES), and the small per-set loops of fluidStructure
// (FORCES, TRANSF).
//
// Node-owned results are gathers over each node's esup segment in ascending
// element order: that is the order in which the reference's sequential
// element-order scatter adds into the node, so sums are bitwise identical and
// no atomics are needed.  All of these are HBM-bound integer/gather work.
#include "cfd_internal.hpp"
#include "reduce.hpp"
#include "crmath.hpp"

namespace cfd {
namespace {

// subrutinas.f90:99-125
__global__ void __launch_bounds__(kBlock)
k_deriv(const int* __restrict__ inpoel, const double* __restrict__ X, const double* __restrict__ Y,
        double* __restrict__ area, double* __restrict__ dNx, double* __restrict__ dNy,
        double* __restrict__ HH, double* __restrict__ HHX, double* __restrict__ HHY, int nelem,
        double* __restrict__ partials) {
    double hm = __builtin_huge_val();
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += gridDim.x * blockDim.x) {
        int n1 = inpoel[3 * e] - 1, n2 = inpoel[3 * e + 1] - 1, n3 = inpoel[3 * e + 2] - 1;
        double X1 = X[n1], X2 = X[n2], X3 = X[n3];
        double Y1 = Y[n1], Y2 = Y[n2], Y3 = Y[n3];
        double a = (X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3)) / 2.0;
        area[e] = a;
        double a2 = 2.0 * a;
        dNx[3 * e + 0] = (Y2 - Y3) / a2;
        dNx[3 * e + 1] = (Y3 - Y1) / a2;
        dNx[3 * e + 2] = (Y1 - Y2) / a2;
        dNy[3 * e + 0] = (X3 - X2) / a2;
        dNy[3 * e + 1] = (X1 - X3) / a2;
        dNy[3 * e + 2] = (X2 - X1) / a2;
        double h = sqrt(a);
        HH[e] = h;
        if (h < hm) hm = h;
        double mx = X3 - X2, my = Y3 - Y2;
        if (X1 - X3 < mx) mx = X1 - X3;
        if (X2 - X1 < mx) mx = X2 - X1;
        if (Y1 - Y3 < my) my = Y1 - Y3;
        if (Y2 - Y1 < my) my = Y2 - Y1;
        HHX[e] = fabs(mx);
        HHY[e] = fabs(my);
    }
    block_partial(hm, OpMin(), partials);
}

// subrutinas.f90:137-152: M(i) = sum of area/3 over i's elements, ascending
__global__ void __launch_bounds__(kBlock)
k_masas(const int* __restrict__ esup1, const int* __restrict__ esup2, const double* __restrict__ area,
        double* __restrict__ M, int npoin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin) return;
    double m = 0.0;
    for (int s = esup2[i]; s < esup2[i + 1]; ++s) m = m + area[esup1[s] - 1] / 3.0;
    M[i] = m;
}

// mLaplace.f90:188-200
__device__ inline double quality_mu(double X1, double X2, double X3, double Y1, double Y2, double Y3) {
    const double TWOSQRT3 = 3.46410161513775;
    double ar = X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3);
    double l1 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
    double l2 = (X1 - X3) * (X1 - X3) + (Y1 - Y3) * (Y1 - Y3);
    double l3 = (X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1);
    double l = l1 + l2 + l3;
    return TWOSQRT3 * ar / l;
}

// mLaplace.f90:120-148, one row per thread; column positions from the slot
// map.  A row of up to kLapRow entries accumulates in LDS (the same sequence
// of additions per entry as in memory), column-major over the block's
// threads so lanes hit distinct banks; longer rows accumulate in memory.
constexpr int kLapRow = 16;

// row i's contributions in esup order; acc(k) is the accumulator of CSR
// entry r0 + k
template <class Acc>
__device__ inline void laplace_row(int i, int r0, const int* __restrict__ inpoel,
                                   const int* __restrict__ esup1, const int* __restrict__ esup2,
                                   const int* __restrict__ lslot, const double* __restrict__ X,
                                   const double* __restrict__ Y, const double* __restrict__ dNx,
                                   const double* __restrict__ dNy, Acc acc) {
    for (int s = esup2[i]; s < esup2[i + 1]; ++s) {
        int e = esup1[s] - 1;
        int n1 = inpoel[3 * e] - 1, n2 = inpoel[3 * e + 1] - 1, n3 = inpoel[3 * e + 2] - 1;
        double mu = quality_mu(X[n1], X[n2], X[n3], Y[n1], Y[n2], Y[n3]);
        double q = 1.0 / (mu * mu);
        int li = (n1 == i) ? 0 : ((n2 == i) ? 1 : 2);
        double ax = dNx[3 * e + li], ay = dNy[3 * e + li];
        for (int j = 0; j < 3; ++j) {
            double& a = acc(lslot[3 * s + j] - r0);
            a = a + (ax * dNx[3 * e + j] + ay * dNy[3 * e + j]) * q;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_laplace(const int* __restrict__ inpoel, const int* __restrict__ esup1, const int* __restrict__ esup2,
          const int* __restrict__ rowptr, const int* __restrict__ lslot, const double* __restrict__ X,
          const double* __restrict__ Y, const double* __restrict__ dNx, const double* __restrict__ dNy,
          double* __restrict__ lap, double* __restrict__ diag, int npoin) {
    __shared__ double srow[kLapRow * kBlock];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin) return;
    const int r0 = rowptr[i], len = rowptr[i + 1] - r0;
    if (len <= kLapRow) {
        double* sr = srow + threadIdx.x;
        for (int j = 0; j < len; ++j) sr[j * kBlock] = 0.0;
        laplace_row(i, r0, inpoel, esup1, esup2, lslot, X, Y, dNx, dNy,
                    [sr](int k) -> double& { return sr[k * kBlock]; });
        for (int j = 0; j < len; ++j) lap[r0 + j] = sr[j * kBlock];
        diag[i] = sr[0];
    } else {
        double* row = lap + r0;
        for (int j = 0; j < len; ++j) row[j] = 0.0;
        laplace_row(i, r0, inpoel, esup1, esup2, lslot, X, Y, dNx, dNy,
                    [row](int k) -> double& { return row[k]; });
        diag[i] = row[0];
    }
}

// gcl.f90:218-236 per node: both divergence sums in ascending element order
__global__ void __launch_bounds__(kBlock)
k_gcl(const int* __restrict__ inpoel, const int* __restrict__ esup1, const int* __restrict__ esup2,
      double* __restrict__ M, const double* __restrict__ Wx, const double* __restrict__ Wy,
      const double* __restrict__ Wxo, const double* __restrict__ Wyo, const double* __restrict__ dNx,
      const double* __restrict__ dNy, const double* __restrict__ area,
      const double* __restrict__ area_old, const double* __restrict__ dtp, int mode, int npoin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin) return;
    const double dt = *dtp;
    const double* wy = mode ? Wy : Wx;     // as written, the y-part reads W_x
    const double* wyo = mode ? Wyo : Wxo;
    double tot1 = 0.0, tot2 = 0.0;
    for (int s = esup2[i]; s < esup2[i + 1]; ++s) {
        int e = esup1[s] - 1;
        int n0 = inpoel[3 * e] - 1, n1 = inpoel[3 * e + 1] - 1, n2 = inpoel[3 * e + 2] - 1;
        const double* nx = dNx + 3 * e;
        const double* ny = dNy + 3 * e;
        double divW = nx[0] * Wx[n0] + nx[1] * Wx[n1] + nx[2] * Wx[n2] + ny[0] * wy[n0] +
                      ny[1] * wy[n1] + ny[2] * wy[n2];
        double divWo = nx[0] * Wxo[n0] + nx[1] * Wxo[n1] + nx[2] * Wxo[n2] + ny[0] * wyo[n0] +
                       ny[1] * wyo[n1] + ny[2] * wyo[n2];
        tot1 = tot1 + divW * area[e] / 3.0;
        tot2 = tot2 + divWo * area_old[e] / 3.0;
    }
    M[i] = M[i] + dt * (tot1 + tot2) / 2.0;
}

// subrutinas.f90:26-63 per wall node; incident wall edges in ascending order
__global__ void k_normales(int nw, const int* __restrict__ wnode_bc, const int* __restrict__ wnode_ptr,
                           const int* __restrict__ wnode_edges, const int* __restrict__ wall,
                           const double* __restrict__ X, const double* __restrict__ Y,
                           int* __restrict__ flags, double* __restrict__ vals) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    double nlx = 0.0, nly = 0.0, den = 0.0;
    for (int q = wnode_ptr[w]; q < wnode_ptr[w + 1]; ++q) {
        int ed = wnode_edges[q];
        int a = wall[2 * ed] - 1, b = wall[2 * ed + 1] - 1;
        double lx = Y[b] - Y[a];
        double ly = -(X[b] - X[a]);
        double l = sqrt(lx * lx + ly * ly);
        nlx = nlx + lx;
        nly = nly + ly;
        den = den + l;
    }
    int bi = wnode_bc[w];
    int f = flags[bi] & ~2;
    if (den > 1.0e-6) {
        double lx = nlx / den, ly = nly / den;
        double ln = sqrt(lx * lx + ly * ly);
        if (ln > 0.2) {
            f |= 2;
            vals[8 * bi + 2] = lx / ln;
            vals[8 * bi + 3] = ly / ln;
        }
    }
    flags[bi] = f;
}

// subrutinas.f90:75-84 on the nodes NORMALES marked (flag bit 1): the
// velocity relative to the mesh keeps its tangential part, then W is added
__global__ void k_normalvel(const int* __restrict__ bcidx, const int* __restrict__ flags,
                            const double* __restrict__ vals, double* __restrict__ VX,
                            double* __restrict__ VY, const double* __restrict__ WX,
                            const double* __restrict__ WY, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int bi = bcidx[i];
    if (bi < 0 || !(flags[bi] & 2)) return;
    const double nx = vals[8 * bi + 2], ny = vals[8 * bi + 3];
    const double vx = VX[i], vy = VY[i], wx = WX[i], wy = WY[i];
    const double p = -ny * (vx - wx) + nx * (vy - wy);
    VX[i] = -ny * p + wx;
    VY[i] = nx * p + wy;
}

// meshMove.f90:167-192 FORCES and ns2DComp.ALE.f90:839-887 FORCE_VISC, one
// set.  FORCE_VISC's traction is that of the full stress tensor (-p I +
// viscous, with the element-average pressure), Sutherland mu at the
// edge-average temperature, on the edge normal (Y2-Y1, -(X2-X1)) / |edge|.
// r = FX, FY, RM, F_VX, F_VY terms of edge i.
template <bool kPress, bool kVisc>
__device__ __forceinline__ void force_terms(int i, const int* __restrict__ iset,
                                            const double* __restrict__ X,
                                            const double* __restrict__ Y,
                                            const double* __restrict__ P, double xref,
                                            double yref, const ViscIn& v, double r[5]) {
    const int n1 = iset[3 * i] - 1, n2 = iset[3 * i + 1] - 1;
    if (kPress) {
        const double d_press = (P[n1] + P[n2]) / 2.0;
        const double rlx = X[n1] - X[n2];
        const double rly = Y[n2] - Y[n1];
        const double dfx = d_press * rly, dfy = d_press * rlx;
        const double xc = (X[n1] + X[n2]) / 2.0;
        const double yc = (Y[n2] + Y[n1]) / 2.0;
        r[0] = dfx;
        r[1] = dfy;
        r[2] = dfy * (xc - xref) - dfx * (yc - yref);
    }
    if (kVisc) {
        const int ie = iset[3 * i + 2] - 1;
        const double temp = (v.T[n1] + v.T[n2]) / 2.0;
        const double smu = 110.0;
        const double fmu = 0.017 * pow_3_2(temp / v.t_inf) * (v.t_inf + smu) / (temp + smu);
        double rly = -(X[n2] - X[n1]);
        double rlx = Y[n2] - Y[n1];
        const double rmod = sqrt(rlx * rlx + rly * rly);
        rlx = rlx / rmod;
        rly = rly / rmod;
        double dux = 0.0, duy = 0.0, dvx = 0.0, dvy = 0.0, press = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int nn = v.inpoel[3 * ie + jj] - 1;
            const double nx = v.dNx[3 * ie + jj], ny = v.dNy[3 * ie + jj];
            const double vx = v.VX[nn], vy = v.VY[nn];
            dux = dux + nx * vx;
            duy = duy + ny * vx;
            dvx = dvx + nx * vy;
            dvy = dvy + ny * vy;
            press = press + P[nn];
        }
        press = press / 3.0;
        const double txx = -press - fmu * (2.0 / 3.0 * (dux + dvy) - 2.0 * dux);
        const double txy = fmu * (duy + dvx);
        const double tyy = -press - fmu * (2.0 / 3.0 * (dux + dvy) - 2.0 * dvy);
        r[3] = (txx * rlx + txy * rly) * rmod;
        r[4] = (txy * rlx + tyy * rly) * rmod;
    }
}

// One wave per 64 edges: lane l forms edge 64 b + l, the wave sums the terms
// by a fixed shuffle tree and lane 0 stores the block's five partials at
// part[q * nb + b].
template <bool kPress, bool kVisc>
__global__ void __launch_bounds__(64)
k_forces_part(int nedge, const int* __restrict__ iset, const double* __restrict__ X,
              const double* __restrict__ Y, const double* __restrict__ P, double xref,
              double yref, ViscIn v, double* __restrict__ part) {
    const int nb = gridDim.x, b = blockIdx.x;
    const int i = b * 64 + threadIdx.x;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < nedge) force_terms<kPress, kVisc>(i, iset, X, Y, P, xref, yref, v, r);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        if (q < 3 && !kPress) continue;
        if (q >= 3 && !kVisc) continue;
        const double t = wave_reduce(r[q], OpSum());
        if (threadIdx.x == 0) part[q * nb + b] = t;
    }
}

// One wave folds the nb block partials: lane l sums partials l, l + 64, ...
// in that order, then the shuffle tree.  zero4 (when given) receives four zeros
// in the same launch (the iteration counts and residuals of mesh solves that
// stop at once on a mesh that does not move).
template <bool kPress, bool kVisc>
__global__ void __launch_bounds__(64)
k_forces_fin(int nb, const double* __restrict__ part, double* __restrict__ out,
             double* __restrict__ outv, double* __restrict__ zero4) {
    const int l = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        if (q < 3 && !kPress) continue;
        if (q >= 3 && !kVisc) continue;
        double a = 0.0;
        for (int k = l; k < nb; k += 64) a = a + part[q * nb + k];
        a = wave_reduce(a, OpSum());
        if (l == 0) {
            if (q < 3)
                out[q] = a;
            else
                outv[q - 3] = a;
        }
    }
    if (zero4 && l < 4) zero4[l] = 0.0;
}

// meshMove.f90:378-389, one set
__global__ void k_transf(int nedge, const int* __restrict__ iset, const double* __restrict__ X,
                         const double* __restrict__ Y, const double* __restrict__ cs,
                         double yposr, double xref, double yref, double* __restrict__ DXPOS,
                         double* __restrict__ DYPOS) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nedge) return;
    const double ca = cs[0], sa = cs[1];
    for (int jj = 0; jj < 2; ++jj) {
        int n = iset[3 * i + jj] - 1;
        double distx = X[n] - xref;
        double disty = Y[n] - yref + yposr;
        DXPOS[n] = ca * distx + sa * disty - distx;
        DYPOS[n] = -(sa * distx) + ca * disty - disty + yposr;
    }
}

}  // namespace

void launch_deriv(const Tables& t, const double* X, const double* Y, double* area, double* dNx,
                  double* dNy, double* HH, double* HHX, double* HHY, double* hmin_out,
                  double* partials, hipStream_t s) {
    const unsigned g = grid_for(t.nelem);
    k_deriv<<<g, kBlock, 0, s>>>(t.inpoel.p, X, Y, area, dNx, dNy, HH, HHX, HHY, t.nelem, partials);
    // hmin = minval(HH), capped at 1e10 (subrutinas.f90:124-125)
    k_min_finalize<<<1, 256, 0, s>>>(partials, (int)g, __builtin_huge_val(), 1.0e10, hmin_out);
    CFD_HIP(hipGetLastError());
}

void launch_masas(const Tables& t, const double* area, double* M, hipStream_t s) {
    k_masas<<<grid_for(t.npoin), kBlock, 0, s>>>(t.esup1.p, t.esup2.p, area, M, t.npoin);
    CFD_HIP(hipGetLastError());
}

void launch_laplace(const Tables& t, const double* X, const double* Y, const double* dNx,
                    const double* dNy, double* lap, double* diag, hipStream_t s) {
    k_laplace<<<grid_for(t.npoin), kBlock, 0, s>>>(t.inpoel.p, t.esup1.p, t.esup2.p, t.rowptr.p,
                                                   t.lslot.p, X, Y, dNx, dNy, lap, diag, t.npoin);
    CFD_HIP(hipGetLastError());
}

void launch_gcl(const Tables& t, double* M, const double* Wx, const double* Wy, const double* Wxo,
                const double* Wyo, const double* dNx, const double* dNy, const double* area,
                const double* area_old, const double* dt, int mode, hipStream_t s) {
    k_gcl<<<grid_for(t.npoin), kBlock, 0, s>>>(t.inpoel.p, t.esup1.p, t.esup2.p, M, Wx, Wy, Wxo, Wyo,
                                               dNx, dNy, area, area_old, dt, mode, t.npoin);
    CFD_HIP(hipGetLastError());
}

void launch_normales(const BcTable& bc, int nw, const int* wnode_bc, const int* wnode_ptr,
                     const int* wnode_edges, const int* wall, const double* X, const double* Y,
                     hipStream_t s) {
    if (nw == 0) return;
    k_normales<<<grid_for(nw), kBlock, 0, s>>>(nw, wnode_bc, wnode_ptr, wnode_edges, wall, X, Y,
                                               bc.flags.p, bc.vals.p);
    CFD_HIP(hipGetLastError());
}

void launch_normalvel(const BcTable& bc, double* VX, double* VY, const double* WX,
                      const double* WY, hipStream_t s) {
    const int n = (int)bc.bcidx.n;
    if (bc.nb == 0 || n == 0) return;
    k_normalvel<<<grid_for(n), kBlock, 0, s>>>(bc.bcidx.p, bc.flags.p, bc.vals.p, VX, VY, WX, WY,
                                               n);
    CFD_HIP(hipGetLastError());
}

size_t forces_part_size(int nedge) { return 5 * (size_t)std::max(1, (nedge + 63) / 64); }

void launch_forces(int nedge, const int* iset, const double* X, const double* Y, const double* P,
                   double xref, double yref, double* out3, const ViscIn* visc, double* out2,
                   double* part, double* zero4, hipStream_t s) {
    if (nedge == 0 || (!out3 && !visc)) {
        if (zero4) {
            k_forces_fin<false, false><<<1, 64, 0, s>>>(0, part, out3, out2, zero4);
            CFD_HIP(hipGetLastError());
        }
        return;
    }
    const ViscIn v = visc ? *visc : ViscIn{};
    const int nb = (nedge + 63) / 64;
    if (out3 && visc) {
        k_forces_part<true, true><<<nb, 64, 0, s>>>(nedge, iset, X, Y, P, xref, yref, v, part);
        k_forces_fin<true, true><<<1, 64, 0, s>>>(nb, part, out3, out2, zero4);
    } else if (out3) {
        k_forces_part<true, false><<<nb, 64, 0, s>>>(nedge, iset, X, Y, P, xref, yref, v, part);
        k_forces_fin<true, false><<<1, 64, 0, s>>>(nb, part, out3, out2, zero4);
    } else {
        k_forces_part<false, true><<<nb, 64, 0, s>>>(nedge, iset, X, Y, P, xref, yref, v, part);
        k_forces_fin<false, true><<<1, 64, 0, s>>>(nb, part, out3, out2, zero4);
    }
    CFD_HIP(hipGetLastError());
}

void launch_transf(int nedge, const int* iset, const double* X, const double* Y, const double* cs,
                   double yposr, double xref, double yref, double* DXPOS, double* DYPOS,
                   hipStream_t s) {
    if (nedge == 0) return;
    k_transf<<<grid_for(nedge), kBlock, 0, s>>>(nedge, iset, X, Y, cs, yposr, xref, yref, DXPOS,
                                                DYPOS);
    CFD_HIP(hipGetLastError());
}

}  // namespace cfd
