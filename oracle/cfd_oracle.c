/* Test infrastructure only: CPU restatement of the reference hot path.
 *
 * Build with -ffp-contract=off (see oracle/Makefile).  Every expression keeps
 * the reference's left-to-right association; x**2 is x*x, x**1.5d0 and
 * x**(-.5d0) call pow(), x**.5d0 is sqrt() (what flang emits for it).  The
 * OpenMP pragmas mirror the reference loops; with one thread the atomic
 * scatters run in element order, which is the order the parity tests use.
 */
#include "cfd_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ */
/* pointNeighbor.f90:5-43 getEsup                                       */
/* ------------------------------------------------------------------ */
void orc_get_esup(int nelem, int npoin, const int *inpoel, int *esup1, int *esup2)
{
    /* esup2 has npoin+1 entries; esup2[k] is Fortran esup2(k+1) */
    memset(esup2, 0, sizeof(int) * (size_t)(npoin + 1));
    for (int e = 0; e < nelem; ++e)               /* :14-19 count at node+1 */
        for (int i = 0; i < 3; ++i)
            esup2[inpoel[3 * e + i]] += 1;          /* (node+1) in 1-based = node in 0-based */
    for (int p = 1; p <= npoin; ++p)              /* :22-24 scan */
        esup2[p] += esup2[p - 1];
    for (int e = 0; e < nelem; ++e)               /* :29-36 fill */
        for (int i = 0; i < 3; ++i) {
            int node = inpoel[3 * e + i];           /* 1-based */
            int istor = esup2[node - 1] + 1;
            esup2[node - 1] = istor;
            esup1[istor - 1] = e + 1;
        }
    for (int p = npoin; p >= 1; --p)              /* :39-42 restore */
        esup2[p] = esup2[p - 1];
    esup2[0] = 0;
}

/* pointNeighbor.f90:45-91 getPsup; psup1 may be NULL (count only).
 * Returns the number of entries. */
int orc_get_psup(int nelem, int npoin, const int *inpoel, const int *esup1,
                 const int *esup2, int *psup1, int *psup2)
{
    (void)nelem;
    int *lpoin = (int *)calloc((size_t)npoin + 1, sizeof(int));
    int istor = 0;
    psup2[0] = 0;
    for (int ip = 1; ip <= npoin; ++ip) {
        for (int k = esup2[ip - 1] + 1; k <= esup2[ip]; ++k) {
            int e = esup1[k - 1];
            for (int i = 0; i < 3; ++i) {
                int jp = inpoel[3 * (e - 1) + i];
                if (jp != ip && lpoin[jp] != ip) {
                    istor += 1;
                    if (psup1) psup1[istor - 1] = jp;
                    lpoin[jp] = ip;
                }
            }
        }
        psup2[ip] = istor;
    }
    free(lpoin);
    return istor;
}

/* mLaplace.f90:60-94 initialize (structure only) */
void orc_lap_structure(int npoin, const int *psup1, const int *psup2,
                       int *lap_rowptr, int *lap_idx)
{
    lap_rowptr[0] = 0;
    for (int i = 2; i <= npoin + 1; ++i)          /* :82-84 */
        lap_rowptr[i - 1] = psup2[i - 1] + i - 1;
    for (int ip = 1; ip <= npoin; ++ip) {         /* :86-93 */
        lap_idx[lap_rowptr[ip - 1]] = ip;
        int ipsup = psup2[ip - 1] + 1;
        for (int i = lap_rowptr[ip - 1] + 2; i <= lap_rowptr[ip]; ++i) {
            lap_idx[i - 1] = psup1[ipsup - 1];
            ipsup += 1;
        }
    }
}

/* ------------------------------------------------------------------ */
/* subrutinas.f90:88-126 deriv                                          */
/* ------------------------------------------------------------------ */
static double min3(double a, double b, double c)
{
    double m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}

void orc_deriv(int nelem, const int *inpoel, const double *X, const double *Y,
               double *area, double *dNx, double *dNy, double *HH, double *HHX,
               double *HHY, double *hmin)
{
#pragma omp parallel for
    for (int e = 0; e < nelem; ++e) {
        int n1 = inpoel[3 * e] - 1, n2 = inpoel[3 * e + 1] - 1, n3 = inpoel[3 * e + 2] - 1;
        double X1 = X[n1], X2 = X[n2], X3 = X[n3];
        double Y1 = Y[n1], Y2 = Y[n2], Y3 = Y[n3];
        double a = (X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3)) / 2.0;
        area[e] = a;
        dNx[3 * e + 0] = (Y2 - Y3) / (2.0 * a);
        dNx[3 * e + 1] = (Y3 - Y1) / (2.0 * a);
        dNx[3 * e + 2] = (Y1 - Y2) / (2.0 * a);
        dNy[3 * e + 0] = (X3 - X2) / (2.0 * a);
        dNy[3 * e + 1] = (X1 - X3) / (2.0 * a);
        dNy[3 * e + 2] = (X2 - X1) / (2.0 * a);
        HH[e] = sqrt(a);
        HHX[e] = fabs(min3(X3 - X2, X1 - X3, X2 - X1));
        HHY[e] = fabs(min3(Y3 - Y2, Y1 - Y3, Y2 - Y1));
    }
    double h = HUGE_VAL;
    for (int e = 0; e < nelem; ++e)
        if (HH[e] < h) h = HH[e];
    if (h > 1.0e10) h = 1.0e10;
    *hmin = h;
}

/* subrutinas.f90:128-153 MASAS */
void orc_masas(int nelem, int npoin, const int *inpoel, const double *area, double *M)
{
    for (int p = 0; p < npoin; ++p) M[p] = 0.0;
#pragma omp parallel for
    for (int e = 0; e < nelem; ++e) {
        for (int i = 0; i < 3; ++i) {
#pragma omp atomic
            M[inpoel[3 * e + i] - 1] += area[e] / 3.0;
        }
    }
}

/* subrutinas.f90:155-218 deltat */
void orc_deltat(const orc_consts *c, int nelem, const int *inpoel, const double *area,
                const double *T, const double *VEL_X, const double *VEL_Y,
                const double *W_X, const double *W_Y, double *DT, double *dtmin)
{
    double dmin = 1.0e20;
    for (int e = 0; e < nelem; ++e) {
        const int *ip = inpoel + 3 * e;
        double vumax = 0.0, vvmax = 0.0;
        double t_iel = (T[ip[0] - 1] + T[ip[1] - 1] + T[ip[2] - 1]) / 3.0;
        for (int i = 0; i < 3; ++i) {
            int p = ip[i] - 1;
            double vu = fabs(VEL_X[p] - W_X[p]);
            double vv = fabs(VEL_Y[p] - W_Y[p]);
            if (vu > vumax) vumax = vu;
            if (vv > vvmax) vvmax = vv;
        }
        double hh = sqrt(2.0 * area[e]);
        double vel = sqrt(vumax * vumax + vvmax * vvmax);
        double smu = 110.0;
        double fmu = 0.017 * pow(t_iel / c->t_inf, 1.5) * (c->t_inf + smu) / (t_iel + smu);
        double et = fmu;
        double pe = (vel * hh) / (2.0 * et);
        double alpha = pe / 3.0 < 1.0 ? pe / 3.0 : 1.0;
        double deltatu = 1.0 / (4.0 * et / (hh * hh) + alpha * vel / hh);
        double deltatc = 1.0 / (4.0 * et / (hh * hh));
        double dtelem = c->fsafe / (1.0 / deltatc + 1.0 / deltatu);
        DT[e] = dtelem;
        if (dtelem < dmin) dmin = dtelem;
    }
    double cota = 10.0 * dmin;
    for (int e = 0; e < nelem; ++e)
        if (DT[e] > cota) DT[e] = cota;
    *dtmin = dmin;
}

/* subrutinas.f90:332-446 ESTAB */
void orc_estab(const orc_consts *c, int nelem, const int *inpoel, const double *U,
               const double *T, const double *GAMM, const double *VEL_X,
               const double *VEL_Y, const double *W_X, const double *W_Y,
               const double *dNx, const double *dNy, double dtmin,
               double *shoc, double *t1, double *t2, double *t3)
{
    const double fr = c->fr, tinf = c->t_inf, rhoinf = c->rho_inf;
    for (int e = 0; e < nelem; ++e) {
        int n1 = inpoel[3 * e] - 1, n2 = inpoel[3 * e + 1] - 1, n3 = inpoel[3 * e + 2] - 1;
        const double *DNX = dNx + 3 * e, *DNY = dNy + 3 * e;
        double gm = (GAMM[n1] + GAMM[n2] + GAMM[n3]) / 3.0;
        double tau = 0.0, h_rgne = 0.0, h_rgn = 0.0, h_jgn = 0.0;
        double rho_elem = (U[4 * n1] + U[4 * n2] + U[4 * n3]) / 3.0;
        double vx = (VEL_X[n1] + VEL_X[n2] + VEL_X[n3]) / 3.0;
        double vy = (VEL_Y[n1] + VEL_Y[n2] + VEL_Y[n3]) / 3.0;
        double wx = (W_X[n1] + W_X[n2] + W_X[n3]) / 3.0;
        double wy = (W_Y[n1] + W_Y[n2] + W_Y[n3]) / 3.0;
        vx = vx - wx;
        vy = vy - wy;
        double vel2 = sqrt(vx * vx + vy * vy);
        double drx = U[4 * n1] * DNX[0] + U[4 * n2] * DNX[1] + U[4 * n3] * DNX[2];
        double dry = U[4 * n1] * DNY[0] + U[4 * n2] * DNY[1] + U[4 * n3] * DNY[2];
        double dr2 = sqrt(drx * drx + dry * dry) + 1.0e-20;
        double dtx = T[n1] * DNX[0] + T[n2] * DNX[1] + T[n3] * DNX[2];
        double dty = T[n1] * DNY[0] + T[n2] * DNY[1] + T[n3] * DNY[2];
        double dt2 = sqrt(dtx * dtx + dty * dty) + 1.0e-20;
        double dux = vel2 * DNX[0] + vel2 * DNX[1] + vel2 * DNX[2];
        double duy = vel2 * DNY[0] + vel2 * DNY[1] + vel2 * DNY[2];
        double du2 = sqrt(dux * dux + duy * duy) + 1.0e-20;
        double rtx = dtx / dt2, rty = dty / dt2;
        double rjx = drx / dr2, rjy = dry / dr2;
        double rux = dux / du2, ruy = duy / du2;
        double temp = (T[n1] + T[n2] + T[n3]) / 3.0;
        double cc = sqrt(gm * fr * temp);
        double smu = 110.0;
        double fmu = 0.017 * pow(temp / tinf, 1.5) * (tinf + smu) / (temp + smu);
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
        shoc[e] = (tr1 + tr1 * tr1) * 0.5 * (cc * cc) * zzz;
        double resumen = 1.0 / (tau * tau) + (2.0 / dtmin) * (2.0 / dtmin);
        double rrr = pow(resumen, -0.5);
        t1[e] = rrr;
        t2[e] = rrr;
        t3[e] = rrr;
        if (fmu != 0.0) {
            double ts3 = (h_rgn * h_rgn) / (4.0 * fmu / rhoinf);
            double ts3e = (h_rgne * h_rgne) / (4.0 * fmu / rhoinf);
            t2[e] = pow(resumen + 1.0 / (ts3 * ts3), -0.5);
            t3[e] = pow(resumen + 1.0 / (ts3e * ts3e), -0.5);
        }
    }
}

/* ------------------------------------------------------------------ */
/* calcRHS.f90:4-154                                                    */
/* ------------------------------------------------------------------ */
void orc_calcrhs(const orc_consts *c, int nelem, int npoin, double *rhs, const double *U,
                 const double *theta, const double *T, const double *dNx,
                 const double *dNy, const double *area, const double *shoc,
                 const double *dtl, const double *t1, const double *t2,
                 const double *t3, const int *inpoel)
{
    (void)npoin;
    const double g = c->gama, Cv = c->fcv, lambda_ref = c->fk, mu_ref = c->fmu;
    const double T_inf = c->t_inf, cte = c->cte;
    /* N(j,k): weight of node j at the midpoint of the edge opposite node k */
    static const double N[3][3] = {{0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};
#pragma omp parallel for
    for (int e = 0; e < nelem; ++e) {
        double rt[3][4] = {{0}};
        int p[3] = {inpoel[3 * e] - 1, inpoel[3 * e + 1] - 1, inpoel[3 * e + 2] - 1};
        double Nx[3] = {dNx[3 * e], dNx[3 * e + 1], dNx[3 * e + 2]};
        double Ny[3] = {dNy[3 * e], dNy[3 * e + 1], dNy[3 * e + 2]};
        double Ux[4], Uy[4];
        for (int q = 0; q < 4; ++q) {
            Ux[q] = U[4 * p[0] + q] * Nx[0] + U[4 * p[1] + q] * Nx[1] + U[4 * p[2] + q] * Nx[2];
            Uy[q] = U[4 * p[0] + q] * Ny[0] + U[4 * p[1] + q] * Ny[1] + U[4 * p[2] + q] * Ny[2];
        }
        double tau[3] = {t1[e], t2[e], t3[e]};
        double nu = shoc[e] * cte;
        double T_avg = (T[p[0]] + T[p[1]] + T[p[2]]) / 3.0;
        double mu = mu_ref * pow(T_avg / T_inf, 1.5) * (T_inf + 110) / (T_avg + 110);
        double lambda = lambda_ref * pow(T_avg / T_inf, 1.5) * (T_inf + 194) / (T_avg + 194);
        for (int k = 0; k < 3; ++k) {
            double Uk[4], th[4];
            for (int q = 0; q < 4; ++q) {
                Uk[q] = N[0][k] * U[4 * p[0] + q] + N[1][k] * U[4 * p[1] + q] + N[2][k] * U[4 * p[2] + q];
                th[q] = N[0][k] * theta[4 * p[0] + q] + N[1][k] * theta[4 * p[1] + q] +
                        N[2][k] * theta[4 * p[2] + q];
            }
            double rho = Uk[0];
            double v1 = Uk[1] / rho;
            double v2 = Uk[2] / rho;
            double en = Uk[3] / rho;
            double V_sq = v1 * v1 + v2 * v2;
            double A[4];
            A[0] = Ux[1] + Uy[2];
            A[1] = (1.0 / 2.0) * Ux[0] * (V_sq * (g - 1) - 2 * (v1 * v1)) - Ux[1] * v1 * (g - 3) -
                   Ux[2] * v2 * (g - 1) + Ux[3] * (g - 1) - Uy[0] * v1 * v2 + Uy[1] * v2 + Uy[2] * v1;
            A[2] = -(Ux[0] * v1 * v2) + Ux[1] * v2 + Ux[2] * v1 +
                   (1.0 / 2.0) * Uy[0] * (V_sq * (g - 1) - 2 * (v2 * v2)) - Uy[1] * v1 * (g - 1) -
                   Uy[2] * v2 * (g - 3) + Uy[3] * (g - 1);
            A[3] = Ux[0] * v1 * (V_sq * (g - 1) - en * g) -
                   (1.0 / 2.0) * Ux[1] * (V_sq * (g - 1) - 2 * en * g + 2 * (v1 * v1) * (g - 1)) -
                   Ux[2] * v1 * v2 * (g - 1) + Ux[3] * g * v1 + Uy[0] * v2 * (V_sq * (g - 1) - en * g) -
                   Uy[1] * v1 * v2 * (g - 1) -
                   (1.0 / 2.0) * Uy[2] * (V_sq * (g - 1) - 2 * en * g + 2 * (v2 * v2) * (g - 1)) +
                   Uy[3] * g * v2;
            double a[4];
            for (int q = 0; q < 4; ++q) a[q] = th[q] + A[q];
            double B1[4], B2[4];
            B1[0] = a[1];
            B1[1] = v1 * (-g + 3) * a[1] - v2 * (g - 1) * a[2] + (g - 1) * a[3] +
                    ((1.0 / 2.0) * V_sq * (g - 1) - (v1 * v1)) * a[0];
            B1[2] = -(v1 * v2 * a[0]) + v1 * a[2] + v2 * a[1];
            B1[3] = g * v1 * a[3] - v1 * v2 * (g - 1) * a[2] + v1 * (V_sq * (g - 1) - en * g) * a[0] +
                    (-((1.0 / 2.0) * V_sq * (g - 1)) + en * g - (v1 * v1) * (g - 1)) * a[1];
            B2[0] = a[2];
            B2[1] = -(v1 * v2 * a[0]) + v1 * a[2] + v2 * a[1];
            B2[2] = -(v1 * (g - 1) * a[1]) + v2 * (-g + 3) * a[2] + (g - 1) * a[3] +
                    ((1.0 / 2.0) * V_sq * (g - 1) - (v2 * v2)) * a[0];
            B2[3] = g * v2 * a[3] - v1 * v2 * (g - 1) * a[1] + v2 * (V_sq * (g - 1) - en * g) * a[0] +
                    (-((1.0 / 2.0) * V_sq * (g - 1)) + en * g - (v2 * v2) * (g - 1)) * a[2];
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < 4; ++q)
                    rt[j][q] = rt[j][q] + N[j][k] * A[q] + tau[j] * (Nx[j] * B1[q] + Ny[j] * B2[q]) +
                               nu * (Nx[j] * Ux[q] + Ny[j] * Uy[q]);
            if (mu_ref > 2.2250738585072014e-308) {
                double K1[4], K2[4];
                K1[1] = (2.0 / 3.0) * mu * (-(2 * Ux[0] * v1) + 2 * Ux[1] + Uy[0] * v2 - Uy[2]) / rho;
                K1[2] = mu * (-(Ux[0] * v2) + Ux[2] - Uy[0] * v1 + Uy[1]) / rho;
                K1[3] = (1.0 / 3.0) *
                        (Cv * mu * (-(Uy[0] * v1 * v2) + 3 * Uy[1] * v2 - 2 * Uy[2] * v1) -
                         Ux[0] * (Cv * mu * (3 * V_sq + (v1 * v1)) - 3 * lambda * (V_sq - en)) +
                         Ux[1] * v1 * (4 * Cv * mu - 3 * lambda) + 3 * Ux[2] * v2 * (Cv * mu - lambda) +
                         3 * Ux[3] * lambda) /
                        (Cv * rho);
                K2[1] = mu * (-(Ux[0] * v2) + Ux[2] - Uy[0] * v1 + Uy[1]) / rho;
                K2[2] = (2.0 / 3.0) * mu * (Ux[0] * v1 - Ux[1] - 2 * Uy[0] * v2 + 2 * Uy[2]) / rho;
                K2[3] = (1.0 / 3.0) *
                        (Cv * mu * (-(Ux[0] * v1 * v2) - 2 * Ux[1] * v2 + 3 * Ux[2] * v1) -
                         Uy[0] * (Cv * mu * (3 * V_sq + (v2 * v2)) - 3 * lambda * (V_sq - en)) +
                         3 * Uy[1] * v1 * (Cv * mu - lambda) + Uy[2] * v2 * (4 * Cv * mu - 3 * lambda) +
                         3 * Uy[3] * lambda) /
                        (Cv * rho);
                for (int j = 0; j < 3; ++j)
                    for (int q = 1; q < 4; ++q)
                        rt[j][q] = rt[j][q] + (Nx[j] * K1[q] + Ny[j] * K2[q]);
            }
        }
        for (int j = 0; j < 3; ++j)
            for (int q = 0; q < 4; ++q)
                rt[j][q] = rt[j][q] * area[e] * dtl[e] / 3.0;
        for (int q = 0; q < 4; ++q) {
#pragma omp atomic
            rhs[4 * p[0] + q] += rt[0][q];
#pragma omp atomic
            rhs[4 * p[1] + q] += rt[1][q];
#pragma omp atomic
            rhs[4 * p[2] + q] += rt[2][q];
        }
    }
}

/* subrutinas.f90:1036-1094 FUENTE */
void orc_fuente(int nelem, double *rhs, const double *U, const double *W_X,
                const double *W_Y, const double *dNx, const double *dNy,
                const double *area, const double *dtl, const int *inpoel)
{
    /* sp(j,k) = weight of local node j at quadrature point k */
    static const double sp[3][3] = {{0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}, {0.0, 0.5, 0.5}};
#pragma omp parallel for
    for (int e = 0; e < nelem; ++e) {
        int p[3] = {inpoel[3 * e] - 1, inpoel[3 * e + 1] - 1, inpoel[3 * e + 2] - 1};
        const double *nx = dNx + 3 * e, *ny = dNy + 3 * e;
        double Ux[4], Uy[4];
        for (int q = 0; q < 4; ++q) {
            Ux[q] = U[4 * p[0] + q] * nx[0] + U[4 * p[1] + q] * nx[1] + U[4 * p[2] + q] * nx[2];
            Uy[q] = U[4 * p[0] + q] * ny[0] + U[4 * p[1] + q] * ny[1] + U[4 * p[2] + q] * ny[2];
        }
        double AR = area[e] * dtl[e] / 3.0;
        double wx[3], wy[3];
        for (int k = 0; k < 3; ++k) {
            double sx = 0.0, sy = 0.0;
            for (int j = 0; j < 3; ++j) {
                sx = sx + sp[j][k] * W_X[p[j]];
                sy = sy + sp[j][k] * W_Y[p[j]];
            }
            wx[k] = sx;
            wy[k] = sy;
        }
        double rt[3][4];
        for (int j = 0; j < 3; ++j)
            for (int q = 0; q < 4; ++q)
                rt[j][q] = -(AR * (sp[j][0] * (Ux[q] * wx[0] + Uy[q] * wy[0]) +
                                   sp[j][1] * (Ux[q] * wx[1] + Uy[q] * wy[1]) +
                                   sp[j][2] * (Ux[q] * wx[2] + Uy[q] * wy[2])));
        for (int j = 0; j < 3; ++j)
            for (int q = 0; q < 4; ++q) {
#pragma omp atomic
                rhs[4 * p[j] + q] += rt[j][q];
            }
    }
}

/* ------------------------------------------------------------------ */
/* subrutinas.f90:645-849 RK (NGAS = 0 branch)                           */
/* ------------------------------------------------------------------ */
/* subrutinas.f90:66-85 NORMALVEL on the nnorm nodes NORMALES listed */
void orc_normalvel(int nnorm, const int *n_ipoin, const double *n_x, const double *n_y,
                   double *VEL_X, double *VEL_Y, const double *W_X, const double *W_Y)
{
    for (int i = 0; i < nnorm; ++i) {
        int ip = n_ipoin[i] - 1;
        double vx = VEL_X[ip], vy = VEL_Y[ip];
        double wx = W_X[ip], wy = W_Y[ip];
        double pp = -n_y[i] * (vx - wx) + n_x[i] * (vy - wy);
        VEL_X[ip] = -n_y[i] * pp + wx;
        VEL_Y[ip] = n_x[i] * pp + wy;
    }
}

void orc_rk(const orc_consts *c, int nelem, int npoin, int nrk, double dtmin,
            const double *dtl, orc_nodes *nd, orc_elems *el, const orc_bcs *bc)
{
    const double fr = c->fr;
    double *theta = (double *)calloc((size_t)4 * npoin, sizeof(double)); /* UN = 0 (:674) */
    for (int irk = 1; irk <= nrk; ++irk) {
        double rk_fact = 1.0 / (nrk + 1 - irk);
        if (irk == 1)
            orc_estab(c, nelem, el->inpoel, nd->U, nd->T, nd->GAMM, nd->VEL_X, nd->VEL_Y,
                      nd->W_X, nd->W_Y, el->dNx, el->dNy, dtmin, el->shoc, el->t1, el->t2, el->t3);
        memset(nd->RHS, 0, sizeof(double) * 4 * (size_t)npoin);
        orc_calcrhs(c, nelem, npoin, nd->RHS, nd->U, theta, nd->T, el->dNx, el->dNy, el->area,
                    el->shoc, dtl, el->t1, el->t2, el->t3, el->inpoel);
        orc_fuente(nelem, nd->RHS, nd->U, nd->W_X, nd->W_Y, el->dNx, el->dNy, el->area, dtl,
                   el->inpoel);
#pragma omp parallel for
        for (int p = 0; p < npoin; ++p)
            for (int q = 0; q < 4; ++q)
                nd->U1[4 * p + q] = nd->U[4 * p + q] - rk_fact / nd->M[p] * nd->RHS[4 * p + q];
#pragma omp parallel for
        for (int p = 0; p < npoin; ++p) {
            double rho = nd->U1[4 * p];
            nd->RHO[p] = rho;
            nd->VEL_X[p] = nd->U1[4 * p + 1] / rho;
            nd->VEL_Y[p] = nd->U1[4 * p + 2] / rho;
            nd->E[p] = nd->U1[4 * p + 3] / rho;
            double vel2 = nd->VEL_X[p] * nd->VEL_X[p] + nd->VEL_Y[p] * nd->VEL_Y[p];
            nd->P[p] = rho * (nd->GAMM[p] - 1.0) * (nd->E[p] - 0.5 * vel2);
            nd->T[p] = nd->P[p] / (rho * fr);
            nd->RMACH[p] = sqrt(vel2 / (nd->T[p] * nd->GAMM[p] * fr));
        }
        /* FIXVEL :601-616 */
        for (int i = 0; i < bc->nfixv; ++i) {
            int j = bc->ifixv_node[i] - 1;
            nd->VEL_X[j] = bc->rfixv_valuex[i];
            nd->VEL_Y[j] = bc->rfixv_valuey[i];
        }
        /* NORMALVEL :66-85 */
        orc_normalvel(bc->nnorm, bc->n_ipoin, bc->n_x, bc->n_y, nd->VEL_X, nd->VEL_Y, nd->W_X,
                      nd->W_Y);
        /* FIX :618-643 */
        for (int i = 0; i < bc->nfixrho; ++i)
            nd->RHO[bc->ifixrho_node[i] - 1] = bc->rfixrho_value[i];
        for (int i = 0; i < bc->nfixt; ++i) {
            int j = bc->ifixt_node[i] - 1;
            double gm = nd->GAMM[j] - 1.0;
            nd->T[j] = bc->rfixt_value[i];
            nd->E[j] = nd->T[j] * fr / gm + 0.5 * (nd->VEL_X[j] * nd->VEL_X[j] + nd->VEL_Y[j] * nd->VEL_Y[j]);
        }
#pragma omp parallel for
        for (int p = 0; p < npoin; ++p) {
            nd->U1[4 * p + 0] = nd->RHO[p];
            nd->U1[4 * p + 1] = nd->VEL_X[p] * nd->RHO[p];
            nd->U1[4 * p + 2] = nd->VEL_Y[p] * nd->RHO[p];
            nd->U1[4 * p + 3] = nd->E[p] * nd->RHO[p];
        }
    }
    free(theta);
}

/* subrutinas.f90:7-64 NORMALES */
int orc_normales(int npoin, int nwall, const int *wall, const double *X, const double *Y,
                 int *n_ipoin, double *n_x, double *n_y)
{
    double *nlx = (double *)calloc((size_t)npoin, sizeof(double));
    double *nly = (double *)calloc((size_t)npoin, sizeof(double));
    double *den = (double *)calloc((size_t)npoin, sizeof(double));
    for (int w = 0; w < nwall; ++w) {
        int a = wall[2 * w] - 1, b = wall[2 * w + 1] - 1;
        double lx = Y[b] - Y[a];
        double ly = -(X[b] - X[a]);
        double l = sqrt(lx * lx + ly * ly);
        nlx[a] += lx; nly[a] += ly; den[a] += l;
        nlx[b] += lx; nly[b] += ly; den[b] += l;
    }
    int m = 0;
    for (int p = 0; p < npoin; ++p) {
        if (den[p] > 1.0e-6) {
            double lx = nlx[p] / den[p], ly = nly[p] / den[p];
            double ln = sqrt(lx * lx + ly * ly);
            if (ln > 0.2) {
                n_ipoin[m] = p + 1;
                n_x[m] = lx / ln;
                n_y[m] = ly / ln;
                m += 1;
            }
        }
    }
    free(nlx); free(nly); free(den);
    return m;
}

/* ------------------------------------------------------------------ */
/* mLaplace.f90:99-150 laplace and :188-200 mu                          */
/* ------------------------------------------------------------------ */
static double quality_mu(double X1, double X2, double X3, double Y1, double Y2, double Y3)
{
    const double TWOSQRT3 = 3.46410161513775;
    double ar = X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3);
    double l1 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
    double l2 = (X1 - X3) * (X1 - X3) + (Y1 - Y3) * (Y1 - Y3);
    double l3 = (X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1);
    double l = l1 + l2 + l3;
    return TWOSQRT3 * ar / l;
}

void orc_laplace(int nelem, int npoin, const int *inpoel, const double *X, const double *Y,
                 const double *dNx, const double *dNy, const int *esup1,
                 const int *esup2, const int *lap_rowptr, const int *lap_idx,
                 double *lap_sparse, double *lap_diag)
{
    (void)nelem;
    memset(lap_sparse, 0, sizeof(double) * (size_t)lap_rowptr[npoin]);
#pragma omp parallel for
    for (int ip = 1; ip <= npoin; ++ip) {
        for (int s = esup2[ip - 1] + 1; s <= esup2[ip]; ++s) {
            int e = esup1[s - 1] - 1;
            const int *n = inpoel + 3 * e;
            double mu = quality_mu(X[n[0] - 1], X[n[1] - 1], X[n[2] - 1], Y[n[0] - 1],
                                   Y[n[1] - 1], Y[n[2] - 1]);
            double q = 1 / (mu * mu);
            for (int i = 0; i < 3; ++i) {
                if (n[i] != ip) continue;
                for (int j = 0; j < 3; ++j) {
                    int kp = n[j];
                    for (int k = lap_rowptr[ip - 1] + 1; k <= lap_rowptr[ip]; ++k)
                        if (lap_idx[k - 1] == kp)
                            lap_sparse[k - 1] = lap_sparse[k - 1] +
                                (dNx[3 * e + i] * dNx[3 * e + j] + dNy[3 * e + i] * dNy[3 * e + j]) * q;
                }
            }
        }
        lap_diag[ip - 1] = lap_sparse[lap_rowptr[ip - 1]];
    }
}

/* ------------------------------------------------------------------ */
/* biconjGrad.f90:8-62 biCG (Jacobi-preconditioned CG)                   */
/* ------------------------------------------------------------------ */
static void spmv(int n, const double *A, const int *idx, const int *rowptr, const double *v,
                 double *y)
{
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
        double dot = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j)
            dot = dot + A[j] * v[idx[j] - 1];
        y[i] = dot;
    }
}

static double vecdot(int n, const double *x, const double *y)
{
    double res = 0.0;
#pragma omp parallel for reduction(+ : res)
    for (int i = 0; i < n; ++i)
        res = res + x[i] * y[i];
    return res;
}

int orc_bicg(int npoin, const double *A, const int *idx, const int *rowptr,
             const double *diag, double *x, const double *b, const double *x_fix,
             const int *fidx, int nfix, double *rz)
{
    double *y = (double *)malloc(sizeof(double) * (size_t)npoin);
    double *p = (double *)malloc(sizeof(double) * (size_t)npoin);
    double *r = (double *)malloc(sizeof(double) * (size_t)npoin);
    double *z = (double *)malloc(sizeof(double) * (size_t)npoin);
    const double tol = 1.0e-10;
    int k = 0;
    double err_old, err_new, alfa, beta, py;
    for (int i = 0; i < nfix; ++i) x[fidx[i] - 1] = 1.0 * x_fix[i];          /* copy1 */
    spmv(npoin, A, idx, rowptr, x, y);
    for (int i = 0; i < nfix; ++i) y[fidx[i] - 1] = 1.0e30 * x[fidx[i] - 1];  /* copy2 */
    for (int i = 0; i < npoin; ++i) r[i] = -1.0 * y[i] + b[i];              /* vecsum */
    for (int i = 0; i < nfix; ++i) r[fidx[i] - 1] = 0.0;                     /* assign2 */
    double rr = vecdot(npoin, r, r);
    if (rr < tol) {
        *rz = rr;
        free(y); free(p); free(r); free(z);
        return 0;
    }
    for (int i = 0; i < npoin; ++i) p[i] = r[i] / diag[i];
    err_new = vecdot(npoin, r, p);
    spmv(npoin, A, idx, rowptr, p, y);
    for (int i = 0; i < nfix; ++i) y[fidx[i] - 1] = 1.0e30 * p[fidx[i] - 1];
    py = vecdot(npoin, p, y);
    alfa = err_new / py;
    for (int i = 0; i < npoin; ++i) x[i] = alfa * p[i] + x[i];
    err_old = err_new;
    while (fabs(err_old) > tol && k < 1000) {
        k = k + 1;
        for (int i = 0; i < npoin; ++i) r[i] = -alfa * y[i] + r[i];
        for (int i = 0; i < npoin; ++i) z[i] = r[i] / diag[i];
        err_new = vecdot(npoin, r, z);
        beta = err_new / err_old;
        for (int i = 0; i < npoin; ++i) p[i] = beta * p[i] + z[i];
        spmv(npoin, A, idx, rowptr, p, y);
        for (int i = 0; i < nfix; ++i) y[fidx[i] - 1] = 1.0e30 * p[fidx[i] - 1];
        py = vecdot(npoin, p, y);
        alfa = err_new / py;
        for (int i = 0; i < npoin; ++i) x[i] = alfa * p[i] + x[i];
        err_old = err_new;
    }
    *rz = err_old;
    free(y); free(p); free(r); free(z);
    return k;
}

/* ------------------------------------------------------------------ */
/* meshMove.f90:153-194 FORCES and :369-392 TRANSF (single set)          */
/* ------------------------------------------------------------------ */
void orc_forces(int nedge, const int *iset, const double *X, const double *Y,
                const double *P, double xref, double yref, double *fx, double *fy, double *rm)
{
    double FX = 0.0, FY = 0.0, RM = 0.0;
    for (int i = 0; i < nedge; ++i) {
        int n1 = iset[3 * i] - 1, n2 = iset[3 * i + 1] - 1;
        double d_press = (P[n1] + P[n2]) / 2.0;
        double rlx = X[n1] - X[n2];
        double rly = Y[n2] - Y[n1];
        double dfx = d_press * rly;
        double dfy = d_press * rlx;
        FX = FX + dfx;
        FY = FY + dfy;
        double xc = (X[n1] + X[n2]) / 2.0;
        double yc = (Y[n2] + Y[n1]) / 2.0;
        RM = RM + dfy * (xc - xref) - dfx * (yc - yref);
    }
    *fx = FX; *fy = FY; *rm = RM;
}

/* ------------------------------------------------------------------ */
/* ns2DComp.ALE.f90:819-893 FORCE_VISC (single set): the traction of the  */
/* full stress tensor, -p I + viscous, with the element-average pressure   */
/* and the set element's P1 velocity gradients, Sutherland mu at the edge  */
/* average temperature, times the edge length, summed in edge order.  The  */
/* edge normal (RLX, RLY) = (Y2-Y1, -(X2-X1)) / |edge| is that of FORCES.  */
/* SKIN.DAT (per-edge skin friction written to a file) is not restated.    */
/* ------------------------------------------------------------------ */
void orc_force_visc(int nedge, const int *iset, const int *inpoel, const double *X,
                    const double *Y, const double *P, const double *T, const double *VEL_X,
                    const double *VEL_Y, const double *dNx, const double *dNy, double t_inf,
                    double *f_vx, double *f_vy)
{
    double FVX = 0.0, FVY = 0.0;
    for (int ii = 0; ii < nedge; ++ii) {
        int nn1 = iset[3 * ii] - 1, nn2 = iset[3 * ii + 1] - 1, ie = iset[3 * ii + 2] - 1;
        double temp = (T[nn1] + T[nn2]) / 2.0;           /* :844 */
        double smu = 110.0;
        double fmu = 0.017 * pow(temp / t_inf, 1.5) * (t_inf + smu) / (temp + smu);
        double rly = -(X[nn2] - X[nn1]);                 /* :849-851 */
        double rlx = Y[nn2] - Y[nn1];
        double rmod = sqrt(rlx * rlx + rly * rly);
        rlx = rlx / rmod;
        rly = rly / rmod;
        double dux = 0.0, duy = 0.0, dvx = 0.0, dvy = 0.0, press = 0.0;
        for (int jj = 0; jj < 3; ++jj) {                 /* :860-869 */
            int nn = inpoel[3 * ie + jj] - 1;
            double nx = dNx[3 * ie + jj], ny = dNy[3 * ie + jj];
            dux = dux + nx * VEL_X[nn];
            duy = duy + ny * VEL_X[nn];
            dvx = dvx + nx * VEL_Y[nn];
            dvy = dvy + ny * VEL_Y[nn];
            press = press + P[nn];
        }
        press = press / 3.0;
        double txx = -press - fmu * (2.0 / 3.0 * (dux + dvy) - 2.0 * dux);   /* :873-877 */
        double txy = fmu * (duy + dvx);
        double tyx = txy;
        double tyy = -press - fmu * (2.0 / 3.0 * (dux + dvy) - 2.0 * dvy);
        double ttx = txx * rlx + txy * rly;              /* :880-881 */
        double tty = tyx * rlx + tyy * rly;
        FVX = FVX + ttx * rmod;                          /* :885-886 */
        FVY = FVY + tty * rmod;
    }
    *f_vx = FVX;
    *f_vy = FVY;
}

void orc_transf(int nedge, const int *iset, const double *X, const double *Y,
                double alpha, double yposr, double xref, double yref,
                double *DXPOS, double *DYPOS)
{
    double ca = cos(alpha), sa = sin(alpha);
    for (int i = 0; i < nedge; ++i)
        for (int jj = 0; jj < 2; ++jj) {
            int n = iset[3 * i + jj] - 1;
            double distx = X[n] - xref;
            double disty = Y[n] - yref + yposr;
            DXPOS[n] = ca * distx + sa * disty - distx;
            DYPOS[n] = -(sa * distx) + ca * disty - disty + yposr;
        }
}

/* ------------------------------------------------------------------ */
/* gcl.f90:8-46 main                                                     */
/* ------------------------------------------------------------------ */
void orc_gcl(int nelem, int npoin, const int *inpoel, double *M, const double *W_x,
             const double *W_y, const double *W_x_old, const double *W_y_old,
             const double *dNx, const double *dNy, const double *area,
             const double *area_old, double dt, int mode)
{
    double *tot1 = (double *)calloc((size_t)npoin, sizeof(double));
    double *tot2 = (double *)calloc((size_t)npoin, sizeof(double));
    for (int e = 0; e < nelem; ++e) {
        const int *n = inpoel + 3 * e;
        const double *nx = dNx + 3 * e, *ny = dNy + 3 * e;
        const double *wy = mode ? W_y : W_x;
        const double *wyo = mode ? W_y_old : W_x_old;
        double divW = nx[0] * W_x[n[0] - 1] + nx[1] * W_x[n[1] - 1] + nx[2] * W_x[n[2] - 1] +
                      ny[0] * wy[n[0] - 1] + ny[1] * wy[n[1] - 1] + ny[2] * wy[n[2] - 1];
        double divWo = nx[0] * W_x_old[n[0] - 1] + nx[1] * W_x_old[n[1] - 1] + nx[2] * W_x_old[n[2] - 1] +
                       ny[0] * wyo[n[0] - 1] + ny[1] * wyo[n[1] - 1] + ny[2] * wyo[n[2] - 1];
        for (int i = 0; i < 3; ++i) {
            tot1[n[i] - 1] = tot1[n[i] - 1] + divW * area[e] / 3.0;
            tot2[n[i] - 1] = tot2[n[i] - 1] + divWo * area_old[e] / 3.0;
        }
    }
    for (int p = 0; p < npoin; ++p)
        M[p] = M[p] + dt * (tot1[p] + tot2[p]) / 2.0;
    free(tot1);
    free(tot2);
}

/* ------------------------------------------------------------------ */
/* smoothing.f90 smoothing_mod: start-up mesh smoothing, sequential      */
/* ------------------------------------------------------------------ */
#define SM_MAXE 20 /* MAX_ELEM_PER_NODE, smoothing.f90:6 */

typedef struct {
    int npoin;
    const int *inpoel, *esup1, *esup2, *psup1, *psup2;
    double *X, *Y;
    int *smoothable;
    const double *tol_move; /* NULL: SWITCH_TOL_MOVE false */
    double hmin_global;
} sm_ctx;

/* smoothing.f90:357-365 h */
static double sm_h(double X1, double X2, double X3, double Y1, double Y2, double Y3)
{
    double area = X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3);
    double d1 = fabs(X3 - X2) + fabs(Y3 - Y2);
    double d2 = fabs(X1 - X3) + fabs(Y1 - Y3);
    double d3 = fabs(X2 - X1) + fabs(Y2 - Y1);
    double d = d1 + d2 + d3;
    return fabs(area) / d;
}

static double sm_mu_elem(const sm_ctx *c, int e)
{
    const int *n = c->inpoel + 3 * e;
    return quality_mu(c->X[n[0] - 1], c->X[n[1] - 1], c->X[n[2] - 1], c->Y[n[0] - 1],
                      c->Y[n[1] - 1], c->Y[n[2] - 1]);
}

/* smoothing.f90:174-197 getMu_vec; returns the count, min_idx 1-based */
static int sm_mu_vec(const sm_ctx *c, int ip, double *mu_vec, int *min_idx)
{
    int cnt = 0;
    double mu_min = 1;
    if (min_idx) *min_idx = 1;
    for (int s = c->esup2[ip - 1] + 1; s <= c->esup2[ip]; ++s) {
        double mu1 = sm_mu_elem(c, c->esup1[s - 1] - 1);
        int k = s - c->esup2[ip - 1];
        mu_vec[k - 1] = mu1;
        cnt = k;
        if (min_idx && mu1 < mu_min) {
            mu_min = mu1;
            *min_idx = k;
        }
    }
    return cnt;
}

/* smoothing.f90:199-215 getMu_min */
static double sm_mu_min(const sm_ctx *c, int ip)
{
    double m = 1;
    for (int s = c->esup2[ip - 1] + 1; s <= c->esup2[ip]; ++s) {
        double mu1 = sm_mu_elem(c, c->esup1[s - 1] - 1);
        if (mu1 <= m) m = mu1;
    }
    return m;
}

/* smoothing.f90:272-287 getH_min */
static double sm_h_min(const sm_ctx *c, int ip)
{
    double m = 1;
    for (int s = c->esup2[ip - 1] + 1; s <= c->esup2[ip]; ++s) {
        const int *n = c->inpoel + 3 * (c->esup1[s - 1] - 1);
        double h1 = sm_h(c->X[n[0] - 1], c->X[n[1] - 1], c->X[n[2] - 1], c->Y[n[0] - 1],
                         c->Y[n[1] - 1], c->Y[n[2] - 1]);
        if (h1 < m) m = h1;
    }
    return m;
}

/* smoothing.f90:77-111 laplacianSmoothing, one pass in node order */
static void sm_laplacian(sm_ctx *c, const int *fixed)
{
    for (int ip = 1; ip <= c->npoin; ++ip) {
        if (!(c->smoothable[ip - 1] && !fixed[ip - 1])) continue;
        double xn = 0, yn = 0;
        double mu_old = sm_mu_min(c, ip);
        double xo = c->X[ip - 1], yo = c->Y[ip - 1];
        int n = c->psup2[ip] - c->psup2[ip - 1];
        for (int s = c->psup2[ip - 1] + 1; s <= c->psup2[ip]; ++s) {
            xn = xn + c->X[c->psup1[s - 1] - 1];
            yn = yn + c->Y[c->psup1[s - 1] - 1];
        }
        xn = xn / n;
        yn = yn / n;
        if (c->tol_move) {
            double dX = xn - c->X[ip - 1], dY = yn - c->Y[ip - 1];
            double tm = c->tol_move[ip - 1];
            c->X[ip - 1] = c->X[ip - 1] + dX / (fabs(dX) + fabs(dY)) * tm / 2 * 3.0;
            c->Y[ip - 1] = c->Y[ip - 1] + dY / (fabs(dX) + fabs(dY)) * tm / 2 * 3.0;
        } else {
            c->X[ip - 1] = xn;
            c->Y[ip - 1] = yn;
        }
        if (sm_mu_min(c, ip) < mu_old) { /* if worse, restore */
            c->X[ip - 1] = xo;
            c->Y[ip - 1] = yo;
        }
    }
}

/* smoothing.f90:217-236 getG: one-sided differences of every element's mu */
static void sm_grad(sm_ctx *c, int ip, const double *mu_vec, double *gx, double *gy)
{
    double delta = c->hmin_global * 1e-2;
    double xo = c->X[ip - 1];
    c->X[ip - 1] = c->X[ip - 1] + delta;
    int cnt = sm_mu_vec(c, ip, gx, NULL);
    for (int k = 0; k < cnt; ++k) gx[k] = (gx[k] - mu_vec[k]) / delta;
    c->X[ip - 1] = xo;
    double yo = c->Y[ip - 1];
    c->Y[ip - 1] = c->Y[ip - 1] + delta;
    sm_mu_vec(c, ip, gy, NULL);
    for (int k = 0; k < cnt; ++k) gy[k] = (gy[k] - mu_vec[k]) / delta;
    c->Y[ip - 1] = yo;
}

/* smoothing.f90:238-270 getStep */
static double sm_step(const sm_ctx *c, int ip, const double *gx, const double *gy,
                      const double *mu_vec, int min_idx, int cnt)
{
    double gxm = gx[min_idx - 1], gym = gy[min_idx - 1], mum = mu_vec[min_idx - 1];
    double g2 = gxm * gxm + gym * gym;
    double step;
    if (c->tol_move)
        step = c->tol_move[ip - 1] / ((fabs(gxm) + fabs(gym)) * 100.0);
    else
        step = sm_h_min(c, ip) * 3.0 / (fabs(gxm) + fabs(gym));
    for (int i = 0; i < cnt; ++i) {
        double gg = gxm * gx[i] + gym * gy[i];
        if (gg < 0) {
            double step1 = (mu_vec[i] - mum) / (g2 - gg);
            if (step1 < step) step = step1;
        }
    }
    return step;
}

/* smoothing.f90:113-171 moveIpoin; returns d_move */
static double sm_move(sm_ctx *c, int ip, double *mu_vec, int min_idx)
{
    double gx[SM_MAXE], gy[SM_MAXE];
    double x0 = c->X[ip - 1], y0 = c->Y[ip - 1];
    int cnt = c->esup2[ip] - c->esup2[ip - 1];
    for (int it = 0; it < 2; ++it) { /* MITER */
        double mu_min = mu_vec[min_idx - 1];
        double xo = c->X[ip - 1], yo = c->Y[ip - 1];
        sm_grad(c, ip, mu_vec, gx, gy);
        double step = sm_step(c, ip, gx, gy, mu_vec, min_idx, cnt);
        int accepted = 0;
        for (int j = 0; j < 8; ++j) { /* NTRY: the tries accumulate */
            double dx = step * gx[min_idx - 1], dy = step * gy[min_idx - 1];
            c->X[ip - 1] = c->X[ip - 1] + dx;
            c->Y[ip - 1] = c->Y[ip - 1] + dy;
            int mi;
            sm_mu_vec(c, ip, mu_vec, &mi);
            if (mu_vec[mi - 1] > mu_min * 1.0) {
                min_idx = mi;
                accepted = 1;
                break;
            }
            step = .5 * step;
        }
        if (!accepted) {
            c->X[ip - 1] = xo;
            c->Y[ip - 1] = yo;
            break;
        }
    }
    double dxm = c->X[ip - 1] - x0, dym = c->Y[ip - 1] - y0;
    double d_move = dxm * dxm + dym * dym;
    if (d_move > DBL_MIN) /* update_list, smoothing.f90:289-298 */
        for (int s = c->psup2[ip - 1] + 1; s <= c->psup2[ip]; ++s)
            c->smoothable[c->psup1[s - 1] - 1] = 1;
    return d_move;
}

int orc_smoothing(int nelem, int npoin, const int *inpoel, const int *esup1, const int *esup2,
                  const int *psup1, const int *psup2, double *X, double *Y, const int *fixed,
                  const double *tol_move, double *hmin_out, int max_sweeps)
{
    sm_ctx c = {npoin, inpoel, esup1, esup2, psup1, psup2, X, Y, NULL, tol_move, 1.0};
    for (int ip = 1; ip <= npoin; ++ip)
        if (esup2[ip] - esup2[ip - 1] > SM_MAXE) return -1;
    int *sm = (int *)calloc((size_t)npoin, sizeof(int));
    c.smoothable = sm;
    /* checkMesh, smoothing.f90:367-385 */
    for (int e = 0; e < nelem; ++e) {
        const int *n = inpoel + 3 * e;
        if (sm_mu_elem(&c, e) < .85)
            for (int k = 0; k < 3; ++k) sm[n[k] - 1] = 1;
        double h1 = sm_h(X[n[0] - 1], X[n[1] - 1], X[n[2] - 1], Y[n[0] - 1], Y[n[1] - 1],
                         Y[n[2] - 1]);
        if (h1 < c.hmin_global) c.hmin_global = h1;
    }
    if (hmin_out) *hmin_out = c.hmin_global;
    int any = 0;
    for (int ip = 0; ip < npoin; ++ip) any |= sm[ip];
    if (!any) {
        free(sm);
        return 0;
    }
    /* smoothing.f90:46-65 */
    for (int it = 0; it < 2; ++it) sm_laplacian(&c, fixed);
    double tol_dist = 1e-2 * c.hmin_global;
    int iter;
    double mu_vec[SM_MAXE];
    for (iter = 1; iter <= max_sweeps; ++iter) { /* NITER = 100 */
        double d_max = 0;
        for (int ip = 1; ip <= npoin; ++ip) {
            if (!(sm[ip - 1] && !fixed[ip - 1])) continue;
            sm[ip - 1] = 0;
            int mi;
            sm_mu_vec(&c, ip, mu_vec, &mi);
            if (mu_vec[mi - 1] < .85) {
                double d = sm_move(&c, ip, mu_vec, mi);
                if (d > d_max) d_max = d;
            }
        }
        if (d_max < tol_dist) break;
    }
    free(sm);
    return iter > max_sweeps ? max_sweeps : iter;
}
