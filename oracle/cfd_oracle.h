/* Test infrastructure only.
 *
 * CPU restatement of the reference solver's hot path (the oracle).  Each
 * function follows one reference routine loop by loop, in the same
 * floating-point evaluation order, so that a sequential build
 * (-ffp-contract=off, one OpenMP thread) reproduces the reference bit for
 * bit.  Only tests and the benchmark's CPU-baseline leg may call it.
 *
 * Conventions follow the reference: node and element ids stored in arrays
 * are 1-based, offset arrays are 0-based with a leading 0, and
 * U(4,npoin) / inpoel(3,nelem) / dNx(3,nelem) are node- or element-major.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double gama, fr, fmu, fk, fcv;          /* gas constants (InputData)     */
    double t_inf, rho_inf, u_inf, v_inf;    /* free stream                   */
    double cte, fsafe;                       /* shock-capturing 1/CTE, CFL    */
} orc_consts;

/* OpenMP thread count: 1 gives the reference's sequential order */
void orc_set_threads(int n);
int orc_get_threads(void);

/* pointNeighbor.f90 / mLaplace.f90 initialize */
void orc_get_esup(int nelem, int npoin, const int *inpoel, int *esup1, int *esup2);
int  orc_get_psup(int nelem, int npoin, const int *inpoel, const int *esup1,
                  const int *esup2, int *psup1, int *psup2);
void orc_lap_structure(int npoin, const int *psup1, const int *psup2,
                       int *lap_rowptr, int *lap_idx);

/* subrutinas.f90 deriv / MASAS / deltat / ESTAB / FUENTE / RK */
void orc_deriv(int nelem, const int *inpoel, const double *X, const double *Y,
               double *area, double *dNx, double *dNy, double *HH, double *HHX,
               double *HHY, double *hmin);
void orc_masas(int nelem, int npoin, const int *inpoel, const double *area, double *M);
void orc_deltat(const orc_consts *c, int nelem, const int *inpoel, const double *area,
                const double *T, const double *VEL_X, const double *VEL_Y,
                const double *W_X, const double *W_Y, double *DT, double *dtmin);
void orc_estab(const orc_consts *c, int nelem, const int *inpoel, const double *U,
               const double *T, const double *GAMM, const double *VEL_X,
               const double *VEL_Y, const double *W_X, const double *W_Y,
               const double *dNx, const double *dNy, double dtmin,
               double *shoc, double *t1, double *t2, double *t3);
void orc_calcrhs(const orc_consts *c, int nelem, int npoin, double *rhs, const double *U,
                 const double *theta, const double *T, const double *dNx,
                 const double *dNy, const double *area, const double *shoc,
                 const double *dtl, const double *t1, const double *t2,
                 const double *t3, const int *inpoel);
void orc_fuente(int nelem, double *rhs, const double *U, const double *W_X,
                const double *W_Y, const double *dNx, const double *dNy,
                const double *area, const double *dtl, const int *inpoel);

/* boundary lists as left by the loader and NORMALES */
typedef struct {
    int nfixv;   const int *ifixv_node;   const double *rfixv_valuex, *rfixv_valuey;
    int nnorm;   const int *n_ipoin;      const double *n_x, *n_y;
    int nfixrho; const int *ifixrho_node; const double *rfixrho_value;
    int nfixt;   const int *ifixt_node;   const double *rfixt_value;
} orc_bcs;

/* Node arrays touched by RK (module state in the reference). */
typedef struct {
    double *U, *U1, *RHS, *T, *VEL_X, *VEL_Y, *W_X, *W_Y, *RHO, *E, *P, *RMACH;
    const double *M, *GAMM;
} orc_nodes;

/* Elements arrays read by RK. */
typedef struct {
    const int *inpoel;
    const double *dNx, *dNy, *area;
    double *shoc, *t1, *t2, *t3;
} orc_elems;

void orc_rk(const orc_consts *c, int nelem, int npoin, int nrk, double dtmin,
            const double *dtl, orc_nodes *nd, orc_elems *el, const orc_bcs *bc);

/* subrutinas.f90 NORMALES: returns m (number of wall nodes kept) */
int orc_normales(int npoin, int nwall, const int *wall, const double *X, const double *Y,
                 int *n_ipoin, double *n_x, double *n_y);

/* mLaplace.f90 laplace (values) */
void orc_laplace(int nelem, int npoin, const int *inpoel, const double *X, const double *Y,
                 const double *dNx, const double *dNy, const int *esup1,
                 const int *esup2, const int *lap_rowptr, const int *lap_idx,
                 double *lap_sparse, double *lap_diag);

/* biconjGrad.f90 biCG: returns iteration count; *rz = final |r.z| */
int orc_bicg(int npoin, const double *spMtx, const int *spIdx, const int *spRowptr,
             const double *diag, double *x, const double *b, const double *x_fix,
             const int *x_fixIdx, int nfix, double *rz);

/* meshMove.f90 FORCES (one set) and TRANSF (one set) */
void orc_forces(int nedge, const int *iset /* (nedge,3): n1 n2 elem */,
                const double *X, const double *Y, const double *P, double xref,
                double yref, double *fx, double *fy, double *rm);
void orc_normalvel(int nnorm, const int *n_ipoin, const double *n_x, const double *n_y,
                   double *VEL_X, double *VEL_Y, const double *W_X, const double *W_Y);
void orc_force_visc(int nedge, const int *iset, const int *inpoel, const double *X,
                    const double *Y, const double *P, const double *T, const double *VEL_X,
                    const double *VEL_Y, const double *dNx, const double *dNy, double t_inf,
                    double *f_vx, double *f_vy);
void orc_transf(int nedge, const int *iset, const double *X, const double *Y,
                double alpha, double yposr, double xref, double yref,
                double *DXPOS, double *DYPOS);

/* gcl.f90 main: mode 0 = as written (y-divergence uses W_x), 1 = corrected */
void orc_gcl(int nelem, int npoin, const int *inpoel, double *M, const double *W_x,
             const double *W_y, const double *W_x_old, const double *W_y_old,
             const double *dNx, const double *dNy, const double *area,
             const double *area_old, double dt, int mode);

/* smoothing.f90 smoothing (start-up mesh smoothing) in the reference's node
 * order: X, Y in place; fixed(npoin) nonzero = fixed; tol_move(npoin) or
 * NULL (setTol_move not called).  Returns the optimisation sweeps run (0:
 * nothing smoothable), -1 when a node has more than 20 elements (the
 * reference's MAX_ELEM_PER_NODE); *hmin_out = H_MIN_GLOBAL.  max_sweeps: the
 * reference's NITER = 100 (fewer only to look at intermediate states). */
int orc_smoothing(int nelem, int npoin, const int *inpoel, const int *esup1, const int *esup2,
                  const int *psup1, const int *psup2, double *X, double *Y, const int *fixed,
                  const double *tol_move, double *hmin_out, int max_sweeps);

#ifdef __cplusplus
}
#endif
#endif
