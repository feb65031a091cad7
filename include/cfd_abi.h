/* cfd_abi.h -- C ABI of libcfd_amd.so, the gfx950 (MI355X) implementation of
 * the chanshing/cfd SUPG/ALE hot path.
 *
 * Two layers:
 *  1. Host-pointer entries, one per reference routine on the path, with the
 *     reference's argument meaning and accumulate/overwrite semantics.  They
 *     take plain host arrays, run the HIP kernels on the calling thread's
 *     mesh context and return when the results are back in host memory.
 *     A Fortran shim module with the reference module/procedure names
 *     (cfd_amd/fortran/cfd_shims.f90) forwards to them via iso_c_binding.
 *  2. A device-resident session that replays the main program's time step
 *     (ns2DComp.ALE.f90:138-281) with all state in HBM.
 *
 * Conventions (the reference's): node and element ids stored in arrays are
 * 1-based, offset arrays are 0-based with a leading 0, arrays are Fortran
 * column-major: U(4,npoin) is 4 contiguous doubles per node, inpoel(3,nelem)
 * and dNx(3,nelem) are 3 contiguous values per element.  Integers are int32,
 * reals are binary64.
 *
 * Errors: every entry returns 0 on success and a non-zero code otherwise;
 * cfd_last_error() returns the message of the last failure on this thread.
 * The reference has no status codes (it STOPs or, for the PCG, silently
 * returns after 1000 iterations); the shims turn non-zero codes into
 * ERROR STOP.  There is no CPU fallback: without a usable GPU every entry
 * fails with CFD_ERR_DEVICE.
 */
#ifndef CFD_ABI_H
#define CFD_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_ABI_VERSION 5

enum {
    CFD_OK = 0,
    CFD_ERR_ARG = 1,      /* invalid argument or size                     */
    CFD_ERR_DEVICE = 2,   /* HIP runtime / device failure, no GPU          */
    CFD_ERR_MESH = 3,     /* mesh rejected (ids out of range, degenerate)  */
    CFD_ERR_STATE = 4     /* call out of sequence (e.g. GCL before putW)   */
};

typedef struct cfd_mesh cfd_mesh;        /* tables + scratch for one mesh  */
typedef struct cfd_session cfd_session;  /* device-resident solver state   */

/* InputData constants the path reads (dataLoader.f90:5-8, derived at :59-64).
 * cte is the shock-capturing factor after the loader's CTE := 1/CTE. */
typedef struct {
    double gama, fr, fmu, fk, fcv;
    double t_inf, rho_inf, u_inf, v_inf;
    double cte, fsafe;
} cfd_consts;

int cfd_abi_version(void);
const char *cfd_last_error(void);
int cfd_device_count(int *count);

/* ---- mesh context ------------------------------------------------------ */
/* Builds esup, psup, the Laplace CSR structure and the gather slot maps on
 * the device from inpoel(3,nelem) (host, 1-based). */
int cfd_mesh_create(cfd_mesh **out, int nelem, int npoin, const int *inpoel);
int cfd_mesh_destroy(cfd_mesh *m);
/* sizes of psup1 and of lap_idx/lap_sparse */
int cfd_mesh_sizes(const cfd_mesh *m, int *npsup, int *nnz);

/* PointNeighbor::getEsup (pointNeighbor.f90:5-43): esup1(3*nelem), esup2(npoin+1) */
int cfd_get_esup(cfd_mesh *m, int *esup1, int *esup2);
/* PointNeighbor::getPsup (pointNeighbor.f90:45-91): psup1(npsup), psup2(npoin+1) */
int cfd_get_psup(cfd_mesh *m, int *psup1, int *psup2);
/* Mlaplace initialize (mLaplace.f90:60-94): lap_rowptr(npoin+1), lap_idx(nnz) */
int cfd_lap_structure(cfd_mesh *m, int *lap_rowptr, int *lap_idx);

/* Mlaplace::laplace (mLaplace.f90:7-58): values of the quality-weighted mesh
 * Laplacian; X, Y are the module coordinates the reference reads by USE. */
int cfd_laplace(cfd_mesh *m, const double *X, const double *Y, const double *dNx,
                const double *dNy, double *lap_sparse, double *lap_diag);

/* deriv (subrutinas.f90:88-126) */
int cfd_deriv(cfd_mesh *m, const double *X, const double *Y, double *area, double *dNx,
              double *dNy, double *HH, double *HHX, double *HHY, double *hmin);

/* MASAS (subrutinas.f90:128-153): lumped mass M(npoin) (overwritten) */
int cfd_masas(cfd_mesh *m, const double *area, double *M);

/* deltat (subrutinas.f90:155-218): DT(nelem) clipped at 10*dtmin */
int cfd_deltat(cfd_mesh *m, const cfd_consts *c, const double *area, const double *T,
               const double *VEL_X, const double *VEL_Y, const double *W_X,
               const double *W_Y, double *DT, double *dtmin);

/* ESTAB (subrutinas.f90:332-446): SHOC, T_SUGN1..3 (nelem) */
int cfd_estab(cfd_mesh *m, const cfd_consts *c, const double *U, const double *T,
              const double *GAMM, const double *VEL_X, const double *VEL_Y,
              const double *W_X, const double *W_Y, const double *dNx, const double *dNy,
              double dtmin, double *shoc, double *t_sugn1, double *t_sugn2,
              double *t_sugn3);

/* calcRHS_mod::calcRHS (calcRHS.f90:4-154): ACCUMULATES into rhs(4,npoin).
 * T(npoin) is the nodal temperature the reference reads from Mvariables. */
int cfd_calcrhs(cfd_mesh *m, const cfd_consts *c, double *rhs, const double *U,
                const double *theta, const double *T, const double *dNx, const double *dNy,
                const double *area, const double *shoc, const double *dtl,
                const double *t_sugn1, const double *t_sugn2, const double *t_sugn3);

/* FUENTE (subrutinas.f90:1036-1094): ALE source, ACCUMULATES into rhs */
int cfd_fuente(cfd_mesh *m, double *rhs, const double *U, const double *W_X,
               const double *W_Y, const double *dNx, const double *dNy, const double *area,
               const double *dtl);

/* Boundary lists as left by dataLoader (dataLoader.f90:130-268). */
typedef struct {
    int nfixv;   const int *ifixv_node;   const double *rfixv_valuex, *rfixv_valuey;
    int nfixrho; const int *ifixrho_node; const double *rfixrho_value;
    int nfixt;   const int *ifixt_node;   const double *rfixt_value;
    int nwall;   const int *wall;         /* wall(2,nwall) edges for NORMALES */
} cfd_bc_lists;

/* Mnormales::normales (subrutinas.f90:7-64): wall normals from X, Y; kept in
 * the mesh context for cfd_rk (the reference keeps them module-private). */
int cfd_normales(cfd_mesh *m, const cfd_bc_lists *bc, const double *X, const double *Y);
/* Mnormales::normalvel (subrutinas.f90:66-85): on the wall nodes the last
 * cfd_normales call found, VEL_X/VEL_Y (npoin, in place) keep only their
 * tangential part relative to the mesh velocity W_X/W_Y, plus W.  Needs a
 * prior cfd_normales on this mesh (CFD_ERR_STATE otherwise). */
int cfd_normalvel(cfd_mesh *m, double *VEL_X, double *VEL_Y, const double *W_X,
                  const double *W_Y);

/* RK (subrutinas.f90:645-849, NGAS=0): the reference passes these through
 * modules; here every array is explicit (host, node-major). */
typedef struct {
    double *U1, *RHS, *T, *VEL_X, *VEL_Y, *RHO, *E, *P, *RMACH;
    const double *U, *W_X, *W_Y, *M, *GAMM;
} cfd_rk_nodes;
typedef struct {
    const double *dNx, *dNy, *area;
    double *shoc, *t_sugn1, *t_sugn2, *t_sugn3;
} cfd_rk_elems;
int cfd_rk(cfd_mesh *m, const cfd_consts *c, const cfd_bc_lists *bc, int nrk, double dtmin,
           const double *dtl, cfd_rk_nodes *nodes, cfd_rk_elems *elems);

/* FORCE_VISC (ns2DComp.ALE.f90:819-893) for one force set: the traction of
 * the full stress tensor (element-average pressure and the set element's P1
 * velocity gradients, Sutherland viscosity at the edge-average temperature)
 * on each edge (n1, n2, elem) of iset(3,nedge), times its length, summed
 * into *f_vx, *f_vy in a fixed order (64-edge partial sums by a fixed tree,
 * then in order; the reference sums in edge order, so the two agree at
 * round-off).  The per-edge skin friction the reference writes to SKIN.DAT
 * is not produced. */
int cfd_force_visc(cfd_mesh *m, int nedge, const int *iset, const double *X, const double *Y,
                   const double *P, const double *T, const double *VEL_X, const double *VEL_Y,
                   const double *dNx, const double *dNy, double t_inf, double *f_vx,
                   double *f_vy);

/* smoothing_mod::smoothing (smoothing.f90:21-75): start-up mesh smoothing of
 * X, Y (host, in place) with the nodes where fixed[i] != 0 held (the main
 * program fixes the moving and fixed-move lists, ns2DComp.ALE.f90:62-76).
 * Nodes of elements with quality below 0.85 become smoothable; two
 * Laplacian passes, then up to 100 sweeps of quality-gradient moves.  dX, dY
 * (both or neither; they are the reference's optional arguments) call
 * setTol_move after the mesh check, as the reference does; the tol_move mode
 * then stays on for this mesh context (module state in the reference).
 * order 0: the reference's node order (one lane walks the nodes: at most
 * 65536 nodes; bit for bit the oracle's sequential restatement of
 * smoothing.f90, which flang cannot compile); order 1: colour by colour over a
 * greedy node colouring, in parallel (the same per-node updates in another
 * order: a different smoothing of the same kind).  *sweeps (may be null):
 * the optimisation sweeps run, 0 when nothing was smoothable; *hmin (may be
 * null): H_MIN_GLOBAL.  Refused when a node has more than 20 elements (the
 * reference's MAX_ELEM_PER_NODE). */
int cfd_smoothing(cfd_mesh *m, double *X, double *Y, const int *fixed, const double *dX,
                  const double *dY, int order, int *sweeps, double *hmin);
/* smoothing_mod::setTol_move (smoothing.f90:386-401): tol_move = |dX| + |dY|
 * per node; the tol_move mode stays on for this mesh context */
int cfd_smoothing_set_tol_move(cfd_mesh *m, const double *dX, const double *dY);

/* BiconjGrad::biCG (biconjGrad.f90:8-62): Jacobi-PCG on a CSR matrix with
 * Dirichlet nodes; x is the initial guess in, the solution out.  Extensions
 * the reference lacks: *iters (iterations run), *rz (final |r.z|). */
int cfd_bicg(int npoin, const double *spMtx, const int *spIdx, const int *spRowptr,
             const double *diagMtx, double *x, const double *b, const double *x_fix,
             const int *x_fixIdx, int nfix, int *iters, double *rz);
/* cfd_bicg with the stopping rule as arguments (no counterpart in the
 * reference, which hard-codes tol = 1e-10 and 1000 iterations): rel = 0 is
 * the reference's absolute rule, rel = 1 the relative rule of
 * cfd_case_desc.pcg_rel; *converged is set when the tolerance was met. */
int cfd_bicg_ex(int npoin, const double *spMtx, const int *spIdx, const int *spRowptr,
                const double *diagMtx, double *x, const double *b, const double *x_fix,
                const int *x_fixIdx, int nfix, double tol, int maxit, int rel, int *iters,
                double *rz, int *converged);

/* cfd_bicg_ex with no allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): no counterpart in
 * the reference.  Every array is a DEVICE pointer (the caller owns them);
 * the solve's work vectors live in a static device array of the library
 * sized for npoin <= 524,288 rows (larger systems are refused with
 * CFD_ERR_ARG before any work); stream is a hipStream_t (null: the default
 * stream); the call returns when the solve has ended.  The row pointer,
 * column and fixed-node indices are checked on the device first.  Calls on
 * one GPU share its scratch array: overlapping calls there race. */
int cfd_bicg_ex_noalloc(int npoin, const double *spMtx, const int *spIdx, const int *spRowptr,
                        const double *diagMtx, double *x, const double *b, const double *x_fix,
                        const int *x_fixIdx, int nfix, double tol, int maxit, int rel,
                        void *stream, int *iters, double *rz, int *converged);

/* The mesh-displacement MLP of the learned warm start (no counterpart in the
 * reference): 16 -> 64 -> 64 -> 16 on the matrix cores, BF16 operands, FP32
 * accumulation, tanh hidden layers; params in the layout W1t[64][16], b1[64],
 * W2t[64][64], b2[64], W3t[16][64], b3[16] (FP32, weights [out][in]); X is
 * n x 16 BF16 bit patterns; two outputs per node.
 * cfd_mlp_forward: out(2,n).  cfd_mlp_gradient: grad(nparams) of
 * sum_nodes sum_{o<2} (out - T)^2 / n and that loss. */
int cfd_mlp_sizes(int *nparams, int *nin, int *nout);
int cfd_mlp_forward(const float *params, const unsigned short *X, int n, float *out);
int cfd_mlp_gradient(const float *params, const unsigned short *X, const float *T, int n,
                     float *grad, float *loss);

/* gcl_mod (gcl.f90:8-62): putW / putArea store the previous mesh velocity and
 * areas in the mesh context; main updates M in place.  mode 0 reproduces the
 * file as written (y-divergence built from W_x), mode 1 uses W_y. */
int cfd_gcl_put_w(cfd_mesh *m, const double *W_x, const double *W_y);
int cfd_gcl_put_area(cfd_mesh *m, const double *area);
int cfd_gcl_main(cfd_mesh *m, double *M, const double *W_x, const double *W_y,
                 const double *dNx, const double *dNy, const double *area, double dt,
                 int mode);

/* ---- device-resident session (ns2DComp.ALE.f90:138-281) ------------------ */
typedef struct {
    cfd_consts c;
    int nrk;                 /* 4 (ns2DComp.ALE.f90:118)                     */
    int moving;              /* MOVING flag: re-run geometry every step       */
    int itlocal;             /* ITLOCAL (local time stepping blend)           */
    double xref, yref;       /* XREF(1), YREF(1): pitch axis of force set 1   */
    int nsetedges;           /* force set 1 edges                             */
    const int *iset;         /* (3,nsetedges): n1 n2 elem                     */
    int nmove, nfix_move;    /* moving / fixed-motion node counts             */
    const int *ilaux;        /* moving nodes, then fixed-motion nodes         */
    cfd_bc_lists bc;
    const double *X, *Y;     /* initial coordinates (npoin)                   */
    /* options beyond the reference (0 = reference behaviour) */
    int motion;              /* 0: DISN(2) = pi/8 sin(10 t) (meshMove.f90:70)
                                1: alpha(t) = motion_amp * sin(motion_omega t) */
    double motion_amp, motion_omega;
    int pcg_rel;             /* 0: r.r < pcg_tol and |r.z| <= pcg_tol absolute
                                   (reference)
                                1: the same tests against pcg_tol^2 times the
                                   r.r and r.z of the guess that is zero off the
                                   fixed nodes (independent of the warm start) */
    double pcg_tol;          /* 0 means 1e-10                                 */
    int pcg_maxit;           /* 0 means 1000                                  */
    int gcl;                 /* -1: no GCL (reference); 0/1: gcl mode          */
    int precision;           /* 0: FP64, element residual, node sums and
                                stage fused per element patch (the residual
                                regrouped, sums per patch: within 1e-12 of
                                the reference order per step);
                                1: FP32 element arithmetic with FP64 node
                                sums and update, fused the same way;
                                2: FP64 in the reference's operation and
                                accumulation order (element kernel, then the
                                node gather in element order).
                                With 0 and 1, fields SHOC, T_SUGN1..3 read
                                back the values the fused kernel used       */
    int warm_start;          /* mesh solves with moving nodes start from
                                0: the previous step's solution (reference)
                                1: the learned prediction (an extrapolation of
                                   the last two solutions plus the MLP's
                                   correction, trained online on the solver's
                                   own solutions) when its initial residual is
                                   smaller, else the previous solution
                                2: the same with the extrapolation alone (no
                                   network: the ablation of 1)                */
    int patch_elems;         /* fused paths (precision 0, 1): elements per
                                patch, 1..682 within the kernel's 64 KB of
                                LDS (0: 512 FP32, 384 FP64)                   */
    int patch_block;         /* threads per patch, a multiple of 64 up to 512
                                (0: 256)                                     */
    int renumber;            /* fused paths (precision 0, 1):
                                0: the session keeps the caller's numbering
                                1: the session numbers nodes and elements in
                                   patch order inside (each patch's elements
                                   and interior nodes contiguous); fields,
                                   boundary lists and the Laplace values go in
                                   and out in the caller's numbering.  Fixed
                                   meshes step bit for bit as with 0; moving
                                   meshes differ at round-off (the PCG dot
                                   products sum in another row order)        */
} cfd_case_desc;

enum {
    CFD_F_U = 0, CFD_F_U1, CFD_F_RHS, CFD_F_T, CFD_F_VEL_X, CFD_F_VEL_Y, CFD_F_W_X,
    CFD_F_W_Y, CFD_F_X, CFD_F_Y, CFD_F_X1, CFD_F_Y1, CFD_F_RHO, CFD_F_E, CFD_F_P,
    CFD_F_RMACH, CFD_F_M, CFD_F_GAMM, CFD_F_XPOS, CFD_F_YPOS, CFD_F_DTL, CFD_F_DT,
    CFD_F_SHOC, CFD_F_T_SUGN1, CFD_F_T_SUGN2, CFD_F_T_SUGN3, CFD_F_AREA, CFD_F_DNX,
    CFD_F_DNY, CFD_F_HH, CFD_F_HHX, CFD_F_HHY, CFD_F_LAP_SPARSE, CFD_F_LAP_DIAG,
    CFD_NFIELDS
};
enum {
    CFD_S_DTMIN = 0, CFD_S_TIME, CFD_S_DTMIN1, CFD_S_BANDERA, CFD_S_ITER, CFD_S_HMIN,
    CFD_S_FX, CFD_S_FY, CFD_S_RM, CFD_S_DISN2, CFD_S_PCG_ITX, CFD_S_PCG_ITY,
    CFD_S_PCG_RZX, CFD_S_PCG_RZY,
    CFD_S_WS_LEARNED,  /* mesh solves started from the learned guess so far  */
    CFD_S_WS_SOLVES,   /* mesh solves with a learned candidate so far        */
    CFD_S_WS_LOSS,     /* the MLP's last training loss                        */
    CFD_S_PCG_ITSUM,   /* mesh-solve iterations summed over both axes so far  */
    CFD_S_FVX,         /* FORCE_VISC F_VX(1), F_VY(1) (ns2DComp.ALE.f90:     */
    CFD_S_FVY,         /*   819-893) after the step's mesh move; 0 if FMU = 0 */
    CFD_S_PCG_UNCONV,  /* mesh solves so far that stopped at the iteration cap
                          without meeting the tolerance                      */
    CFD_NSCALARS
};
/* kernel timing slots (HIP events on the session stream) */
enum {
    CFD_K_ELEMENT = 0,  /* calcRHS+FUENTE element kernel (fused paths: the
                           fused stage launches, one event pair around each
                           step's back-to-back stages)                       */
    CFD_K_STAGE,        /* gather + stage update                              */
    CFD_K_ESTAB, CFD_K_DELTAT, CFD_K_PCG, CFD_K_GEOM, CFD_K_STEP, CFD_NKTIMERS
};

int cfd_session_create(cfd_session **out, cfd_mesh *m, const cfd_case_desc *d);
int cfd_session_destroy(cfd_session *s);
/* RESTART free-stream branch (ns2DComp.ALE.f90:404-420) and the start-up
 * geometry (NORMALES, DERIV, MASAS, laplace: :88-100) */
int cfd_session_init_freestream(cfd_session *s);
/* nsteps time steps; returns after the work is queued (cfd_session_sync
 * waits).  The step's scalars (DTMIN, TIME, forces, PCG counts, ...) stay on
 * the device: a fixed-mesh step makes no host round trip; with moving nodes
 * the host waits for each mesh solve's convergence test. */
int cfd_session_step(cfd_session *s, int nsteps);
int cfd_session_sync(cfd_session *s);
/* field size: npoin, 4*npoin, nelem, 3*nelem or nnz doubles */
int cfd_session_field_size(cfd_session *s, int field, long long *count);
int cfd_session_get_field(cfd_session *s, int field, double *host);
int cfd_session_set_field(cfd_session *s, int field, const double *host);
/* Restart files (ns2DComp.ALE.f90:394-435 RESTART, :898-917 PRINTREST): a
 * Fortran unformatted sequential file, each record framed by its byte length
 * (4-byte integer) before and after; record 0 holds ITER (int32) and TIME,
 * then one record per node U(1:4,i), T(i), GAMM(i).  Host-only: no device is
 * used.  cfd_read_restart reads the first npoin node records (more are
 * ignored, fewer or misframed ones are an error) and returns ITER and TIME
 * too, which the reference's RESTART reads and discards. */
int cfd_write_restart(const char *path, int iter, double time, int npoin, const double *U,
                      const double *T, const double *GAMM);
int cfd_read_restart(const char *path, int npoin, int *iter, double *time, double *U,
                     double *T, double *GAMM);
/* A session's U, T, GAMM and its ITER, TIME to a restart file (the caller's
 * numbering), and back: U and T from the file, GAMM left at GAMA (the main
 * program replaces the loaded GAMM by GAMA when NGAS /= 1,
 * ns2DComp.ALE.f90:78), and, beyond
 * the reference (which leaves them unset), U1 = U and the primitives RHO,
 * VEL_X, VEL_Y, E of U.  The step control restarts (DTMIN, DTMIN1 = 0,
 * BANDERA = 1); keep_clock = 0 sets TIME = ITER = 0 as the main program does
 * (:107-110), 1 keeps the file's. */
int cfd_session_write_restart(cfd_session *s, const char *path);
int cfd_session_read_restart(cfd_session *s, const char *path, int keep_clock);
/* the CFD_S_* scalars; get waits for queued steps */
int cfd_session_get_scalars(cfd_session *s, double *scalars /* CFD_NSCALARS */);
int cfd_session_set_scalars(cfd_session *s, const double *scalars);
/* per-slot kernel time: enable (all slots) resets the accumulators and
 * creates the session's events (a pool of 256, each recorded once) before
 * any timed step */
int cfd_session_timing(cfd_session *s, int enable);
/* the same for the slots whose bit (1 << CFD_K_*) is set in mask */
int cfd_session_timing_mask(cfd_session *s, unsigned mask);
/* total_ms over the slot's timed intervals since enable; launches = the
 * kernel launches those intervals hold (total_ms / launches = mean launch) */
int cfd_session_kernel_time(cfd_session *s, int slot, double *total_ms, long long *launches);

/* ---- batches of independent fixed-mesh cases ----------------------------
 * No counterpart in the reference, which runs one case per process: a sweep
 * of cases of uneven mesh size advanced together on one GPU.  Each case
 * keeps its own step control (DTMIN, hysteresis, clock) and free stream;
 * every case's results agree bit for bit with the same case run alone in a
 * session of the same precision and patch shape.  One step of every case is
 * one replay of a recorded hipGraph (cfd_batch_set_graph(b, 0): plain
 * launches); per RK stage one resident-grid launch covers every case. */
typedef struct cfd_batch cfd_batch;
typedef struct {
    int nelem, npoin;
    const int *inpoel;       /* inpoel(3,nelem), 1-based ids of this case     */
    const double *X, *Y;     /* node coordinates                              */
    cfd_consts c;            /* this case's constants (a sweep over Mach and
                                Reynolds numbers changes FR, FCV, FK and the
                                free stream from case to case)                */
    cfd_bc_lists bc;         /* this case's boundary lists (its own ids)      */
} cfd_batch_case;
/* precision 0 (FP64) or 1 (FP32 element), as cfd_case_desc.precision;
 * patch_elems / patch_block as there (0: defaults); itlocal as the case
 * files' ITLOCAL (0: global time step).  Starts from the free stream. */
int cfd_batch_create(cfd_batch **out, int ncases, const cfd_batch_case *cases, int precision,
                     int itlocal, int patch_elems, int patch_block);
int cfd_batch_destroy(cfd_batch *b);
int cfd_batch_init_freestream(cfd_batch *b);
int cfd_batch_set_graph(cfd_batch *b, int use_graph);
int cfd_batch_step(cfd_batch *b, int nsteps);
int cfd_batch_sync(cfd_batch *b);
/* one case's part of a field: CFD_F_U, U1, RHS, T, VEL_X, VEL_Y, W_X, W_Y,
 * X, Y, RHO, E, P, RMACH, M, GAMM (nodes), DTL, DT, SHOC, T_SUGN1..3, AREA,
 * DNX, DNY, HH, HHX, HHY (elements) */
int cfd_batch_case_size(cfd_batch *b, int icase, int field, long long *count);
int cfd_batch_get_field(cfd_batch *b, int icase, int field, double *host);
/* one case's CFD_S_* scalars (DTMIN, TIME, DTMIN1, BANDERA, ITER) */
int cfd_batch_get_scalars(cfd_batch *b, int icase, double *scalars /* CFD_NSCALARS */);
int cfd_batch_info(cfd_batch *b, long long *npoin, long long *nelem, int *npatch,
                   int *patch_elems, int *patch_block);
/* Autotuner: times nsteps graph-replayed steps for each candidate patch
 * shape (patch_elems[i] elements on patch_block[i] threads), writes the
 * ms per step of each (-1 for a shape whose patches exceed the fused
 * kernel's 64 KB of LDS), keeps the fastest (*best) and restarts every case
 * from the free stream. */
int cfd_batch_autotune(cfd_batch *b, int ncand, const int *patch_elems, const int *patch_block,
                       int nsteps, double *ms_per_step, int *best);

#ifdef __cplusplus
}
#endif
#endif /* CFD_ABI_H */
