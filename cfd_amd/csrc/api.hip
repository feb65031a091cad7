// C ABI of libcfd_amd.so: mesh context, host-pointer entries for the
// reference routines, and the device-resident session (time step of
// ns2DComp.ALE.f90:138-281).  See include/cfd_abi.h for the contract.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

static Consts to_consts(const cfd_consts* c) { return consts_of(c); }

// ---------------------------------------------------------------------------
// small kernels used by the entries and the session
// ---------------------------------------------------------------------------
namespace {
// POS_AUX(1:NMOVE) = DPOS(ilaux), POS_AUX(NMOVE+1:NNMOVE) = 0 (meshMove.f90:79-89)
__global__ void k_pos_aux(const int* __restrict__ ilaux, const double* __restrict__ dpos,
                          double* __restrict__ pos, int nmove, int nn) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    pos[q] = q < nmove ? dpos[ilaux[q] - 1] : 0.0;
}
// X += XPOS; X1 += XPOS; W = XPOS/DTMIN (meshMove.f90:99-104)
__global__ void k_move(double* __restrict__ X, double* __restrict__ X1, double* __restrict__ W,
                       const double* __restrict__ pos, const double* __restrict__ dtminp, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dtmin = *dtminp;
    double d = pos[i];
    X[i] = X[i] + d;
    X1[i] = X1[i] + d;
    W[i] = d / dtmin;
}

}  // namespace

}  // namespace cfd

using namespace cfd;

// ---------------------------------------------------------------------------
// mesh context
// ---------------------------------------------------------------------------
struct cfd_mesh {
    hipStream_t s = nullptr;
    Tables t;
    std::map<std::string, DBuf<double>> dscratch;
    DBuf<double> partials;
    DBuf<double> dscal;
    // NORMALES / RK state
    std::unique_ptr<BcDevice> bc;
    // gcl_mod saved state
    DBuf<double> w_x_old, w_y_old, area_old;
    bool has_w = false, has_area = false;
    PcgWork pcg;
    // smoothing_mod saved state: tol_move and SWITCH_TOL_MOVE stay set once
    // setTol_move has run (smoothing.f90:15-17)
    SmoothPlan smooth;
    DBuf<double> tol_move;
    bool switch_tol_move = false;

    double* scratch(const char* name, size_t n) {
        auto& b = dscratch[name];
        if (b.n < n) b.alloc(n);
        return b.p;
    }
    template <class T>
    T* put(const char* name, const T* h, size_t n) {
        static_assert(sizeof(T) <= sizeof(double), "scratch element too large");
        double* d = scratch(name, (n * sizeof(T) + 7) / 8 + 1);
        if (n) CFD_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
        return reinterpret_cast<T*>(d);
    }
    template <class T>
    void get(T* h, const T* d, size_t n) {
        if (n) CFD_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    }
    void sync() { CFD_HIP(hipStreamSynchronize(s)); }
};



extern "C" {

int cfd_abi_version(void) { return CFD_ABI_VERSION; }
const char* cfd_last_error(void) { return g_last_error.c_str(); }

int cfd_device_count(int* count) {
    CFD_TRY({
        int n = 0;
        CFD_HIP(hipGetDeviceCount(&n));
        *count = n;
    })
}

int cfd_mesh_create(cfd_mesh** out, int nelem, int npoin, const int* inpoel) {
    CFD_TRY({
        CFD_ARG(out && inpoel && nelem > 0 && npoin > 0, "mesh sizes / pointers");
        int ndev = 0;
        CFD_HIP(hipGetDeviceCount(&ndev));
        CFD_CHECK(ndev > 0, "no GPU device");
        auto m = std::make_unique<cfd_mesh>();
        CFD_HIP(hipStreamCreate(&m->s));
        build_tables(m->t, inpoel, nelem, npoin, m->s);
        m->partials.alloc(grid_for(std::max(nelem, npoin)));
        m->dscal.alloc(8);
        m->sync();
        *out = m.release();
    })
}

int cfd_mesh_destroy(cfd_mesh* m) {
    CFD_TRY({
        if (!m) return CFD_OK;
        hipStream_t s = m->s;
        (void)hipStreamSynchronize(s);
        delete m;
        (void)hipStreamDestroy(s);
    })
}

int cfd_mesh_sizes(const cfd_mesh* m, int* npsup, int* nnz) {
    CFD_TRY({
        CFD_ARG(m, "null mesh");
        if (npsup) *npsup = m->t.npsup;
        if (nnz) *nnz = m->t.nnz;
    })
}

int cfd_get_esup(cfd_mesh* m, int* esup1, int* esup2) {
    CFD_TRY({
        m->t.esup1.download(esup1, 3 * (size_t)m->t.nelem, m->s);
        m->t.esup2.download(esup2, m->t.npoin + 1, m->s);
        m->sync();
    })
}

int cfd_get_psup(cfd_mesh* m, int* psup1, int* psup2) {
    CFD_TRY({
        m->t.psup1.download(psup1, m->t.npsup, m->s);
        m->t.psup2.download(psup2, m->t.npoin + 1, m->s);
        m->sync();
    })
}

int cfd_lap_structure(cfd_mesh* m, int* rowptr, int* idx) {
    CFD_TRY({
        m->t.rowptr.download(rowptr, m->t.npoin + 1, m->s);
        m->t.idx.download(idx, m->t.nnz, m->s);
        m->sync();
    })
}

int cfd_laplace(cfd_mesh* m, const double* X, const double* Y, const double* dNx, const double* dNy,
                double* lap_sparse, double* lap_diag) {
    CFD_TRY({
        const int np = m->t.npoin, ne = m->t.nelem;
        double* dX = m->put("X", X, np);
        double* dY = m->put("Y", Y, np);
        double* dnx = m->put("dNx", dNx, 3 * (size_t)ne);
        double* dny = m->put("dNy", dNy, 3 * (size_t)ne);
        double* lap = m->scratch("lap", m->t.nnz);
        double* diag = m->scratch("diag", np);
        launch_laplace(m->t, dX, dY, dnx, dny, lap, diag, m->s);
        m->get(lap_sparse, lap, m->t.nnz);
        m->get(lap_diag, diag, np);
        m->sync();
    })
}

int cfd_deriv(cfd_mesh* m, const double* X, const double* Y, double* area, double* dNx, double* dNy,
              double* HH, double* HHX, double* HHY, double* hmin) {
    CFD_TRY({
        const int np = m->t.npoin, ne = m->t.nelem;
        double* dX = m->put("X", X, np);
        double* dY = m->put("Y", Y, np);
        double* a = m->scratch("area", ne);
        double* nx = m->scratch("dNx", 3 * (size_t)ne);
        double* ny = m->scratch("dNy", 3 * (size_t)ne);
        double* h = m->scratch("HH", ne);
        double* hx = m->scratch("HHX", ne);
        double* hy = m->scratch("HHY", ne);
        launch_deriv(m->t, dX, dY, a, nx, ny, h, hx, hy, m->dscal.p, m->partials.p, m->s);
        m->get(area, a, ne);
        m->get(dNx, nx, 3 * (size_t)ne);
        m->get(dNy, ny, 3 * (size_t)ne);
        if (HH) m->get(HH, h, ne);
        if (HHX) m->get(HHX, hx, ne);
        if (HHY) m->get(HHY, hy, ne);
        if (hmin) m->get(hmin, (const double*)m->dscal.p, 1);
        m->sync();
    })
}

int cfd_masas(cfd_mesh* m, const double* area, double* M) {
    CFD_TRY({
        double* a = m->put("area", area, m->t.nelem);
        double* dM = m->scratch("M", m->t.npoin);
        launch_masas(m->t, a, dM, m->s);
        m->get(M, dM, m->t.npoin);
        m->sync();
    })
}

int cfd_deltat(cfd_mesh* m, const cfd_consts* c, const double* area, const double* T,
               const double* VEL_X, const double* VEL_Y, const double* W_X, const double* W_Y,
               double* DT, double* dtmin) {
    CFD_TRY({
        CFD_ARG(c, "null consts");
        const int np = m->t.npoin, ne = m->t.nelem;
        Consts k = to_consts(c);
        double* a = m->put("area", area, ne);
        double* t = m->put("T", T, np);
        double* vx = m->put("VX", VEL_X, np);
        double* vy = m->put("VY", VEL_Y, np);
        double* wx = m->put("WX", W_X, np);
        double* wy = m->put("WY", W_Y, np);
        double* dt = m->scratch("DT", ne);
        launch_deltat(m->t, k, a, t, vx, vy, wx, wy, dt, m->dscal.p, m->partials.p, m->s);
        m->get(DT, dt, ne);
        m->get(dtmin, (const double*)m->dscal.p, 1);
        m->sync();
    })
}

int cfd_estab(cfd_mesh* m, const cfd_consts* c, const double* U, const double* T, const double* GAMM,
              const double* VEL_X, const double* VEL_Y, const double* W_X, const double* W_Y,
              const double* dNx, const double* dNy, double dtmin, double* shoc, double* t1,
              double* t2, double* t3) {
    CFD_TRY({
        CFD_ARG(c, "null consts");
        const int np = m->t.npoin, ne = m->t.nelem;
        Consts k = to_consts(c);
        double* u = m->put("U", U, 4 * (size_t)np);
        double* t = m->put("T", T, np);
        double* g = m->put("GAMM", GAMM, np);
        double* vx = m->put("VX", VEL_X, np);
        double* vy = m->put("VY", VEL_Y, np);
        double* wx = m->put("WX", W_X, np);
        double* wy = m->put("WY", W_Y, np);
        double* nx = m->put("dNx", dNx, 3 * (size_t)ne);
        double* ny = m->put("dNy", dNy, 3 * (size_t)ne);
        double* o0 = m->scratch("shoc", ne);
        double* o1 = m->scratch("t1", ne);
        double* o2 = m->scratch("t2", ne);
        double* o3 = m->scratch("t3", ne);
        const double* dtm = m->put("dtmin", &dtmin, 1);
        launch_estab(m->t, k, u, t, g, vx, vy, wx, wy, nx, ny, dtm, o0, o1, o2, o3, m->s);
        m->get(shoc, o0, ne);
        m->get(t1, o1, ne);
        m->get(t2, o2, ne);
        m->get(t3, o3, ne);
        m->sync();
    })
}

static void element_inputs(cfd_mesh* m, ElemIn& in, const double* U, const double* dNx,
                           const double* dNy, const double* area, const double* dtl) {
    const int np = m->t.npoin, ne = m->t.nelem;
    in.U = m->put("U", U, 4 * (size_t)np);
    in.dNx = m->put("dNx", dNx, 3 * (size_t)ne);
    in.dNy = m->put("dNy", dNy, 3 * (size_t)ne);
    in.area = m->put("area", area, ne);
    in.dtl = m->put("dtl", dtl, ne);
}

int cfd_calcrhs(cfd_mesh* m, const cfd_consts* c, double* rhs, const double* U, const double* theta,
                const double* T, const double* dNx, const double* dNy, const double* area,
                const double* shoc, const double* dtl, const double* t1, const double* t2,
                const double* t3) {
    CFD_TRY({
        CFD_ARG(c && rhs && U && T, "null argument");
        const int np = m->t.npoin, ne = m->t.nelem;
        Consts k = to_consts(c);
        ElemIn in{};
        element_inputs(m, in, U, dNx, dNy, area, dtl);
        in.theta = theta ? m->put("theta", theta, 4 * (size_t)np) : nullptr;
        in.T = m->put("T", T, np);
        in.shoc = m->put("shoc", shoc, ne);
        in.t1 = m->put("t1", t1, ne);
        in.t2 = m->put("t2", t2, ne);
        in.t3 = m->put("t3", t3, ne);
        in.WX = in.WY = nullptr;
        double* cc = m->scratch("c_calc", 12 * (size_t)ne);
        double* r = m->put("rhs", rhs, 4 * (size_t)np);
        launch_element_rhs(m->t, k, in, cc, nullptr, m->s);
        launch_gather_rhs(m->t, r, cc, nullptr, r, m->s);
        m->get(rhs, (const double*)r, 4 * (size_t)np);
        m->sync();
    })
}

int cfd_fuente(cfd_mesh* m, double* rhs, const double* U, const double* W_X, const double* W_Y,
               const double* dNx, const double* dNy, const double* area, const double* dtl) {
    CFD_TRY({
        const int np = m->t.npoin, ne = m->t.nelem;
        ElemIn in{};
        element_inputs(m, in, U, dNx, dNy, area, dtl);
        in.WX = m->put("WX", W_X, np);
        in.WY = m->put("WY", W_Y, np);
        double* cf = m->scratch("c_fuente", 12 * (size_t)ne);
        launch_fuente(m->t, in, cf, m->s);
        double* r = m->put("rhs", rhs, 4 * (size_t)np);
        launch_gather_rhs(m->t, r, cf, nullptr, r, m->s);
        m->get(rhs, (const double*)r, 4 * (size_t)np);
        m->sync();
    })
}

int cfd_normales(cfd_mesh* m, const cfd_bc_lists* bc, const double* X, const double* Y) {
    CFD_TRY({
        const int np = m->t.npoin;
        BcHost h = build_bc_host(np, bc);
        m->bc = std::make_unique<BcDevice>();
        m->bc->upload(h, np, bc, m->s);
        double* dX = m->put("X", X, np);
        double* dY = m->put("Y", Y, np);
        m->bc->normales(dX, dY, m->s);
        m->sync();
    })
}

int cfd_normalvel(cfd_mesh* m, double* VEL_X, double* VEL_Y, const double* W_X, const double* W_Y) {
    CFD_TRY({
        CFD_ARG(m && VEL_X && VEL_Y && W_X && W_Y, "null argument");
        CFD_CHECK(m->bc, "state: cfd_normales must run before cfd_normalvel");
        const int np = m->t.npoin;
        double* vx = m->put("VX", VEL_X, np);
        double* vy = m->put("VY", VEL_Y, np);
        double* wx = m->put("WX", W_X, np);
        double* wy = m->put("WY", W_Y, np);
        launch_normalvel(m->bc->tab, vx, vy, wx, wy, m->s);
        m->get(VEL_X, (const double*)vx, np);
        m->get(VEL_Y, (const double*)vy, np);
        m->sync();
    })
}

// smoothing_mod::setTol_move (smoothing.f90:386-401)
static void set_tol_move(cfd_mesh* m, const double* dX, const double* dY) {
    const int np = m->t.npoin;
    std::vector<double> tm(np);
    for (int i = 0; i < np; ++i) tm[i] = std::fabs(dX[i]) + std::fabs(dY[i]);
    if (m->tol_move.n < (size_t)np) m->tol_move.alloc(np);
    m->tol_move.upload(tm.data(), np, m->s);
    m->switch_tol_move = true;
}

int cfd_smoothing_set_tol_move(cfd_mesh* m, const double* dX, const double* dY) {
    CFD_TRY({
        CFD_ARG(m && dX && dY, "null argument");
        set_tol_move(m, dX, dY);
        m->sync();
    })
}

int cfd_smoothing(cfd_mesh* m, double* X, double* Y, const int* fixed, const double* dX,
                  const double* dY, int order, int* sweeps, double* hmin) {
    CFD_TRY({
        CFD_ARG(m && X && Y && fixed, "null argument");
        CFD_ARG(!dX == !dY, "dX and dY go together");
        CFD_ARG(order == 0 || order == 1, "order must be 0 (node order) or 1 (by colour)");
        const int np = m->t.npoin;
        CFD_ARG(order == 1 || np <= kSmoothSeqMaxNodes,
                "order 0 walks the nodes on one lane: at most 65536 nodes (use order 1)");
        m->smooth.build(m->t, fixed, m->s);
        CFD_ARG(m->smooth.max_elems <= 20,
                "a node has more than 20 elements (the reference's MAX_ELEM_PER_NODE)");
        double* dXs = m->put("smX", X, np);
        double* dYs = m->put("smY", Y, np);
        int n = 0;
        if (smooth_check(m->t, m->smooth, dXs, dYs, hmin, m->s)) {
            if (dX) set_tol_move(m, dX, dY);  // after checkMesh (smoothing.f90:44-46)
            n = smooth_run(m->t, m->smooth, dXs, dYs, m->switch_tol_move ? m->tol_move.p : nullptr,
                           order, m->s);
            m->get(X, (const double*)dXs, np);
            m->get(Y, (const double*)dYs, np);
            m->sync();
        }
        if (sweeps) *sweeps = n;
    })
}

int cfd_force_visc(cfd_mesh* m, int nedge, const int* iset, const double* X, const double* Y,
                   const double* P, const double* T, const double* VEL_X, const double* VEL_Y,
                   const double* dNx, const double* dNy, double t_inf, double* f_vx, double* f_vy) {
    CFD_TRY({
        CFD_ARG(m && f_vx && f_vy && nedge >= 0 && (nedge == 0 || iset), "null argument");
        const int np = m->t.npoin, ne = m->t.nelem;
        for (int q = 0; q < nedge; ++q) {
            CFD_ARG(iset[3 * q] >= 1 && iset[3 * q] <= np && iset[3 * q + 1] >= 1 &&
                        iset[3 * q + 1] <= np, "iset node out of range");
            CFD_ARG(iset[3 * q + 2] >= 1 && iset[3 * q + 2] <= ne, "iset element out of range");
        }
        *f_vx = 0.0;
        *f_vy = 0.0;
        if (nedge == 0) return CFD_OK;
        const int* is = m->put("iset", iset, 3 * (size_t)nedge);
        ViscIn v{m->t.inpoel.p, m->put("T", T, np), m->put("VX", VEL_X, np),
                 m->put("VY", VEL_Y, np), m->put("dNx", dNx, 3 * (size_t)ne),
                 m->put("dNy", dNy, 3 * (size_t)ne), t_inf};
        double* dX = m->put("X", X, np);
        double* dY = m->put("Y", Y, np);
        double* dP = m->put("P", P, np);
        launch_forces(nedge, is, dX, dY, dP, 0.0, 0.0, nullptr, &v, m->dscal.p,
                      m->scratch("fpart", forces_part_size(nedge)), nullptr, m->s);
        double out[2];
        m->get(out, (const double*)m->dscal.p, 2);
        m->sync();
        *f_vx = out[0];
        *f_vy = out[1];
    })
}

int cfd_rk(cfd_mesh* m, const cfd_consts* c, const cfd_bc_lists* bc, int nrk, double dtmin,
           const double* dtl, cfd_rk_nodes* nd, cfd_rk_elems* el) {
    CFD_TRY({
        CFD_ARG(c && nd && el && nrk >= 1, "null argument");
        CFD_CHECK(m->bc, "state: cfd_normales must run before cfd_rk");
        (void)bc;
        const int np = m->t.npoin, ne = m->t.nelem;
        const size_t n4 = 4 * (size_t)np;
        Consts k = to_consts(c);
        double* U = m->put("U", nd->U, n4);
        double* T = m->put("T", nd->T, np);
        double* VX = m->put("VX", nd->VEL_X, np);
        double* VY = m->put("VY", nd->VEL_Y, np);
        double* WX = m->put("WX", nd->W_X, np);
        double* WY = m->put("WY", nd->W_Y, np);
        double* M = m->put("M", nd->M, np);
        double* G = m->put("GAMM", nd->GAMM, np);
        double* nx = m->put("dNx", el->dNx, 3 * (size_t)ne);
        double* ny = m->put("dNy", el->dNy, 3 * (size_t)ne);
        double* ar = m->put("area", el->area, ne);
        double* dl = m->put("dtl", dtl, ne);
        double* sh = m->scratch("shoc", ne);
        double* t1 = m->scratch("t1", ne);
        double* t2 = m->scratch("t2", ne);
        double* t3 = m->scratch("t3", ne);
        double* U1 = m->scratch("U1", n4);
        double* RHS = m->scratch("RHS", n4);
        double* RHO = m->scratch("RHO", np);
        double* E = m->scratch("E", np);
        double* P = m->scratch("P", np);
        double* RM = m->scratch("RMACH", np);
        double* cc = m->scratch("c_calc", 12 * (size_t)ne);
        double* cf = m->scratch("c_fuente", 12 * (size_t)ne);
        // ESTAB at stage 1 with the start-of-step U and T (subrutinas.f90:671-677)
        const double* dtm = m->put("dtmin", &dtmin, 1);
        launch_estab(m->t, k, U, T, G, VX, VY, WX, WY, nx, ny, dtm, sh, t1, t2, t3, m->s);
        ElemIn in{U, nullptr, T, nx, ny, ar, sh, dl, t1, t2, t3, WX, WY};
        StageNodes sn{U, U1, RHS, T, VX, VY, RHO, E, P, RM, M, G, WX, WY};
        for (int irk = 1; irk <= nrk; ++irk) {
            double rk_fact = 1.0 / (nrk + 1 - irk);
            launch_element_rhs(m->t, k, in, cc, cf, m->s);
            launch_stage_update(m->t, k, m->bc->tab, rk_fact, cc, cf, sn, irk == nrk, m->s);
        }
        m->get(nd->U1, (const double*)U1, n4);
        m->get(nd->RHS, (const double*)RHS, n4);
        m->get(nd->T, (const double*)T, np);
        m->get(nd->VEL_X, (const double*)VX, np);
        m->get(nd->VEL_Y, (const double*)VY, np);
        m->get(nd->RHO, (const double*)RHO, np);
        m->get(nd->E, (const double*)E, np);
        m->get(nd->P, (const double*)P, np);
        m->get(nd->RMACH, (const double*)RM, np);
        m->get(el->shoc, (const double*)sh, ne);
        m->get(el->t_sugn1, (const double*)t1, ne);
        m->get(el->t_sugn2, (const double*)t2, ne);
        m->get(el->t_sugn3, (const double*)t3, ne);
        m->sync();
    })
}

int cfd_bicg(int npoin, const double* spMtx, const int* spIdx, const int* spRowptr,
             const double* diagMtx, double* x, const double* b, const double* x_fix,
             const int* x_fixIdx, int nfix, int* iters, double* rz) {
    return cfd_bicg_ex(npoin, spMtx, spIdx, spRowptr, diagMtx, x, b, x_fix, x_fixIdx, nfix, 1.0e-10,
                       1000, 0, iters, rz, nullptr);
}

int cfd_bicg_ex(int npoin, const double* spMtx, const int* spIdx, const int* spRowptr,
                const double* diagMtx, double* x, const double* b, const double* x_fix,
                const int* x_fixIdx, int nfix, double tol, int maxit, int rel, int* iters,
                double* rz, int* converged) {
    CFD_TRY({
        CFD_ARG(npoin > 0 && spMtx && spIdx && spRowptr && diagMtx && x && b, "null argument");
        CFD_ARG(tol > 0.0 && maxit >= 0 && (rel == 0 || rel == 1), "tol > 0, maxit >= 0, rel 0/1");
        CFD_ARG(nfix >= 0 && (nfix == 0 || (x_fix && x_fixIdx)), "fixed-node lists");
        for (int q = 0; q < nfix; ++q)
            CFD_ARG(x_fixIdx[q] >= 1 && x_fixIdx[q] <= npoin, "x_fixIdx out of range");
        const int nnz = spRowptr[npoin];
        CFD_ARG(nnz >= 0 && spRowptr[0] == 0, "row pointer");
        for (int i = 0; i < nnz; ++i)
            CFD_ARG(spIdx[i] >= 1 && spIdx[i] <= npoin, "column index out of range");
        hipStream_t s;
        CFD_HIP(hipStreamCreate(&s));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        DBuf<double> A(nnz), D(npoin), X(npoin), B(npoin), XF(std::max(1, nfix));
        DBuf<int> I(nnz), R(npoin + 1), FI(std::max(1, nfix));
        A.upload(spMtx, nnz, s);
        I.upload(spIdx, nnz, s);
        R.upload(spRowptr, npoin + 1, s);
        D.upload(diagMtx, npoin, s);
        X.upload(x, npoin, s);
        B.upload(b, npoin, s);
        XF.upload(x_fix, nfix, s);
        FI.upload(x_fixIdx, nfix, s);
        PcgWork w;
        pcg_launch(w, npoin, A.p, I.p, R.p, D.p, X.p, B.p, XF.p, FI.p, nfix, tol, maxit, rel, s);
        PcgResult res = pcg_result(w, s);
        X.download(x, npoin, s);
        CFD_HIP(hipStreamSynchronize(s));
        if (iters) *iters = res.iters;
        if (rz) *rz = res.rz;
        if (converged) *converged = res.converged;
    })
}

int cfd_mlp_sizes(int* nparams, int* nin, int* nout) {
    CFD_TRY({
        if (nparams) *nparams = kMlpParams;
        if (nin) *nin = kMlpIn;
        if (nout) *nout = kMlpOutUsed;
    })
}

int cfd_mlp_forward(const float* params, const unsigned short* X, int n, float* out) {
    CFD_TRY({
        CFD_ARG(params && X && out && n > 0, "null argument or n <= 0");
        hipStream_t s;
        CFD_HIP(hipStreamCreate(&s));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        MlpWeights w;
        w.init(params, s);
        DBuf<unsigned short> dX(16 * (size_t)n);
        DBuf<float> dO(2 * (size_t)n);
        dX.upload(X, 16 * (size_t)n, s);
        mlp_forward(w, dX.p, n, dO.p, s);
        dO.download(out, 2 * (size_t)n, s);
        CFD_HIP(hipStreamSynchronize(s));
    })
}

int cfd_mlp_gradient(const float* params, const unsigned short* X, const float* T, int n, float* grad,
                     float* loss) {
    CFD_TRY({
        CFD_ARG(params && X && T && grad && n > 0, "null argument or n <= 0");
        hipStream_t s;
        CFD_HIP(hipStreamCreate(&s));
        struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{s};
        MlpWeights w;
        w.init(params, s);
        DBuf<unsigned short> dX(16 * (size_t)n);
        DBuf<float> dT(2 * (size_t)n);
        dX.upload(X, 16 * (size_t)n, s);
        dT.upload(T, 2 * (size_t)n, s);
        mlp_gradient(w, dX.p, dT.p, n, 8, s);
        w.grad.download(grad, kMlpParams, s);
        float sum = 0.f;
        w.loss.download(&sum, 1, s);
        CFD_HIP(hipStreamSynchronize(s));
        if (loss) *loss = sum / (float)n;
    })
}

int cfd_gcl_put_w(cfd_mesh* m, const double* W_x, const double* W_y) {
    CFD_TRY({
        const int np = m->t.npoin;
        if (!m->has_w) {
            m->w_x_old.alloc(np);
            m->w_y_old.alloc(np);
        }
        m->w_x_old.upload(W_x, np, m->s);
        m->w_y_old.upload(W_y, np, m->s);
        m->has_w = true;
        m->sync();
    })
}

int cfd_gcl_put_area(cfd_mesh* m, const double* area) {
    CFD_TRY({
        const int ne = m->t.nelem;
        if (!m->has_area) m->area_old.alloc(ne);
        m->area_old.upload(area, ne, m->s);
        m->has_area = true;
        m->sync();
    })
}

int cfd_gcl_main(cfd_mesh* m, double* M, const double* W_x, const double* W_y, const double* dNx,
                 const double* dNy, const double* area, double dt, int mode) {
    CFD_TRY({
        // gcl.f90:23-25 stops when putW/putArea have not run
        CFD_CHECK(m->has_w && m->has_area, "state: Faltan valores (GCL): putW/putArea first");
        const int np = m->t.npoin, ne = m->t.nelem;
        double* dM = m->put("M", M, np);
        double* wx = m->put("WX", W_x, np);
        double* wy = m->put("WY", W_y, np);
        double* nx = m->put("dNx", dNx, 3 * (size_t)ne);
        double* ny = m->put("dNy", dNy, 3 * (size_t)ne);
        double* a = m->put("area", area, ne);
        const double* dtp = m->put("dt", &dt, 1);
        launch_gcl(m->t, dM, wx, wy, m->w_x_old.p, m->w_y_old.p, nx, ny, a, m->area_old.p, dtp,
                   mode ? 1 : 0, m->s);
        m->get(M, (const double*)dM, np);
        m->sync();
    })
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device-resident session: the main program's time step
// (ns2DComp.ALE.f90:140-281) with every array in HBM.
//
// Per step: deltat -> DTMIN, its hysteresis, DTL and the clock on the device
// (:142-166) -> RK (ESTAB, then nrk x [element residual, gather + stage
// update]) -> fluidStructure (FORCES, TRANSF, two PCG mesh solves, X += XPOS,
// W = XPOS/DTMIN) -> on moving meshes NORMALES, DERIV, MASAS (or the GCL
// update), laplace -> U = U1 (a buffer swap).  All scalars stay in device
// memory (ctl), so a step is queued without waiting for the device.
//
// The reference copies U1 = U before RK (:168-172); RK overwrites all of U1
// in its last stage before anything reads it, so no copy is made here.
// FUENTE adds exactly +-0 while the mesh velocity is identically zero
// (fixed mesh, no moving nodes, W never set by the caller); it is skipped then.
// ---------------------------------------------------------------------------
// Patch-order numbering inside a session (cfd_case_desc.renumber = 1): the
// session runs on its own copy of the mesh renumbered so that each patch's
// elements and interior nodes are contiguous; fields cross the boundary in
// the caller's numbering.
struct Renumbering {
    cfd_mesh* mesh = nullptr;       // the renumbered mesh, owned
    std::vector<int> node_old;      // node_old[i]: the caller's node at position i
    std::vector<int> elem_old;      // elem_old[q]: the caller's element at position q
    std::vector<long long> lap_old; // the caller's CSR entry of each Laplace entry
    ~Renumbering() {
        if (mesh) cfd_mesh_destroy(mesh);
    }
};

struct cfd_session {
    std::unique_ptr<Renumbering> rn;  // first: destroyed after everything using its mesh
    cfd_mesh* m = nullptr;
    hipStream_t s = nullptr;
    Consts c{};
    cfd_case_desc d{};
    int np = 0, ne = 0, nnz = 0;
    std::vector<int> iset_h, ilaux_h;
    DBuf<double> Ub[2];
    int ucur = 0;
    DBuf<double> RHS, T, VX, VY, WX, WY, X, Y, X1, Y1, RHO, E, P, RMACH, M, GAMM, XPOS, YPOS;
    DBuf<double> DTL, DT, SHOC, T1, T2, T3, AREA, DNX, DNY, HH, HHX, HHY, AREA_OLD, WXO, WYO;
    DBuf<double> LAP, LDIAG, c_calc, c_fuente, DXPOS, DYPOS, POSAUX, POSAUY, BZERO, partials;
    // fused paths: shape gradients and ESTAB outputs in patch order, binary32
    // (precision 1) or FP64 (precision 0)
    DBuf<float> DN32, NT32, FAC32;
    DBuf<double> DN6, DN64, NT64, FAC64;
    PatchSet patches;
    DBuf<int> ILAUX, ISET;
    DBuf<double> FPART;  // the forces' block partials
    BcDevice bc;
    PcgWork pcg, pcg_y;  // the x and y mesh solves
    EllLap ell;          // the Laplace values again in ELL order, for the solves' SpMV
    bool ell_on = false;
    WarmStart ws;        // learned warm start (warm_start = 1, moving nodes)
    bool ws_on = false;
    DBuf<double> ctl;                  // session scalars on the device (X_NCTL slots)
    double sc[CFD_NSCALARS] = {0};     // host copy, refreshed by sync_scalars()
    bool w_dirty = false;
    // The mesh solves of a mesh without moving nodes return at their first
    // residual test while XPOS/YPOS are zero (the right-hand side and every
    // fixed value are zero): XPOS/YPOS stay zero and X, Y, W do not change
    // once one X += 0, W = 0/DTMIN pass has normalised signed zeros.
    bool xpos_zero = true, move_normalized = false;
    // Under the reference's absolute PCG rule most warm-started solves stop at
    // their first residual test; after a step whose two solves both did, the
    // next step skips the learned candidate (and its training), as nothing
    // can be gained.  The relative rule always uses it.
    bool last_solves_trivial = false;
    // With ITLOCAL = 0 the fused paths skip DT's clip and the DTL fill (DTL
    // would be DTMIN everywhere, and no kernel of theirs reads DT or DTL);
    // the two fields are formed when a caller reads them (dt_stale).
    bool dt_stale = false;
    bool lazy_dt() const { return fused() && d.itlocal == 0; }
    // timing
    bool timing = false;
    unsigned tmask = 0;
    double tacc[CFD_NKTIMERS] = {0};
    long long tcnt[CFD_NKTIMERS] = {0};
    std::vector<hipEvent_t> pool;
    hipEvent_t open_ev[CFD_NKTIMERS] = {};
    struct Pend { hipEvent_t a, b; int slot, launches; };
    std::vector<Pend> pending;

    double* U() { return Ub[ucur].p; }
    double* U1() { return Ub[ucur ^ 1].p; }
    hipEvent_t ev() {
        if (pool.empty()) {
            hipEvent_t e;
            CFD_HIP(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    void tb(int slot) {
        if (!timing || !((tmask >> slot) & 1u)) return;
        open_ev[slot] = ev();
        CFD_HIP(hipEventRecord(open_ev[slot], s));
    }
    // closes the slot's interval, which held `launches` back-to-back launches
    void te(int slot, int launches = 1) {
        if (!timing || !((tmask >> slot) & 1u)) return;
        hipEvent_t b = ev();
        CFD_HIP(hipEventRecord(b, s));
        pending.push_back({open_ev[slot], b, slot, launches});
    }
    void resolve() {
        for (auto& p : pending) {
            CFD_HIP(hipEventSynchronize(p.b));
            float ms = 0.f;
            CFD_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            tacc[p.slot] += ms;
            tcnt[p.slot] += p.launches;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~cfd_session() {
        if (s) (void)hipStreamSynchronize(s);
        for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    bool fuente_needed() const { return d.moving || d.nmove > 0 || w_dirty; }
    double* scal(int k) { return ctl.p + k; }
    void sync_scalars() {  // device scalars -> sc (waits for the stream)
        ctl.download(sc, CFD_NSCALARS, s);
        CFD_HIP(hipStreamSynchronize(s));
    }
    void put_scalars() {  // sc -> device scalars
        ctl.upload(sc, CFD_NSCALARS, s);
        CFD_HIP(hipStreamSynchronize(s));
    }

    void geometry() {  // NORMALES, DERIV, MASAS, laplace (ns2DComp.ALE.f90:88-100, 259-275)
        const Tables& t = m->t;
        bc.normales(X.p, Y.p, s);
        launch_deriv(t, X.p, Y.p, AREA.p, DNX.p, DNY.p, HH.p, HHX.p, HHY.p, scal(CFD_S_HMIN),
                     partials.p, s);
        refresh_patch_geometry();
        launch_masas(t, AREA.p, M.p, s);
        launch_laplace(t, X.p, Y.p, DNX.p, DNY.p, LAP.p, LDIAG.p, s);
        if (ell_on) ell.refresh(t, LAP.p, s);
    }

    // the fused paths' ESTAB outputs (patch order)
    template <class R>
    ElemP<R> elem_patch(R* nt, R* fac) {
        ElemP<R> o;
        o.nt = nt;
        o.fac = fac;
        o.area = AREA.p;
        o.dtl = lazy_dt() ? nullptr : DTL.p;
        o.inv_perm = patches.inv_perm.p;
        o.cte = c.cte;
        return o;
    }
    Elem32 elem32() { return elem_patch(NT32.p, FAC32.p); }
    Elem64 elem64() { return elem_patch(NT64.p, FAC64.p); }
    bool fused() const { return d.precision == 0 || d.precision == 1; }

    // the shape gradients read by the fused kernels, in patch order
    void refresh_patch_geometry() {
        if (d.precision == 1) {
            launch_to_patch_grad(DNX.p, DNY.p, DN32.p, patches.perm.p, ne, s);
        } else if (d.precision == 0) {
            launch_to_patch64(DNX.p, DNY.p, DN6.p, DN64.p, patches.perm.p, ne, s);
        }
    }

    // The two biCG mesh solves of fluidStructure (meshMove.f90:91-126), then
    // X += XPOS, X1 += XPOS, W_X = XPOS/DTMIN and the same for y.  The y
    // solve reads nothing the x update writes (same matrix, DYPOS from
    // TRANSF), so the two run interleaved.
    void mesh_solves() {
        const int nn = d.nmove + d.nfix_move;
        if (nn > 0) {
            k_pos_aux<<<grid_for(nn), kBlock, 0, s>>>(ILAUX.p, DXPOS.p, POSAUX.p, d.nmove, nn);
            k_pos_aux<<<grid_for(nn), kBlock, 0, s>>>(ILAUX.p, DYPOS.p, POSAUY.p, d.nmove, nn);
        }
        tb(CFD_K_PCG);
        const double tol = d.pcg_tol > 0.0 ? d.pcg_tol : 1.0e-10;
        const int maxit = d.pcg_maxit > 0 ? d.pcg_maxit : 1000;
        WsSolve q;
        const bool ws_now = ws_on && !(d.pcg_rel == 0 && last_solves_trivial);
        if (ws_on) {
            q.XPOS = XPOS.p;
            q.YPOS = YPOS.p;
            q.X = X.p;
            q.Y = Y.p;
            q.xref = d.xref;
            q.yref = d.yref;
            q.dtmin = scal(CFD_S_DTMIN);
            q.wx = &pcg;
            q.wy = &pcg_y;
            q.A = LAP.p;
            q.diag = LDIAG.p;
            q.b = BZERO.p;
            q.fixx = POSAUX.p;
            q.fixy = POSAUY.p;
            q.idx = m->t.idx.p;
            q.rowptr = m->t.rowptr.p;
            q.fix_idx = ILAUX.p;
            q.nfix = nn;
            q.n_learned = scal(CFD_S_WS_LEARNED);
            q.n_solves = scal(CFD_S_WS_SOLVES);
            q.loss = scal(CFD_S_WS_LOSS);
            if (ws_now) ws.before_solves(q, s);  // else the reference's guess
        }
        const PcgSystem sys[2] = {{&pcg, XPOS.p, BZERO.p, POSAUX.p, ILAUX.p, nn},
                                  {&pcg_y, YPOS.p, BZERO.p, POSAUY.p, ILAUX.p, nn}};
        const EllMat em = ell.mat(m->t);
        pcg_launch_multi(sys, 2, np, LAP.p, m->t.idx.p, m->t.rowptr.p, LDIAG.p, tol, maxit,
                         d.pcg_rel, s, ell_on ? &em : nullptr);
        last_solves_trivial = pcg_polled_iterations(pcg) == 0.0 && pcg_polled_iterations(pcg_y) == 0.0;
        te(CFD_K_PCG);
        pcg_store_result(pcg, scal(CFD_S_PCG_ITX), scal(CFD_S_PCG_RZX), scal(CFD_S_PCG_ITSUM), s,
                         scal(CFD_S_PCG_UNCONV));
        pcg_store_result(pcg_y, scal(CFD_S_PCG_ITY), scal(CFD_S_PCG_RZY), scal(CFD_S_PCG_ITSUM), s,
                         scal(CFD_S_PCG_UNCONV));
        if (ws_on) ws.after_solves(q, s);
        k_move<<<grid_for(np), kBlock, 0, s>>>(X.p, X1.p, WX.p, XPOS.p, scal(CFD_S_DTMIN), np);
        k_move<<<grid_for(np), kBlock, 0, s>>>(Y.p, Y1.p, WY.p, YPOS.p, scal(CFD_S_DTMIN), np);
        CFD_HIP(hipGetLastError());
    }

    // One time step, queued on the stream without waiting for the device
    // (except inside the mesh solves: the host polls PCG's done flag).
    void step() {
        const Tables& t = m->t;
        tb(CFD_K_STEP);
        // deltat, DTMIN with its hysteresis, DTL, the clock (ns2DComp.ALE.f90:142-166)
        tb(CFD_K_DELTAT);
        // fused paths: in patch order with the patch's nodes in LDS (the same DT);
        // W not read while it is zero (no moving nodes, W never set)
        const bool fu = fuente_needed();
        const double* wx = fu ? WX.p : nullptr;
        const double* wy = fu ? WY.p : nullptr;
        const int nb = fused() && patches.npatch <= (int)partials.n
                           ? launch_deltat_patch(patches, c, AREA.p, T.p, VX.p, VY.p, wx, wy,
                                                 DT.p, partials.p, s)
                           : launch_deltat_partials(t, c, AREA.p, T.p, VX.p, VY.p, WX.p, WY.p,
                                                    DT.p, partials.p, s);
        launch_step_ctl(partials.p, nullptr, nb, ctl.p, 1, d.itlocal, d.motion, d.motion_amp,
                        d.motion_omega, s);
        if (lazy_dt()) {
            dt_stale = true;
        } else {
            launch_dt_finish(DT.p, DTL.p, ctl.p, nullptr, d.itlocal != 0, ne, s);
        }
        te(CFD_K_DELTAT);
        const double* dtmin = scal(CFD_S_DTMIN);

        // RK (subrutinas.f90:645-849)
        tb(CFD_K_ESTAB);
        if (d.precision == 1)
            launch_estab_patch(t, patches, c, U(), T.p, GAMM.p, VX.p, VY.p, wx, wy, DNX.p,
                               DNY.p, nullptr, dtmin, elem32(), s);
        else if (d.precision == 0)
            launch_estab_patch(t, patches, c, U(), T.p, GAMM.p, VX.p, VY.p, wx, wy, DNX.p,
                               DNY.p, DN6.p, dtmin, elem64(), s);
        else
            launch_estab(t, c, U(), T.p, GAMM.p, VX.p, VY.p, WX.p, WY.p, DNX.p, DNY.p, dtmin,
                         SHOC.p, T1.p, T2.p, T3.p, s);
        te(CFD_K_ESTAB);
        ElemIn in{U(), nullptr, T.p, DNX.p, DNY.p, AREA.p, SHOC.p, DTL.p, T1.p, T2.p, T3.p,
                  fu ? WX.p : nullptr, fu ? WY.p : nullptr};
        PatchIn<float> in32{U(), T.p, fu ? WX.p : nullptr, fu ? WY.p : nullptr, DN32.p,
                            FAC32.p, NT32.p};
        PatchIn<double> in64{U(), T.p, fu ? WX.p : nullptr, fu ? WY.p : nullptr, DN64.p,
                             FAC64.p, NT64.p};
        StageNodes sn{U(), U1(), RHS.p, T.p, VX.p, VY.p, RHO.p, E.p, P.p, RMACH.p, M.p, GAMM.p, WX.p,
                      WY.p};
        const int nrk = d.nrk > 0 ? d.nrk : 4;
        // fused paths: residual + gather + stage in one launch per stage, the
        // stages back to back, timed as the element kernel by one event pair
        // around all of them (an event record costs the stream a few us)
        if (fused()) tb(CFD_K_ELEMENT);
        for (int irk = 1; irk <= nrk; ++irk) {
            double rk_fact = 1.0 / (nrk + 1 - irk);
            if (fused()) {
                if (d.precision == 1)
                    launch_patch_stage(t, patches, c, bc.tab, rk_fact, in32, sn, irk == nrk, s);
                else
                    launch_patch_stage(t, patches, c, bc.tab, rk_fact, in64, sn, irk == nrk, s);
                continue;
            }
            tb(CFD_K_ELEMENT);
            launch_element_rhs(t, c, in, c_calc.p, fu ? c_fuente.p : nullptr, s);
            te(CFD_K_ELEMENT);
            tb(CFD_K_STAGE);
            launch_stage_update(t, c, bc.tab, rk_fact, c_calc.p, fu ? c_fuente.p : nullptr, sn,
                                irk == nrk, s);
            te(CFD_K_STAGE);
        }
        if (fused()) te(CFD_K_ELEMENT, nrk);

        // fluidStructure (meshMove.f90:52-127)
        if (d.gcl >= 0) {
            CFD_HIP(hipMemcpyAsync(WXO.p, WX.p, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
            CFD_HIP(hipMemcpyAsync(WYO.p, WY.p, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
        }
        // FORCE_VISC as the reference's print block calls it (ns2DComp.ALE.f90:
        // 231-236): after the mesh move, with the step's velocities, pressure
        // and temperature and the shape gradients of the start of the step;
        // only when FMU /= 0.  When nothing moves it shares FORCES' launch.
        const ViscIn visc{t.inpoel.p, T.p, VX.p, VY.p, DNX.p, DNY.p, c.t_inf};
        const bool fvisc = c.fmu != 0.0;
        const bool still = d.nmove == 0 && xpos_zero && move_normalized;
        // when nothing moves (see xpos_zero) both solves return at once: their
        // iteration counts and residuals (PCG_ITX, _ITY, _RZX, _RZY) are zeroed
        // by the forces' launch
        static_assert(CFD_S_PCG_ITY == CFD_S_PCG_ITX + 1 && CFD_S_PCG_RZX == CFD_S_PCG_ITX + 2 &&
                          CFD_S_PCG_RZY == CFD_S_PCG_ITX + 3,
                      "mesh-solve scalars are contiguous");
        launch_forces(d.nsetedges, ISET.p, X.p, Y.p, P.p, d.xref, d.yref, scal(CFD_S_FX),
                      fvisc && still ? &visc : nullptr, scal(CFD_S_FVX), FPART.p,
                      still ? scal(CFD_S_PCG_ITX) : nullptr, s);
        if (!still) {
            // TRANSF's displacements reach the solves only through moving nodes
            if (d.nmove > 0)
                launch_transf(d.nsetedges, ISET.p, X.p, Y.p, scal(X_CA), 0.0, d.xref, d.yref,
                              DXPOS.p, DYPOS.p, s);
            mesh_solves();
            move_normalized = true;
            if (fvisc)
                launch_forces(d.nsetedges, ISET.p, X.p, Y.p, P.p, d.xref, d.yref, nullptr, &visc,
                              scal(CFD_S_FVX), FPART.p, nullptr, s);
        }

        if (d.moving == 1) {  // ns2DComp.ALE.f90:259-275
            tb(CFD_K_GEOM);
            if (d.gcl >= 0) {
                CFD_HIP(hipMemcpyAsync(AREA_OLD.p, AREA.p, sizeof(double) * ne,
                                       hipMemcpyDeviceToDevice, s));
                bc.normales(X.p, Y.p, s);
                launch_deriv(t, X.p, Y.p, AREA.p, DNX.p, DNY.p, HH.p, HHX.p, HHY.p,
                             scal(CFD_S_HMIN), partials.p, s);
                refresh_patch_geometry();
                launch_gcl(t, M.p, WX.p, WY.p, WXO.p, WYO.p, DNX.p, DNY.p, AREA.p, AREA_OLD.p, dtmin,
                           d.gcl, s);
                launch_laplace(t, X.p, Y.p, DNX.p, DNY.p, LAP.p, LDIAG.p, s);
                if (ell_on) ell.refresh(t, LAP.p, s);
            } else {
                geometry();
            }
            te(CFD_K_GEOM);
        }
        ucur ^= 1;  // U = U1 (:277-281)
        te(CFD_K_STEP);
    }

    DBuf<double>* field(int f, long long& count) {
        const long long n = np, e = ne;
        switch (f) {
            case CFD_F_U: count = 4 * n; return &Ub[ucur];
            case CFD_F_U1: count = 4 * n; return &Ub[ucur ^ 1];
            case CFD_F_RHS: count = 4 * n; return &RHS;
            case CFD_F_T: count = n; return &T;
            case CFD_F_VEL_X: count = n; return &VX;
            case CFD_F_VEL_Y: count = n; return &VY;
            case CFD_F_W_X: count = n; return &WX;
            case CFD_F_W_Y: count = n; return &WY;
            case CFD_F_X: count = n; return &X;
            case CFD_F_Y: count = n; return &Y;
            case CFD_F_X1: count = n; return &X1;
            case CFD_F_Y1: count = n; return &Y1;
            case CFD_F_RHO: count = n; return &RHO;
            case CFD_F_E: count = n; return &E;
            case CFD_F_P: count = n; return &P;
            case CFD_F_RMACH: count = n; return &RMACH;
            case CFD_F_M: count = n; return &M;
            case CFD_F_GAMM: count = n; return &GAMM;
            case CFD_F_XPOS: count = n; return &XPOS;
            case CFD_F_YPOS: count = n; return &YPOS;
            case CFD_F_DTL: count = e; return &DTL;
            case CFD_F_DT: count = e; return &DT;
            case CFD_F_SHOC: count = e; return &SHOC;
            case CFD_F_T_SUGN1: count = e; return &T1;
            case CFD_F_T_SUGN2: count = e; return &T2;
            case CFD_F_T_SUGN3: count = e; return &T3;
            case CFD_F_AREA: count = e; return &AREA;
            case CFD_F_DNX: count = 3 * e; return &DNX;
            case CFD_F_DNY: count = 3 * e; return &DNY;
            case CFD_F_HH: count = e; return &HH;
            case CFD_F_HHX: count = e; return &HHX;
            case CFD_F_HHY: count = e; return &HHY;
            case CFD_F_LAP_SPARSE: count = nnz; return &LAP;
            case CFD_F_LAP_DIAG: count = n; return &LDIAG;
            default: throw Error("arg: unknown field id");
        }
    }
};

extern "C" {

}  // extern "C"

// A session on the patch-order renumbering of m (cfd_case_desc.renumber = 1):
// the renumbered mesh and description, the session on them, and the maps
// back to the caller's numbering.
static cfd_session* create_renumbered(cfd_mesh* m, const cfd_case_desc* d) {
    const int np = m->t.npoin, ne = m->t.nelem;
    hipStream_t s = m->s;
    std::vector<int> inp(3 * (size_t)ne);
    m->t.inpoel.download(inp.data(), inp.size(), s);
    CFD_HIP(hipStreamSynchronize(s));
    const int pe = d->patch_elems > 0 ? d->patch_elems
                                      : (d->precision == 1 ? kPatchElems : kPatchElems64);
    CFD_ARG(pe >= 1 && pe <= 682, "patch_elems must be 1..682");
    auto rn = std::make_unique<Renumbering>();
    patch_numbering(inp.data(), d->X, d->Y, ne, np, pe, rn->elem_old, rn->node_old);
    CFD_CHECK((int)rn->node_old.size() == np && (int)rn->elem_old.size() == ne,
              "internal: patch numbering");
    std::vector<int> node_new(np), elem_new(ne);
    for (int i = 0; i < np; ++i) node_new[rn->node_old[i]] = i;
    for (int q = 0; q < ne; ++q) elem_new[rn->elem_old[q]] = q;
    std::vector<int> inp2(3 * (size_t)ne);
    for (int q = 0; q < ne; ++q)
        for (int k = 0; k < 3; ++k)  // each element keeps its node order
            inp2[3 * (size_t)q + k] = node_new[inp[3 * (size_t)rn->elem_old[q] + k] - 1] + 1;
    if (cfd_mesh_create(&rn->mesh, ne, np, inp2.data()) != CFD_OK)
        throw Error(std::string("mesh: renumbered mesh: ") + g_last_error);
    // per-node sums (MASAS, laplace, GCL) in the caller's element order, as
    // the caller-numbered session sums them
    reorder_esup(rn->mesh->t, rn->elem_old, rn->mesh->s);
    // the description in the new numbering (node lists mapped, order kept)
    auto nodes = [&](const int* a, size_t n, const char* what) {
        std::vector<int> v(n);
        for (size_t k = 0; k < n; ++k) {
            CFD_ARG(a[k] >= 1 && a[k] <= np, std::string(what) + " node out of range");
            v[k] = node_new[a[k] - 1] + 1;
        }
        return v;
    };
    cfd_case_desc d2 = *d;
    d2.renumber = 0;
    std::vector<double> X2(np), Y2(np);
    for (int i = 0; i < np; ++i) {
        X2[i] = d->X[rn->node_old[i]];
        Y2[i] = d->Y[rn->node_old[i]];
    }
    d2.X = X2.data();
    d2.Y = Y2.data();
    const int nn = d->nmove + d->nfix_move;
    const std::vector<int> ilaux2 = nodes(d->ilaux, nn > 0 ? nn : 0, "ilaux");
    d2.ilaux = nn > 0 ? ilaux2.data() : nullptr;
    std::vector<int> iset2(3 * (size_t)std::max(0, d->nsetedges));
    for (int q = 0; q < d->nsetedges; ++q) {
        for (int j = 0; j < 2; ++j) {
            const int a = d->iset[3 * q + j];
            CFD_ARG(a >= 1 && a <= np, "iset node out of range");
            iset2[3 * q + j] = node_new[a - 1] + 1;
        }
        const int e = d->iset[3 * q + 2];
        iset2[3 * q + 2] = e >= 1 && e <= ne ? elem_new[e - 1] + 1 : e;
    }
    d2.iset = d->nsetedges > 0 ? iset2.data() : nullptr;
    const cfd_bc_lists& b = d->bc;
    const std::vector<int> fv = nodes(b.ifixv_node, std::max(0, b.nfixv), "ifixv_node");
    const std::vector<int> fr = nodes(b.ifixrho_node, std::max(0, b.nfixrho), "ifixrho_node");
    const std::vector<int> ft = nodes(b.ifixt_node, std::max(0, b.nfixt), "ifixt_node");
    const std::vector<int> wl = nodes(b.wall, 2 * (size_t)std::max(0, b.nwall), "wall");
    d2.bc.ifixv_node = b.nfixv > 0 ? fv.data() : b.ifixv_node;
    d2.bc.ifixrho_node = b.nfixrho > 0 ? fr.data() : b.ifixrho_node;
    d2.bc.ifixt_node = b.nfixt > 0 ? ft.data() : b.ifixt_node;
    d2.bc.wall = b.nwall > 0 ? wl.data() : b.wall;
    cfd_session* S = nullptr;
    if (cfd_session_create(&S, rn->mesh, &d2) != CFD_OK)
        throw Error(std::string("arg: renumbered session: ") + g_last_error);
    std::unique_ptr<cfd_session> own(S);
    // each Laplace entry's position in the caller's CSR
    const Tables& t0 = m->t;
    const Tables& t1 = rn->mesh->t;
    std::vector<int> rp0(np + 1), ix0(t0.nnz), rp1(np + 1), ix1(t1.nnz);
    t0.rowptr.download(rp0.data(), np + 1, s);
    t0.idx.download(ix0.data(), t0.nnz, s);
    t1.rowptr.download(rp1.data(), np + 1, rn->mesh->s);
    t1.idx.download(ix1.data(), t1.nnz, rn->mesh->s);
    CFD_HIP(hipStreamSynchronize(s));
    CFD_HIP(hipStreamSynchronize(rn->mesh->s));
    CFD_CHECK(t0.nnz == t1.nnz, "internal: renumbered Laplace structure");
    rn->lap_old.assign(t1.nnz, -1);
    for (int i = 0; i < np; ++i) {
        const int io = rn->node_old[i];
        for (int q = rp1[i]; q < rp1[i + 1]; ++q) {
            const int jo = rn->node_old[ix1[q] - 1] + 1;
            for (int r = rp0[io]; r < rp0[io + 1]; ++r)
                if (ix0[r] == jo) {
                    rn->lap_old[q] = r;
                    break;
                }
            CFD_CHECK(rn->lap_old[q] >= 0, "internal: renumbered Laplace entry");
        }
    }
    own->rn = std::move(rn);
    return own.release();
}

extern "C" {

int cfd_session_create(cfd_session** out, cfd_mesh* m, const cfd_case_desc* d) {
    CFD_TRY({
        CFD_ARG(out && m && d && d->X && d->Y, "null argument");
        CFD_ARG(d->nmove >= 0 && d->nfix_move >= 0 && d->nsetedges >= 0, "list sizes");
        CFD_ARG(d->nmove + d->nfix_move == 0 || d->ilaux, "ilaux");
        CFD_ARG(d->nsetedges == 0 || d->iset, "iset");
        CFD_ARG(d->precision >= 0 && d->precision <= 2,
                "precision must be 0 (FP64), 1 (FP32 element) or 2 (FP64, reference order)");
        CFD_ARG(d->warm_start >= 0 && d->warm_start <= 2, "warm_start must be 0, 1 or 2");
        CFD_ARG(d->renumber == 0 || d->renumber == 1, "renumber must be 0 or 1");
        CFD_ARG(d->renumber == 0 || d->precision == 0 || d->precision == 1,
                "renumber = 1 needs a fused path (precision 0 or 1)");
        if (d->renumber == 1) {
            *out = create_renumbered(m, d);
            return CFD_OK;
        }
        auto S = std::make_unique<cfd_session>();
        S->m = m;
        S->s = m->s;
        S->d = *d;
        S->c = to_consts(&d->c);
        const int np = m->t.npoin, ne = m->t.nelem;
        S->np = np;
        S->ne = ne;
        S->nnz = m->t.nnz;
        const int nn = d->nmove + d->nfix_move;
        for (int q = 0; q < nn; ++q)
            CFD_ARG(d->ilaux[q] >= 1 && d->ilaux[q] <= np, "ilaux node out of range");
        for (int q = 0; q < d->nsetedges; ++q)
            for (int j = 0; j < 2; ++j)
                CFD_ARG(d->iset[3 * q + j] >= 1 && d->iset[3 * q + j] <= np, "iset node out of range");
        S->ilaux_h.assign(d->ilaux, d->ilaux + nn);
        S->iset_h.assign(d->iset, d->iset + 3 * (size_t)d->nsetedges);
        S->d.ilaux = nullptr;
        S->d.iset = nullptr;
        S->d.X = S->d.Y = nullptr;
        S->d.bc = cfd_bc_lists{};
        hipStream_t s = S->s;
        for (auto* b : {&S->Ub[0], &S->Ub[1], &S->RHS}) b->alloc(4 * (size_t)np);
        for (auto* b : {&S->T, &S->VX, &S->VY, &S->WX, &S->WY, &S->X, &S->Y, &S->X1, &S->Y1, &S->RHO,
                        &S->E, &S->P, &S->RMACH, &S->M, &S->GAMM, &S->XPOS, &S->YPOS, &S->WXO,
                        &S->WYO, &S->LDIAG, &S->DXPOS, &S->DYPOS, &S->BZERO})
            b->alloc(np);
        for (auto* b : {&S->DTL, &S->DT, &S->SHOC, &S->T1, &S->T2, &S->T3, &S->AREA, &S->HH, &S->HHX,
                        &S->HHY, &S->AREA_OLD})
            b->alloc(ne);
        S->DNX.alloc(3 * (size_t)ne);
        S->DNY.alloc(3 * (size_t)ne);
        S->LAP.alloc(m->t.nnz);
        if (d->precision == 1) {
            S->DN32.alloc(4 * (size_t)ne);
            S->NT32.alloc(4 * (size_t)ne);
            S->FAC32.alloc(ne);
        } else if (d->precision == 0) {
            S->DN6.alloc(6 * (size_t)ne);
            S->DN64.alloc(4 * (size_t)ne);
            S->NT64.alloc(4 * (size_t)ne);
            S->FAC64.alloc(ne);
        }
        if (S->fused()) {
            std::vector<int> inp(3 * (size_t)ne);
            m->t.inpoel.download(inp.data(), inp.size(), s);
            CFD_HIP(hipStreamSynchronize(s));
            const int pe = d->patch_elems > 0 ? d->patch_elems
                                              : (d->precision == 1 ? kPatchElems : kPatchElems64);
            const int blk = d->patch_block > 0 ? d->patch_block : 256;
            CFD_ARG(pe <= 682, "patch_elems must be at most 682");
            CFD_ARG(blk % 64 == 0 && blk <= 512, "patch_block must be a multiple of 64 up to 512");
            build_patches(S->patches, inp.data(), d->X, d->Y, ne, np, pe, blk, s);
            CFD_ARG(patch_lds_fits(S->patches.max_nodes, pe,
                                   d->precision == 1 ? sizeof(float4) : sizeof(double4)),
                    "patch_elems too large: a patch exceeds the fused kernel's 64 KB of LDS");
        } else {
            S->c_calc.alloc(12 * (size_t)ne);
            S->c_fuente.alloc(12 * (size_t)ne);
        }
        S->POSAUX.alloc(std::max(1, nn));
        S->POSAUY.alloc(std::max(1, nn));
        S->ILAUX.alloc(std::max(1, nn));
        S->ILAUX.upload(S->ilaux_h.data(), nn, s);
        S->ISET.alloc(std::max<size_t>(3, S->iset_h.size()));
        S->ISET.upload(S->iset_h.data(), S->iset_h.size(), s);
        S->FPART.alloc(forces_part_size((int)(S->iset_h.size() / 3)));
        S->ctl.alloc(X_NCTL);
        S->ctl.zero(s);
        S->partials.alloc(grid_for(std::max(ne, np)));
        // zero everything once (DXPOS/DYPOS keep zeros off the set nodes,
        // meshMove.f90:52-53; B stays zero, :91-95)
        for (auto* b : {&S->Ub[0], &S->Ub[1], &S->RHS, &S->T, &S->VX, &S->VY, &S->WX, &S->WY, &S->X1,
                        &S->Y1, &S->RHO, &S->E, &S->P, &S->RMACH, &S->M, &S->GAMM, &S->XPOS, &S->YPOS,
                        &S->WXO, &S->WYO, &S->LDIAG, &S->DXPOS, &S->DYPOS, &S->BZERO, &S->DTL, &S->DT,
                        &S->SHOC, &S->T1, &S->T2, &S->T3, &S->AREA, &S->HH, &S->HHX, &S->HHY,
                        &S->AREA_OLD, &S->DNX, &S->DNY, &S->LAP})
            b->zero(s);
        S->X.upload(d->X, np, s);
        S->Y.upload(d->Y, np, s);
        BcHost h = build_bc_host(np, &d->bc);
        S->bc.upload(h, np, &d->bc, s);
        S->pcg.ensure(np);
        S->pcg_y.ensure(np);
        // the ELL copy serves the mesh solves, which iterate only with moving nodes
        if (d->nmove > 0) S->ell_on = S->ell.build(m->t, s);
        S->ws_on = d->warm_start != 0 && d->nmove > 0;
        if (S->ws_on) {  // length scale: farthest node from the pivot
            double l = 0.0;
            for (int i = 0; i < np; ++i)
                l = std::max(l, std::hypot(d->X[i] - d->xref, d->Y[i] - d->yref));
            S->ws.init(np, l, s);
            S->ws.use_mlp = d->warm_start == 1;
        }
        S->sc[CFD_S_BANDERA] = 1.0;
        S->put_scalars();
        *out = S.release();
    })
}

int cfd_session_destroy(cfd_session* S) {
    CFD_TRY({ delete S; })
}

int cfd_session_init_freestream(cfd_session* S) {
    CFD_TRY({
        CFD_ARG(S, "null session");
        launch_freestream({S->U(), S->U1(), S->VX.p, S->VY.p, S->T.p, S->GAMM.p, S->RHO.p, S->E.p,
                           S->WX.p, S->WY.p},
                          S->c, S->np, S->s);
        S->geometry();
        CFD_HIP(hipMemcpyAsync(S->AREA_OLD.p, S->AREA.p, sizeof(double) * S->ne,
                               hipMemcpyDeviceToDevice, S->s));
        S->sync_scalars();
        for (int k : {CFD_S_TIME, CFD_S_ITER, CFD_S_DTMIN, CFD_S_DTMIN1, CFD_S_DISN2}) S->sc[k] = 0.0;
        S->sc[CFD_S_BANDERA] = 1.0;
        S->put_scalars();
        S->w_dirty = false;
        S->move_normalized = false;
        S->last_solves_trivial = false;
        S->ws.nhist = 0;  // the learned start's history restarts with the run
        S->ws.predicted = S->ws.net_used = false;
    })
}

int cfd_session_step(cfd_session* S, int nsteps) {
    CFD_TRY({
        CFD_ARG(S && nsteps >= 0, "arguments");
        for (int k = 0; k < nsteps; ++k) S->step();
    })
}

int cfd_session_sync(cfd_session* S) {
    CFD_TRY({
        CFD_HIP(hipStreamSynchronize(S->s));
        S->resolve();
    })
}

int cfd_session_field_size(cfd_session* S, int f, long long* count) {
    CFD_TRY({ S->field(f, *count); })
}

}  // extern "C"

// Layout of a field for the caller's numbering: per node or per element,
// values per entity; the Laplace values through their entry map.
enum FieldKind { FK_NODE, FK_ELEM, FK_LAP };
static FieldKind field_kind(int f, int& width) {
    width = 1;
    switch (f) {
        case CFD_F_U: case CFD_F_U1: case CFD_F_RHS: width = 4; return FK_NODE;
        case CFD_F_DTL: case CFD_F_DT: case CFD_F_SHOC: case CFD_F_T_SUGN1: case CFD_F_T_SUGN2:
        case CFD_F_T_SUGN3: case CFD_F_AREA: case CFD_F_HH: case CFD_F_HHX: case CFD_F_HHY:
            return FK_ELEM;
        case CFD_F_DNX: case CFD_F_DNY: width = 3; return FK_ELEM;
        case CFD_F_LAP_SPARSE: return FK_LAP;
        default: return FK_NODE;
    }
}

// in[position] -> out[caller's position] (to_caller) or back
static void renumber_field(const Renumbering& rn, int f, const double* in, double* out,
                           bool to_caller) {
    int w = 1;
    const FieldKind k = field_kind(f, w);
    if (k == FK_LAP) {
        for (size_t q = 0; q < rn.lap_old.size(); ++q) {
            if (to_caller) out[rn.lap_old[q]] = in[q];
            else out[q] = in[rn.lap_old[q]];
        }
        return;
    }
    const std::vector<int>& old = k == FK_NODE ? rn.node_old : rn.elem_old;
    for (size_t i = 0; i < old.size(); ++i)
        for (int c = 0; c < w; ++c) {
            if (to_caller) out[w * (size_t)old[i] + c] = in[w * i + c];
            else out[w * i + c] = in[w * (size_t)old[i] + c];
        }
}

extern "C" {

int cfd_session_get_field(cfd_session* S, int f, double* host_out) {
    CFD_TRY({
        CFD_ARG(S && host_out, "null argument");
        long long n = 0;
        S->field(f, n);
        std::vector<double> tmp(S->rn ? (size_t)n : 0);
        double* host = S->rn ? tmp.data() : host_out;
        DBuf<double>* b = S->field(f, n);
        const bool estab_field = f == CFD_F_SHOC || f == CFD_F_T_SUGN1 || f == CFD_F_T_SUGN2 ||
                                 f == CFD_F_T_SUGN3;
        if (S->d.precision == 1 && estab_field)  // the binary32 values the FP32 path used
            launch_estab_widen(S->elem32(), S->patches.inv_perm.p, S->ne, S->SHOC.p, S->T1.p,
                               S->T2.p, S->T3.p, S->s);
        if (S->d.precision == 0 && estab_field)  // from the fused FP64 path's patch order
            launch_estab_widen(S->elem64(), S->patches.inv_perm.p, S->ne, S->SHOC.p, S->T1.p,
                               S->T2.p, S->T3.p, S->s);
        if ((f == CFD_F_DT || f == CFD_F_DTL) && S->dt_stale) {  // the last step's clip and DTL
            launch_dt_finish(S->DT.p, S->DTL.p, S->ctl.p, nullptr, false, S->ne, S->s);
            S->dt_stale = false;
        }
        b->download(host, n, S->s);
        CFD_HIP(hipStreamSynchronize(S->s));
        if (S->rn) renumber_field(*S->rn, f, tmp.data(), host_out, true);
    })
}

int cfd_session_set_field(cfd_session* S, int f, const double* host_in) {
    CFD_TRY({
        CFD_ARG(S && host_in, "null argument");
        long long n = 0;
        DBuf<double>* b = S->field(f, n);
        std::vector<double> tmp(S->rn ? (size_t)n : 0);
        if (S->rn) renumber_field(*S->rn, f, host_in, tmp.data(), false);
        const double* host = S->rn ? tmp.data() : host_in;
        b->upload(host, n, S->s);
        CFD_HIP(hipStreamSynchronize(S->s));
        if (f == CFD_F_W_X || f == CFD_F_W_Y || f == CFD_F_XPOS || f == CFD_F_YPOS) S->w_dirty = true;
        if (f == CFD_F_XPOS || f == CFD_F_YPOS) S->xpos_zero = false;
        if (f == CFD_F_X || f == CFD_F_Y || f == CFD_F_X1 || f == CFD_F_Y1 || f == CFD_F_W_X ||
            f == CFD_F_W_Y)
            S->move_normalized = false;
        if (f == CFD_F_DNX || f == CFD_F_DNY) {
            S->refresh_patch_geometry();
            CFD_HIP(hipStreamSynchronize(S->s));
        }
        if (f == CFD_F_DT || f == CFD_F_DTL) S->dt_stale = false;  // the caller's values
        if (f == CFD_F_LAP_SPARSE && S->ell_on) {
            S->ell.refresh(S->m->t, S->LAP.p, S->s);
            CFD_HIP(hipStreamSynchronize(S->s));
        }
    })
}

int cfd_session_get_scalars(cfd_session* S, double* out) {
    CFD_TRY({
        S->sync_scalars();
        std::memcpy(out, S->sc, sizeof(S->sc));
    })
}

int cfd_session_set_scalars(cfd_session* S, const double* in) {
    CFD_TRY({
        std::memcpy(S->sc, in, sizeof(S->sc));
        S->put_scalars();
    })
}

int cfd_session_timing_mask(cfd_session* S, unsigned mask) {
    CFD_TRY({
        CFD_HIP(hipStreamSynchronize(S->s));
        S->resolve();
        S->timing = mask != 0;
        S->tmask = mask;
        // the events come from the pool: created here and recorded once,
        // so that no event is created or used for the first time while the
        // timed steps are queued
        constexpr size_t kEventPool = 256;
        bool fresh = false;
        while (S->timing && S->pool.size() < kEventPool) {
            hipEvent_t e;
            CFD_HIP(hipEventCreate(&e));
            CFD_HIP(hipEventRecord(e, S->s));
            S->pool.push_back(e);
            fresh = true;
        }
        if (fresh) CFD_HIP(hipStreamSynchronize(S->s));
        for (int k = 0; k < CFD_NKTIMERS; ++k) {
            S->tacc[k] = 0.0;
            S->tcnt[k] = 0;
        }
    })
}

int cfd_session_timing(cfd_session* S, int enable) {
    return cfd_session_timing_mask(S, enable ? (1u << CFD_NKTIMERS) - 1u : 0u);
}

int cfd_session_kernel_time(cfd_session* S, int slot, double* total_ms, long long* launches) {
    CFD_TRY({
        CFD_ARG(slot >= 0 && slot < CFD_NKTIMERS, "timer slot");
        *total_ms = S->tacc[slot];
        *launches = S->tcnt[slot];
    })
}

}  // extern "C"
