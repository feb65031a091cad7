// Session state helpers shared by the single-case session and the batch:
// the boundary-condition table, the free-stream start, step control.
#include <map>

#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {

namespace {
// free stream (ns2DComp.ALE.f90:409-420) plus the start values of the work arrays
__global__ void k_freestream(double* __restrict__ U, double* __restrict__ U1, double* __restrict__ VX,
                             double* __restrict__ VY, double* __restrict__ T, double* __restrict__ GAMM,
                             double* __restrict__ RHO, double* __restrict__ E, double* __restrict__ WX,
                             double* __restrict__ WY, double rho, double u, double v, double t,
                             double gama, double pamb, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GAMM[i] = gama;
    double energia = pamb / ((gama - 1.0) * rho) + 0.5 * (u * u + v * v);
    double u4 = energia * rho;
    double4 q = make_double4(rho, rho * u, rho * v, u4);
    reinterpret_cast<double4*>(U)[i] = q;
    reinterpret_cast<double4*>(U1)[i] = q;
    VX[i] = u;
    VY[i] = v;
    T[i] = t;
    RHO[i] = rho;
    E[i] = u4 / rho;
    WX[i] = -0.0;
    WY[i] = 0.0;
}
// Step control (ns2DComp.ALE.f90:142-166, meshMove.f90:55-70): DTMIN from
// the deltat partials, the 1 % hysteresis against DTMIN1, the clock, the
// local-time-step blend factor and the body's rotation increment for TRANSF.
// BANDERA's increment after fluidStructure (ns2DComp.ALE.f90:250) is applied
// here too: nothing in between reads it.  Block c serves case c: its
// partials are [pb[c], pb[c+1]) (all nb when pb is null), its scalars the
// X_NCTL slots from ctl_all + c * X_NCTL.
__global__ void __launch_bounds__(256)
k_step_ctl(const double* __restrict__ partials, const int* __restrict__ pb, int nb,
           double* __restrict__ ctl_all, int itlocal, int motion, double amp, double omega) {
    const int cs = blockIdx.x;
    const int b0 = pb ? pb[cs] : 0, b1 = pb ? pb[cs + 1] : nb;
    const double m = combine_partials(partials + b0, b1 - b0, 1.0e20, OpMin());
    if (threadIdx.x != 0) return;
    double* ctl = ctl_all + (size_t)cs * X_NCTL;
    double dtmin = m;
    ctl[X_RAW] = dtmin;
    const double iter = ctl[CFD_S_ITER] + 1.0;
    ctl[CFD_S_ITER] = iter;
    double bandera = ctl[CFD_S_BANDERA];
    double dtmin1 = ctl[CFD_S_DTMIN1];
    if ((int)bandera == 1) {
        dtmin1 = dtmin;
        bandera = 2.0;
    }
    const double porc = fabs((dtmin - dtmin1) / dtmin);
    if (100.0 * porc <= 1.0) {
        dtmin = dtmin1;
    } else {
        dtmin1 = dtmin;
        bandera = 2.0;
    }
    ctl[CFD_S_DTMIN1] = dtmin1;
    ctl[CFD_S_BANDERA] = bandera + 1.0;
    ctl[CFD_S_DTMIN] = dtmin;
    ctl[X_FDTL] = itlocal != 0 ? 1.0 - exp(-iter * 4.6 / itlocal) : 1.0;
    const double time = ctl[CFD_S_TIME] + dtmin;
    ctl[CFD_S_TIME] = time;
    // meshMove.f90:55-57,70: ampli*sin(10 t), ampli = pi/8 (motion 0)
    const double pi = 3.141592653589793;
    const double ang = motion == 1 ? amp * sin(omega * time) : pi / 8.0 * sin(10.0 * time);
    const double alpha = ang - ctl[CFD_S_DISN2];
    ctl[CFD_S_DISN2] = ang;
    ctl[X_CA] = cos(alpha);
    ctl[X_SA] = sin(alpha);
}

// DT clipped at 10*min(DT) (subrutinas.f90:210-216), then DTL = DTMIN, or
// DTMIN*f + DT*(1-f) with local time stepping (ns2DComp.ALE.f90:158-164);
// element e reads the scalars of case ecase[e] (case 0 when ecase is null)
__global__ void k_dt_finish(double* __restrict__ DT, double* __restrict__ DTL,
                            const double* __restrict__ ctl_all, const int* __restrict__ ecase,
                            int blend, int n) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* ctl = ctl_all + (ecase ? (size_t)ecase[e] * X_NCTL : 0);
    const double cota = 10.0 * ctl[X_RAW], dtmin = ctl[CFD_S_DTMIN];
    double d = DT[e];
    if (d > cota) {
        d = cota;
        DT[e] = d;
    }
    if (blend) {
        const double f = ctl[X_FDTL];
        DTL[e] = dtmin * f + d * (1.0 - f);
    } else {
        DTL[e] = dtmin;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Boundary-condition table (host construction, reference list semantics)
// ---------------------------------------------------------------------------


BcHost build_bc_host(int npoin, const cfd_bc_lists* bc) {
    BcHost h;
    h.bcidx.assign(npoin, -1);
    auto entry = [&](int node) -> int {
        CFD_CHECK(node >= 1 && node <= npoin, "boundary list node id out of range");
        int& b = h.bcidx[node - 1];
        if (b < 0) {
            b = (int)h.nodes.size();
            h.nodes.push_back(node);
            h.flags.push_back(0);
            h.vals.insert(h.vals.end(), 8, 0.0);
        }
        return b;
    };
    if (!bc) return h;
    for (int i = 0; i < bc->nfixv; ++i) {           // FIXVEL: last occurrence wins
        int b = entry(bc->ifixv_node[i]);
        h.flags[b] |= 1;
        h.vals[8 * b + 0] = bc->rfixv_valuex[i];
        h.vals[8 * b + 1] = bc->rfixv_valuey[i];
    }
    for (int i = 0; i < bc->nfixrho; ++i) {
        int b = entry(bc->ifixrho_node[i]);
        h.flags[b] |= 4;
        h.vals[8 * b + 4] = bc->rfixrho_value[i];
    }
    for (int i = 0; i < bc->nfixt; ++i) {
        int b = entry(bc->ifixt_node[i]);
        h.flags[b] |= 8;
        h.vals[8 * b + 5] = bc->rfixt_value[i];
    }
    // wall nodes: incident edges in ascending edge order (NORMALES sums)
    std::map<int, std::vector<int>> inc;
    for (int e = 0; e < bc->nwall; ++e)
        for (int j = 0; j < 2; ++j) inc[bc->wall[2 * e + j]].push_back(e);
    h.wnode_ptr.push_back(0);
    for (auto& kv : inc) {
        int b = entry(kv.first);
        h.wnode_bc.push_back(b);
        for (int e : kv.second) h.wnode_edges.push_back(e);
        h.wnode_ptr.push_back((int)h.wnode_edges.size());
    }
    return h;
}

Consts consts_of(const cfd_consts* c) {
    Consts k;
    k.gama = c->gama; k.fr = c->fr; k.fmu = c->fmu; k.fk = c->fk; k.fcv = c->fcv;
    k.t_inf = c->t_inf; k.rho_inf = c->rho_inf; k.u_inf = c->u_inf; k.v_inf = c->v_inf;
    k.cte = c->cte; k.fsafe = c->fsafe;
    return k;
}

void launch_freestream(const NodeState& ns, const Consts& c, int n, hipStream_t s) {
    if (n <= 0) return;
    const double pamb = c.rho_inf * c.fr * c.t_inf;
    k_freestream<<<grid_for(n), kBlock, 0, s>>>(ns.U, ns.U1, ns.VX, ns.VY, ns.T, ns.GAMM, ns.RHO, ns.E,
                                               ns.WX, ns.WY, c.rho_inf, c.u_inf, c.v_inf, c.t_inf,
                                               c.gama, pamb, n);
    CFD_HIP(hipGetLastError());
}

void launch_step_ctl(const double* partials, const int* pb, int nb, double* ctl, int ncases,
                     int itlocal, int motion, double amp, double omega, hipStream_t s) {
    k_step_ctl<<<ncases, 256, 0, s>>>(partials, pb, nb, ctl, itlocal, motion, amp, omega);
    CFD_HIP(hipGetLastError());
}

void launch_dt_finish(double* DT, double* DTL, const double* ctl, const int* ecase, bool blend,
                      int nelem, hipStream_t s) {
    k_dt_finish<<<grid_for(nelem), kBlock, 0, s>>>(DT, DTL, ctl, ecase, blend ? 1 : 0, nelem);
    CFD_HIP(hipGetLastError());
}

}  // namespace cfd
