// Learned warm start of the two mesh-Laplace solves of fluidStructure
// (meshMove.f90:91-126), the reference's guess being the previous step's
// solution.
//
// With x1, x2, x3 the last three solutions (displacements over their steps
// dt1, dt2, dt3) and dt this step's, per axis:
//   E   = dt (2 x1/dt1 - x2/dt2), the rates extrapolated linearly (FP64);
//   D2  = dt (x1/dt1 - 2 x2/dt2 + x3/dt3), the rates' second difference;
//   y   = the MLP's output for this node (matrix cores, BF16, mlp.hip), from
//         the features E/se, (x1 - x2)/sd, D2/s2 (both axes), node position
//         relative to the pivot / L, distance / L, 1  (se, sd, s2: the
//         maxima of |E|, |x1 - x2|, |D2| over the nodes and both axes);
//   xhat = E + beta y, beta = (y . r_E) / (y . A y) over the free nodes, r_E
//         the residual of E: the step along y that minimises the energy
//         norm of the error, so xhat's error is never larger than E's in
//         that norm (Galerkin projection of the network's correction).
// The network is trained online, kWsAdamSteps Adam steps per time step on
// strided node samples, to map the features to the shape (x - E)/sc of the
// correction the solver has just computed (sc = max |x - E| over the nodes
// and both axes; beta sets the size).  Each solve starts from xhat when xhat's
// first residual r.r is smaller than that of the previous solution, else
// from the previous solution, so the solve never starts worse than the
// reference's start.  With two solutions of history only (the third step),
// and without the network (use_mlp false, the ablation), xhat = E.
#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {

namespace {

constexpr int kWsTrainNodes = 32768;
constexpr int kWsAdamSteps = 4;

enum {
    W_DT1 = 0, W_DT2, W_DT3, W_SE, W_SD, W_S2, W_SC, W_BX, W_BY, W_RRPX, W_RRLX, W_RRPY,
    W_RRLY, W_RREX, W_RREY, W_N = 16
};

__device__ __forceinline__ unsigned short to_bf16(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// E = dt (2 x1/dt1 - x2/dt2) for both axes
__global__ void k_ws_extrap(const double* __restrict__ h1x, const double* __restrict__ h2x,
                            const double* __restrict__ h1y, const double* __restrict__ h2y,
                            const double* __restrict__ dtp, const double* __restrict__ wsc,
                            double* __restrict__ ex, double* __restrict__ ey, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = *dtp, a = 2.0 * dt / wsc[W_DT1], b = dt / wsc[W_DT2];
    ex[i] = a * h1x[i] - b * h2x[i];
    ey[i] = a * h1y[i] - b * h2y[i];
}

// D2 = dt (x1/dt1 - 2 x2/dt2 + x3/dt3) of one node and axis
__device__ __forceinline__ double ws_d2(const double* h1, const double* h2, const double* h3, int i,
                                        double dt, const double* wsc) {
    return dt * (h1[i] / wsc[W_DT1] - 2.0 * h2[i] / wsc[W_DT2] + h3[i] / wsc[W_DT3]);
}

// block partials of max |D2| over both axes
__global__ void __launch_bounds__(kBlock)
k_ws_maxd2(const double* __restrict__ h1x, const double* __restrict__ h2x,
           const double* __restrict__ h3x, const double* __restrict__ h1y,
           const double* __restrict__ h2y, const double* __restrict__ h3y,
           const double* __restrict__ dtp, const double* __restrict__ wsc, int n,
           double* __restrict__ partials) {
    const double dt = *dtp;
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = fmax(m, fmax(fabs(ws_d2(h1x, h2x, h3x, i, dt, wsc)), fabs(ws_d2(h1y, h2y, h3y, i, dt, wsc))));
    block_partial(m, OpMax(), partials);
}

// block partials of max(|a - b|, |c - d|) (b, d may be null: zero)
__global__ void __launch_bounds__(kBlock)
k_maxabs2(const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ c,
          const double* __restrict__ d, int n, double* __restrict__ partials) {
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double u = fabs(a[i] - (b ? b[i] : 0.0)), v = fabs(c[i] - (d ? d[i] : 0.0));
        m = fmax(m, fmax(u, v));
    }
    block_partial(m, OpMax(), partials);
}

__global__ void __launch_bounds__(256)
k_max_to(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    const double m = combine_partials(partials, nb, 0.0, OpMax());
    if (threadIdx.x == 0) *out = m > 0.0 ? m : 1.0;
}

__global__ void k_ws_features(const double* __restrict__ ex, const double* __restrict__ ey,
                              const double* __restrict__ h1x, const double* __restrict__ h2x,
                              const double* __restrict__ h3x, const double* __restrict__ h1y,
                              const double* __restrict__ h2y, const double* __restrict__ h3y,
                              const double* __restrict__ X, const double* __restrict__ Y, double xref,
                              double yref, double inv_l, const double* __restrict__ dtp,
                              const double* __restrict__ wsc, unsigned short* __restrict__ feat,
                              int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = *dtp;
    const double ise = 1.0 / wsc[W_SE], isd = 1.0 / wsc[W_SD], is2 = 1.0 / wsc[W_S2];
    const double dx = X[i] - xref, dy = Y[i] - yref;
    float f[kMlpIn] = {0.f};
    f[0] = (float)(ex[i] * ise);
    f[1] = (float)(ey[i] * ise);
    f[2] = (float)((h1x[i] - h2x[i]) * isd);
    f[3] = (float)((h1y[i] - h2y[i]) * isd);
    f[4] = (float)(ws_d2(h1x, h2x, h3x, i, dt, wsc) * is2);
    f[5] = (float)(ws_d2(h1y, h2y, h3y, i, dt, wsc) * is2);
    f[6] = (float)(dx * inv_l);
    f[7] = (float)(dy * inv_l);
    f[8] = (float)(sqrt(dx * dx + dy * dy) * inv_l);
    f[9] = 1.0f;
    unsigned short* p = feat + (size_t)i * kMlpIn;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) p[k] = to_bf16(f[k]);
}

// Partials of y.r and y.(A y) over the free rows of one axis (y = the
// network's output for that axis, taken as zero on the fixed nodes; r = the
// residual of E, zero on the fixed rows)
__global__ void __launch_bounds__(kBlock)
k_ws_galerkin(int n, const double* __restrict__ A, const int* __restrict__ idx,
              const int* __restrict__ rowptr, const int* __restrict__ fixflag,
              const float* __restrict__ out, int axis, const double* __restrict__ r,
              double* __restrict__ partials) {
    double yr = 0.0, yay = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (fixflag[i]) continue;
        double ay = 0.0;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int j = idx[q] - 1;
            if (!fixflag[j]) ay = ay + A[q] * (double)out[2 * (size_t)j + axis];
        }
        const double yi = (double)out[2 * (size_t)i + axis];
        yr = yr + yi * r[i];
        yay = yay + yi * ay;
    }
    block_partial(yr, OpSum(), partials);
    block_partial(yay, OpSum(), partials + gridDim.x);
}

// beta = y.r / y.(A y) (0 when y.(A y) is not positive)
__global__ void __launch_bounds__(256)
k_ws_beta(const double* __restrict__ partials, int nb, double* __restrict__ beta) {
    const double yr = combine_partials(partials, nb, 0.0, OpSum());
    const double yay = combine_partials(partials + nb, nb, 0.0, OpSum());
    if (threadIdx.x == 0) *beta = yay > 0.0 ? yr / yay : 0.0;
}

// xhat = E + beta y for both axes (beta zero: xhat = E)
__global__ void k_ws_predict(const double* __restrict__ ex, const double* __restrict__ ey,
                             const float* __restrict__ out, const double* __restrict__ wsc,
                             int use_net, double* __restrict__ hatx, double* __restrict__ haty,
                             int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double bx = use_net ? wsc[W_BX] : 0.0, by = use_net ? wsc[W_BY] : 0.0;
    hatx[i] = use_net ? ex[i] + bx * (double)out[2 * (size_t)i] : ex[i];
    haty[i] = use_net ? ey[i] + by * (double)out[2 * (size_t)i + 1] : ey[i];
}

// x = hat when the learned first residual is smaller (a choice every block
// makes alike); block 0 counts the choices in the session scalars
__global__ void k_ws_select(double* __restrict__ x, const double* __restrict__ hat,
                            const double* __restrict__ wsc, int rr_prev, int rr_learn,
                            double* __restrict__ n_learned, double* __restrict__ n_solves, int n) {
    const bool use = wsc[rr_learn] < wsc[rr_prev];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (use && i < n) x[i] = hat[i];
    if (i == 0) {
        *n_solves = *n_solves + 1.0;
        if (use) *n_learned = *n_learned + 1.0;
    }
}

__global__ void k_ws_target(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ ex, const double* __restrict__ ey,
                            const double* __restrict__ wsc, float* __restrict__ tgt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double isc = 1.0 / wsc[W_SC];
    tgt[2 * (size_t)i] = (float)((x[i] - ex[i]) * isc);
    tgt[2 * (size_t)i + 1] = (float)((y[i] - ey[i]) * isc);
}

// history shift and the rates' time steps
__global__ void k_ws_shift(const double* __restrict__ x, const double* __restrict__ y,
                           double* __restrict__ h1x, double* __restrict__ h2x,
                           double* __restrict__ h3x, double* __restrict__ h1y,
                           double* __restrict__ h2y, double* __restrict__ h3y,
                           const double* __restrict__ dtp, double* __restrict__ wsc, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        h3x[i] = h2x[i];
        h2x[i] = h1x[i];
        h1x[i] = x[i];
        h3y[i] = h2y[i];
        h2y[i] = h1y[i];
        h1y[i] = y[i];
    }
    if (i == 0) {  // read by the next step only
        wsc[W_DT3] = wsc[W_DT2];
        wsc[W_DT2] = wsc[W_DT1];
        wsc[W_DT1] = *dtp;
    }
}

__global__ void k_copy_loss(const float* __restrict__ loss, double inv_n, double* __restrict__ dst) {
    *dst = (double)*loss * inv_n;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

void WarmStart::init(int npoin, double l, hipStream_t s) {
    np = npoin;
    nhist = 0;
    predicted = net_used = false;
    inv_l = l > 0.0 ? 1.0 / l : 1.0;
    for (auto* b : {&h1x, &h2x, &h3x, &h1y, &h2y, &h3y, &ex, &ey, &hatx, &haty}) b->alloc(np);
    for (auto* b : {&h1x, &h2x, &h3x, &h1y, &h2y, &h3y}) b->zero(s);
    feat.alloc(kMlpIn * (size_t)np);
    out.alloc(2 * (size_t)np);
    tgt.alloc(2 * (size_t)np);
    wsc.alloc(W_N);
    wsc.zero(s);
    partials.alloc(2 * kRedBlocks);
    // Glorot-uniform hidden layers, zero output layer: the first prediction is E
    std::vector<float> p(kMlpParams, 0.0f);
    uint64_t seed = 20261202ull;
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int k = 0; k < fan_out * fan_in; ++k) {
            const double u = (double)(splitmix(seed) >> 11) * (1.0 / 9007199254740992.0);
            p[off + k] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(kMlpW1, kMlpHid, kMlpIn);
    fill(kMlpW2, kMlpHid, kMlpHid);
    mlp.init(p.data(), s);
}

void WarmStart::before_solves(const WsSolve& q, hipStream_t s) {
    if (nhist < 2) return;  // no rate history yet: the reference's guess
    const unsigned g = grid_for(np);
    const int nb = std::min<int>(grid_for(np), kRedBlocks);
    net_used = use_mlp && nhist >= 3;
    k_ws_extrap<<<g, kBlock, 0, s>>>(h1x.p, h2x.p, h1y.p, h2y.p, q.dtmin, wsc.p, ex.p, ey.p, np);
    if (net_used) {
        k_maxabs2<<<nb, kBlock, 0, s>>>(ex.p, nullptr, ey.p, nullptr, np, partials.p);
        k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SE);
        k_maxabs2<<<nb, kBlock, 0, s>>>(h1x.p, h2x.p, h1y.p, h2y.p, np, partials.p);
        k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SD);
        k_ws_maxd2<<<nb, kBlock, 0, s>>>(h1x.p, h2x.p, h3x.p, h1y.p, h2y.p, h3y.p, q.dtmin, wsc.p,
                                         np, partials.p);
        k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_S2);
        k_ws_features<<<g, kBlock, 0, s>>>(ex.p, ey.p, h1x.p, h2x.p, h3x.p, h1y.p, h2y.p, h3y.p,
                                           q.X, q.Y, q.xref, q.yref, inv_l, q.dtmin, wsc.p, feat.p,
                                           np);
        mlp_forward(mlp, feat.p, np, out.p, s);
        // the energy-optimal step along the network's correction, per axis,
        // from the residual of E (left in the work space's r by the call)
        struct Axis { PcgWork* w; const double *e, *fix; int slot_rr, slot_b; };
        for (const Axis a : {Axis{q.wx, ex.p, q.fixx, W_RREX, W_BX}, Axis{q.wy, ey.p, q.fixy, W_RREY, W_BY}}) {
            pcg_initial_rr(*a.w, np, q.A, q.idx, q.rowptr, q.diag, a.e, q.b, a.fix, q.fix_idx,
                           q.nfix, wsc.p + a.slot_rr, s);
            k_ws_galerkin<<<nb, kBlock, 0, s>>>(np, q.A, q.idx, q.rowptr, a.w->fixflag.p, out.p,
                                                a.slot_b == W_BX ? 0 : 1, a.w->r.p, partials.p);
            k_ws_beta<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + a.slot_b);
        }
    }
    k_ws_predict<<<g, kBlock, 0, s>>>(ex.p, ey.p, out.p, wsc.p, net_used ? 1 : 0, hatx.p, haty.p, np);
    // first residuals of both candidates, both axes
    pcg_initial_rr(*q.wx, np, q.A, q.idx, q.rowptr, q.diag, q.XPOS, q.b, q.fixx, q.fix_idx, q.nfix,
                   wsc.p + W_RRPX, s);
    pcg_initial_rr(*q.wx, np, q.A, q.idx, q.rowptr, q.diag, hatx.p, q.b, q.fixx, q.fix_idx, q.nfix,
                   wsc.p + W_RRLX, s);
    pcg_initial_rr(*q.wy, np, q.A, q.idx, q.rowptr, q.diag, q.YPOS, q.b, q.fixy, q.fix_idx, q.nfix,
                   wsc.p + W_RRPY, s);
    pcg_initial_rr(*q.wy, np, q.A, q.idx, q.rowptr, q.diag, haty.p, q.b, q.fixy, q.fix_idx, q.nfix,
                   wsc.p + W_RRLY, s);
    k_ws_select<<<g, kBlock, 0, s>>>(q.XPOS, hatx.p, wsc.p, W_RRPX, W_RRLX, q.n_learned, q.n_solves, np);
    k_ws_select<<<g, kBlock, 0, s>>>(q.YPOS, haty.p, wsc.p, W_RRPY, W_RRLY, q.n_learned, q.n_solves, np);
    CFD_HIP(hipGetLastError());
    predicted = true;
}

void WarmStart::after_solves(const WsSolve& q, hipStream_t s) {
    const unsigned g = grid_for(np);
    if (predicted && net_used) {  // train on this step's correction
        const int nb = std::min<int>(grid_for(np), kRedBlocks);
        k_maxabs2<<<nb, kBlock, 0, s>>>(q.XPOS, ex.p, q.YPOS, ey.p, np, partials.p);
        k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SC);
        k_ws_target<<<g, kBlock, 0, s>>>(q.XPOS, q.YPOS, ex.p, ey.p, wsc.p, tgt.p, np);
        // Adam steps on evenly strided samples of about kWsTrainNodes nodes,
        // each sample shifted by a fraction of the stride
        const int stride = std::max(1, np / kWsTrainNodes);
        int ns = 0;
        for (int k = 0; k < kWsAdamSteps; ++k) {
            const int off = (int)((long long)stride * k / kWsAdamSteps);
            ns = mlp_gradient(mlp, feat.p + (size_t)off * kMlpIn, tgt.p + 2 * (size_t)off, np - off,
                              2, s, stride);
            mlp_adam(mlp, 1.0e-3f, s);
        }
        k_copy_loss<<<1, 1, 0, s>>>(mlp.loss.p, 1.0 / ns, q.loss);
    }
    predicted = net_used = false;
    k_ws_shift<<<g, kBlock, 0, s>>>(q.XPOS, q.YPOS, h1x.p, h2x.p, h3x.p, h1y.p, h2y.p, h3y.p,
                                    q.dtmin, wsc.p, np);
    CFD_HIP(hipGetLastError());
    nhist = std::min(nhist + 1, 3);
}

}  // namespace cfd
