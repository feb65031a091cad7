// Learned warm start of the two mesh-Laplace solves of fluidStructure
// (meshMove.f90:91-126), the reference's guess being the previous step's
// solution.
//
// Per axis the candidate is  xhat = E + sc * y,  where
//   E   = dt * (2 x1/dt1 - x2/dt2), the last two solutions extrapolated as
//         displacement rates (FP64);
//   y   = the MLP's output for this node (matrix cores, BF16, mlp.hip), from
//         the features [Ex, Ey]/se, [x1 - x2 (x and y)]/sd, node position
//         relative to the pivot / L, distance / L, 1;
//   sc  = the size of the last step's correction, max |x - E|.
// The network is trained online, one Adam step per time step, to map the
// features to the normalised correction (x - E)/sc of the solution the solver
// has just computed.  Each solve starts from xhat when xhat's first residual
// r.r is smaller than that of the previous solution, else from the previous
// solution, so the solve never starts worse than the reference's start.
#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {

namespace {

constexpr int kWsTrainNodes = 32768;

enum { W_DT1 = 0, W_DT2, W_SE, W_SD, W_SCX, W_SCY, W_RRPX, W_RRLX, W_RRPY, W_RRLY, W_N = 16 };

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
                              const double* __restrict__ h1y, const double* __restrict__ h2y,
                              const double* __restrict__ X, const double* __restrict__ Y, double xref,
                              double yref, double inv_l, const double* __restrict__ wsc,
                              unsigned short* __restrict__ feat, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ise = 1.0 / wsc[W_SE], isd = 1.0 / wsc[W_SD];
    const double dx = X[i] - xref, dy = Y[i] - yref;
    float f[kMlpIn] = {0.f};
    f[0] = (float)(ex[i] * ise);
    f[1] = (float)(ey[i] * ise);
    f[2] = (float)((h1x[i] - h2x[i]) * isd);
    f[3] = (float)((h1y[i] - h2y[i]) * isd);
    f[4] = (float)(dx * inv_l);
    f[5] = (float)(dy * inv_l);
    f[6] = (float)(sqrt(dx * dx + dy * dy) * inv_l);
    f[7] = 1.0f;
    unsigned short* p = feat + (size_t)i * kMlpIn;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) p[k] = to_bf16(f[k]);
}

__global__ void k_ws_predict(const double* __restrict__ ex, const double* __restrict__ ey,
                             const float* __restrict__ out, const double* __restrict__ wsc,
                             double* __restrict__ hatx, double* __restrict__ haty, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    hatx[i] = ex[i] + wsc[W_SCX] * (double)out[2 * (size_t)i];
    haty[i] = ey[i] + wsc[W_SCY] * (double)out[2 * (size_t)i + 1];
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
    tgt[2 * (size_t)i] = (float)((x[i] - ex[i]) / wsc[W_SCX]);
    tgt[2 * (size_t)i + 1] = (float)((y[i] - ey[i]) / wsc[W_SCY]);
}

// history shift and the rates' time steps
__global__ void k_ws_shift(const double* __restrict__ x, const double* __restrict__ y,
                           double* __restrict__ h1x, double* __restrict__ h2x,
                           double* __restrict__ h1y, double* __restrict__ h2y,
                           const double* __restrict__ dtp, double* __restrict__ wsc, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        h2x[i] = h1x[i];
        h1x[i] = x[i];
        h2y[i] = h1y[i];
        h1y[i] = y[i];
    }
    if (i == 0) {  // read by the next step only
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
    inv_l = l > 0.0 ? 1.0 / l : 1.0;
    for (auto* b : {&h1x, &h2x, &h1y, &h2y, &ex, &ey, &hatx, &haty}) b->alloc(np);
    feat.alloc(kMlpIn * (size_t)np);
    out.alloc(2 * (size_t)np);
    tgt.alloc(2 * (size_t)np);
    wsc.alloc(W_N);
    wsc.zero(s);
    partials.alloc(kRedBlocks);
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
    k_ws_extrap<<<g, kBlock, 0, s>>>(h1x.p, h2x.p, h1y.p, h2y.p, q.dtmin, wsc.p, ex.p, ey.p, np);
    k_maxabs2<<<nb, kBlock, 0, s>>>(ex.p, nullptr, ey.p, nullptr, np, partials.p);
    k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SE);
    k_maxabs2<<<nb, kBlock, 0, s>>>(h1x.p, h2x.p, h1y.p, h2y.p, np, partials.p);
    k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SD);
    k_ws_features<<<g, kBlock, 0, s>>>(ex.p, ey.p, h1x.p, h2x.p, h1y.p, h2y.p, q.X, q.Y, q.xref,
                                       q.yref, inv_l, wsc.p, feat.p, np);
    mlp_forward(mlp, feat.p, np, out.p, s);
    k_ws_predict<<<g, kBlock, 0, s>>>(ex.p, ey.p, out.p, wsc.p, hatx.p, haty.p, np);
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
    const int nb = std::min<int>(grid_for(np), kRedBlocks);
    if (predicted) {  // train on this step's correction
        k_maxabs2<<<nb, kBlock, 0, s>>>(q.XPOS, ex.p, q.XPOS, ex.p, np, partials.p);
        k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SCX);
        k_maxabs2<<<nb, kBlock, 0, s>>>(q.YPOS, ey.p, q.YPOS, ey.p, np, partials.p);
        k_max_to<<<1, 256, 0, s>>>(partials.p, nb, wsc.p + W_SCY);
        k_ws_target<<<g, kBlock, 0, s>>>(q.XPOS, q.YPOS, ex.p, ey.p, wsc.p, tgt.p, np);
        // one Adam step on an evenly strided sample of about kWsTrainNodes nodes
        const int stride = std::max(1, np / kWsTrainNodes);
        const int ns = mlp_gradient(mlp, feat.p, tgt.p, np, 2, s, stride);
        mlp_adam(mlp, 1.0e-3f, s);
        k_copy_loss<<<1, 1, 0, s>>>(mlp.loss.p, 1.0 / ns, q.loss);
    }
    predicted = false;
    k_ws_shift<<<g, kBlock, 0, s>>>(q.XPOS, q.YPOS, h1x.p, h2x.p, h1y.p, h2y.p, q.dtmin, wsc.p, np);
    CFD_HIP(hipGetLastError());
    nhist = std::min(nhist + 1, 2);
}

}  // namespace cfd
