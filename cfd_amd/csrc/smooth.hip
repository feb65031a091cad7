// Start-up mesh smoothing: smoothing_mod::smoothing (smoothing.f90:21-75),
// the quality-driven node relocation the main program runs once before the
// time loop (ns2DComp.ALE.f90:76).
//
// The reference:
//   checkMesh: nodes of elements with quality mu < 0.85 become smoothable;
//     H_MIN_GLOBAL = min(1, min over elements of h);
//   two Laplacian passes: each smoothable free node moves to the mean of its
//     neighbours (or a tol_move-sized step towards it), restored when its
//     worst element gets worse;
//   up to 100 sweeps: each smoothable free node clears its flag and, when its
//     worst element is below 0.85, climbs the worst element's quality along a
//     one-sided-difference gradient (2 steps of up to 8 halving tries, the
//     tries accumulating), re-marking its neighbours when it moved; the
//     sweeps stop when the largest squared move is below 0.01 H_MIN_GLOBAL.
// Every pass updates the nodes in place, in index order, so each node reads
// the new positions of the lower-numbered neighbours moved before it.  On an
// O-grid numbered ring by ring that order is a chain through every node of
// the mesh (node (r, 0) follows node (r-1, last) through the seam), so the
// reference's result is inherently sequential.
//
// Two schedules share the per-node code below (same operations, same order,
// contraction off):
//   order 0 - the reference's node order, one lane walking each pass: the
//             reference's result bit for bit, for small meshes;
//   order 1 - colour by colour: a greedy colouring of the node graph (no two
//             nodes of one colour share an element, so nodes of a colour read
//             and write disjoint data), one lane per node, colours in
//             sequence.  The same per-node updates in another order: a
//             different, equally admissible smoothing (every accepted move
//             still raises the node's worst element).
#include <algorithm>

#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {

namespace {

constexpr int kSmMaxE = 20;          // MAX_ELEM_PER_NODE (smoothing.f90:6)
constexpr int kSmSweeps = 100;       // NITER
constexpr int kSmSteps = 2;          // MITER
constexpr int kSmTries = 8;          // NTRY
constexpr double kSmTolMetric = .85; // TOL_METRIC
constexpr double kSmTolDist = 1e-2;  // FACTOR_TOL_DIST
constexpr double kSmDelta = 1e-2;    // FACTOR_DELTA
constexpr double kSmPlus = 1.0;      // FACTOR_PLUS
constexpr double kSmStep = 3.0;      // FACTOR_STEP
constexpr double kTwoSqrt3 = 3.46410161513775;

struct SmDev {
    const int *inpoel, *esup1, *esup2, *psup1, *psup2, *fixed;
    double *X, *Y;
    int* smoothable;
    const double* tol_move;  // null: setTol_move not called
    const double* hmin;      // H_MIN_GLOBAL, device scalar
};

// smoothing.f90:344-355 mu and :357-365 h of the element's three nodes
__device__ inline double sm_mu(const SmDev& d, int e) {
    const int* n = d.inpoel + 3 * e;
    const double X1 = d.X[n[0] - 1], X2 = d.X[n[1] - 1], X3 = d.X[n[2] - 1];
    const double Y1 = d.Y[n[0] - 1], Y2 = d.Y[n[1] - 1], Y3 = d.Y[n[2] - 1];
    const double area = X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3);
    const double l1 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
    const double l2 = (X1 - X3) * (X1 - X3) + (Y1 - Y3) * (Y1 - Y3);
    const double l3 = (X2 - X1) * (X2 - X1) + (Y2 - Y1) * (Y2 - Y1);
    const double l = l1 + l2 + l3;
    return kTwoSqrt3 * area / l;
}

__device__ inline double sm_h(const SmDev& d, int e) {
    const int* n = d.inpoel + 3 * e;
    const double X1 = d.X[n[0] - 1], X2 = d.X[n[1] - 1], X3 = d.X[n[2] - 1];
    const double Y1 = d.Y[n[0] - 1], Y2 = d.Y[n[1] - 1], Y3 = d.Y[n[2] - 1];
    const double area = X2 * Y3 + X3 * Y1 + X1 * Y2 - (X2 * Y1 + X3 * Y2 + X1 * Y3);
    const double d1 = fabs(X3 - X2) + fabs(Y3 - Y2);
    const double d2 = fabs(X1 - X3) + fabs(Y1 - Y3);
    const double d3 = fabs(X2 - X1) + fabs(Y2 - Y1);
    const double dd = d1 + d2 + d3;
    return fabs(area) / dd;
}

// getMu_vec (smoothing.f90:174-197): the qualities of node ip's elements in
// esup order; *min_idx (0-based here) the first strict minimum below 1
__device__ inline int sm_mu_vec(const SmDev& d, int ip, double* v, int* min_idx) {
    const int b = d.esup2[ip - 1], cnt = d.esup2[ip] - b;
    double mu_min = 1;
    if (min_idx) *min_idx = 0;
    for (int k = 0; k < cnt; ++k) {
        const double mu1 = sm_mu(d, d.esup1[b + k] - 1);
        v[k] = mu1;
        if (min_idx && mu1 < mu_min) {
            mu_min = mu1;
            *min_idx = k;
        }
    }
    return cnt;
}

// getMu_min (:199-215) and getH_min (:272-287)
__device__ inline double sm_mu_min(const SmDev& d, int ip) {
    double m = 1;
    for (int s = d.esup2[ip - 1]; s < d.esup2[ip]; ++s) {
        const double mu1 = sm_mu(d, d.esup1[s] - 1);
        if (mu1 <= m) m = mu1;
    }
    return m;
}

__device__ inline double sm_h_min(const SmDev& d, int ip) {
    double m = 1;
    for (int s = d.esup2[ip - 1]; s < d.esup2[ip]; ++s) {
        const double h1 = sm_h(d, d.esup1[s] - 1);
        if (h1 < m) m = h1;
    }
    return m;
}

// laplacianSmoothing (:77-111), node ip
__device__ void sm_laplacian_node(const SmDev& d, int ip) {
    if (!(d.smoothable[ip - 1] && !d.fixed[ip - 1])) return;
    double xn = 0, yn = 0;
    const double mu_old = sm_mu_min(d, ip);
    const double xo = d.X[ip - 1], yo = d.Y[ip - 1];
    const int n = d.psup2[ip] - d.psup2[ip - 1];
    for (int s = d.psup2[ip - 1]; s < d.psup2[ip]; ++s) {
        xn = xn + d.X[d.psup1[s] - 1];
        yn = yn + d.Y[d.psup1[s] - 1];
    }
    xn = xn / n;
    yn = yn / n;
    if (d.tol_move) {
        const double dX = xn - d.X[ip - 1], dY = yn - d.Y[ip - 1];
        const double tm = d.tol_move[ip - 1];
        d.X[ip - 1] = d.X[ip - 1] + dX / (fabs(dX) + fabs(dY)) * tm / 2 * kSmStep;
        d.Y[ip - 1] = d.Y[ip - 1] + dY / (fabs(dX) + fabs(dY)) * tm / 2 * kSmStep;
    } else {
        d.X[ip - 1] = xn;
        d.Y[ip - 1] = yn;
    }
    if (sm_mu_min(d, ip) < mu_old) {  // if worse, restore
        d.X[ip - 1] = xo;
        d.Y[ip - 1] = yo;
    }
}

// moveIpoin (:113-171) with getG (:217-236) and getStep (:238-270); returns
// the squared distance moved.  The tries accumulate: a failed try is not
// undone before the next, half as long, is added.
__device__ double sm_move_node(const SmDev& d, int ip, double* mu_vec, int min_idx) {
    double gx[kSmMaxE], gy[kSmMaxE];
    const double x0 = d.X[ip - 1], y0 = d.Y[ip - 1];
    const int cnt = d.esup2[ip] - d.esup2[ip - 1];
    const double delta = *d.hmin * kSmDelta;
    for (int it = 0; it < kSmSteps; ++it) {
        const double mu_min = mu_vec[min_idx];
        const double xo = d.X[ip - 1], yo = d.Y[ip - 1];
        // getG: one-sided differences of every element's quality
        d.X[ip - 1] = d.X[ip - 1] + delta;
        sm_mu_vec(d, ip, gx, nullptr);
        for (int k = 0; k < cnt; ++k) gx[k] = (gx[k] - mu_vec[k]) / delta;
        d.X[ip - 1] = xo;
        d.Y[ip - 1] = d.Y[ip - 1] + delta;
        sm_mu_vec(d, ip, gy, nullptr);
        for (int k = 0; k < cnt; ++k) gy[k] = (gy[k] - mu_vec[k]) / delta;
        d.Y[ip - 1] = yo;
        // getStep: the largest step along the worst element's gradient before
        // another element's linearised quality falls to the worst one's
        const double gxm = gx[min_idx], gym = gy[min_idx];
        const double g2 = gxm * gxm + gym * gym;
        double step = d.tol_move ? d.tol_move[ip - 1] / ((fabs(gxm) + fabs(gym)) * (double)kSmSweeps)
                                 : sm_h_min(d, ip) * kSmStep / (fabs(gxm) + fabs(gym));
        for (int i = 0; i < cnt; ++i) {
            const double gg = gxm * gx[i] + gym * gy[i];
            if (gg < 0) {
                const double step1 = (mu_vec[i] - mu_min) / (g2 - gg);
                if (step1 < step) step = step1;
            }
        }
        bool accepted = false;
        for (int j = 0; j < kSmTries; ++j) {
            const double dx = step * gxm, dy = step * gym;
            d.X[ip - 1] = d.X[ip - 1] + dx;
            d.Y[ip - 1] = d.Y[ip - 1] + dy;
            int mi;
            sm_mu_vec(d, ip, mu_vec, &mi);
            if (mu_vec[mi] > mu_min * kSmPlus) {
                min_idx = mi;
                accepted = true;
                break;
            }
            step = .5 * step;
        }
        if (!accepted) {  // restore and stop
            d.X[ip - 1] = xo;
            d.Y[ip - 1] = yo;
            break;
        }
    }
    const double ex = d.X[ip - 1] - x0, ey = d.Y[ip - 1] - y0;
    const double d_move = ex * ex + ey * ey;
    if (d_move > __DBL_MIN__)  // update_list (:289-298): the neighbours again
        for (int s = d.psup2[ip - 1]; s < d.psup2[ip]; ++s) d.smoothable[d.psup1[s] - 1] = 1;
    return d_move;
}

// One node of an optimisation sweep (smoothing.f90:54-63); returns the
// squared move, 0 when the node was not moved
__device__ double sm_sweep_node(const SmDev& d, int ip) {
    if (!(d.smoothable[ip - 1] && !d.fixed[ip - 1])) return 0.0;
    d.smoothable[ip - 1] = 0;
    double mu_vec[kSmMaxE];
    int mi;
    sm_mu_vec(d, ip, mu_vec, &mi);
    if (mu_vec[mi] < kSmTolMetric) return sm_move_node(d, ip, mu_vec, mi);
    return 0.0;
}

// checkMesh (:367-385): smoothable marks and per-block partials of min h
// and of "some element below 0.85"
__global__ void __launch_bounds__(kBlock)
k_sm_check(int nelem, SmDev d, double* __restrict__ hpart, double* __restrict__ anypart) {
    double hm = 1.0, any = 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += gridDim.x * blockDim.x) {
        if (sm_mu(d, e) < kSmTolMetric) {
            const int* n = d.inpoel + 3 * e;
            for (int k = 0; k < 3; ++k) d.smoothable[n[k] - 1] = 1;
            any = 1.0;
        }
        const double h1 = sm_h(d, e);
        if (h1 < hm) hm = h1;
    }
    block_partial(hm, OpMin(), hpart);
    block_partial(any, OpMax(), anypart);
}

// H_MIN_GLOBAL = min(1, partials) and the "any" flag
__global__ void __launch_bounds__(256)
k_sm_check_fin(const double* __restrict__ hpart, const double* __restrict__ anypart, int nb,
               double* __restrict__ out) {
    const double h = combine_partials(hpart, nb, 1.0, OpMin());
    const double a = combine_partials(anypart, nb, 0.0, OpMax());
    if (threadIdx.x == 0) {
        out[0] = h;
        out[1] = a;
    }
}

// order 0: one lane walks a pass in node order
__global__ void k_sm_laplacian_seq(SmDev d, int npoin) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int ip = 1; ip <= npoin; ++ip) sm_laplacian_node(d, ip);
}

__global__ void k_sm_sweep_seq(SmDev d, int npoin, double* __restrict__ dmax) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double d_max = 0.0;
    for (int ip = 1; ip <= npoin; ++ip) {
        const double dm = sm_sweep_node(d, ip);
        if (dm > d_max) d_max = dm;
    }
    *dmax = d_max;
}

// order 1: one lane per node of a colour
__global__ void __launch_bounds__(kBlock)
k_sm_laplacian_colour(SmDev d, const int* __restrict__ nodes, int count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < count) sm_laplacian_node(d, nodes[q]);
}

__global__ void __launch_bounds__(kBlock)
k_sm_sweep_colour(SmDev d, const int* __restrict__ nodes, int count,
                  double* __restrict__ dmove) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < count) {
        const int ip = nodes[q];
        dmove[ip - 1] = sm_sweep_node(d, ip);
    }
}

// largest squared move of a colour sweep (moves are >= 0 or NaN; NaN is
// skipped as the reference's "d_move > d_max" skips it)
__global__ void __launch_bounds__(kBlock)
k_sm_max(const double* __restrict__ v, int n, double* __restrict__ partials) {
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (v[i] > m) m = v[i];
    block_partial(m, OpMax(), partials);
}

__global__ void __launch_bounds__(256)
k_sm_max_fin(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    const double m = combine_partials(partials, nb, 0.0, OpMax());
    if (threadIdx.x == 0) *out = m;
}

}  // namespace

void SmoothPlan::build(const Tables& t, const int* fixed_host, hipStream_t s) {
    const int n = t.npoin;
    std::vector<int> e2(n + 1), p2(n + 1), p1(t.npsup);
    t.esup2.download(e2.data(), n + 1, s);
    t.psup2.download(p2.data(), n + 1, s);
    t.psup1.download(p1.data(), t.npsup, s);
    CFD_HIP(hipStreamSynchronize(s));
    max_elems = 0;
    for (int i = 0; i < n; ++i) max_elems = std::max(max_elems, e2[i + 1] - e2[i]);
    // greedy colouring in node order: the smallest colour no lower-numbered
    // neighbour holds
    std::vector<int> colour(n, -1), seen;
    int ncol = 0;
    for (int i = 0; i < n; ++i) {
        seen.assign(ncol + 1, 0);
        for (int q = p2[i]; q < p2[i + 1]; ++q) {
            const int c = colour[p1[q] - 1];
            if (c >= 0) seen[c] = 1;
        }
        int c = 0;
        while (seen[c]) ++c;
        colour[i] = c;
        ncol = std::max(ncol, c + 1);
    }
    // free nodes by colour, ascending within a colour
    colour_ptr.assign(ncol + 1, 0);
    for (int i = 0; i < n; ++i)
        if (!fixed_host[i]) ++colour_ptr[colour[i] + 1];
    for (int c = 0; c < ncol; ++c) colour_ptr[c + 1] += colour_ptr[c];
    std::vector<int> nodes(std::max(1, colour_ptr[ncol])), fill(colour_ptr.begin(), colour_ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        if (!fixed_host[i]) nodes[fill[colour[i]]++] = i + 1;
    colour_nodes.alloc(nodes.size());
    colour_nodes.upload(nodes.data(), nodes.size(), s);
    fixed.alloc(std::max(1, n));
    fixed.upload(fixed_host, n, s);
    smoothable.alloc(std::max(1, n));
    dmove.alloc(std::max(1, n));
    partials.alloc(2 * (size_t)kRedBlocks);
    scal.alloc(4);
}

bool smooth_check(const Tables& t, SmoothPlan& plan, double* X, double* Y, double* hmin_host,
                  hipStream_t s) {
    SmDev d{t.inpoel.p, t.esup1.p, t.esup2.p, t.psup1.p, t.psup2.p, plan.fixed.p,
            X, Y, plan.smoothable.p, nullptr, plan.scal.p};
    plan.smoothable.zero(s);
    const int nb = std::min<int>(grid_for(t.nelem), kRedBlocks);
    k_sm_check<<<nb, kBlock, 0, s>>>(t.nelem, d, plan.partials.p, plan.partials.p + kRedBlocks);
    k_sm_check_fin<<<1, 256, 0, s>>>(plan.partials.p, plan.partials.p + kRedBlocks, nb, plan.scal.p);
    CFD_HIP(hipGetLastError());
    double hs[2];
    CFD_HIP(hipMemcpyAsync(hs, plan.scal.p, sizeof(hs), hipMemcpyDeviceToHost, s));
    CFD_HIP(hipStreamSynchronize(s));
    plan.hmin = hs[0];
    if (hmin_host) *hmin_host = hs[0];
    return hs[1] != 0.0;
}

int smooth_run(const Tables& t, SmoothPlan& plan, double* X, double* Y, const double* tol_move,
               int order, hipStream_t s) {
    const int n = t.npoin;
    SmDev d{t.inpoel.p, t.esup1.p, t.esup2.p, t.psup1.p, t.psup2.p, plan.fixed.p,
            X, Y, plan.smoothable.p, tol_move, plan.scal.p};
    const int ncol = (int)plan.colour_ptr.size() - 1;
    auto colour_pass = [&](auto launch) {
        for (int c = 0; c < ncol; ++c) {
            const int b = plan.colour_ptr[c], cnt = plan.colour_ptr[c + 1] - b;
            if (cnt > 0) launch(plan.colour_nodes.p + b, cnt);
        }
    };
    for (int it = 0; it < 2; ++it) {
        if (order == 0)
            k_sm_laplacian_seq<<<1, 1, 0, s>>>(d, n);
        else
            colour_pass([&](const int* nodes, int cnt) {
                k_sm_laplacian_colour<<<grid_for(cnt), kBlock, 0, s>>>(d, nodes, cnt);
            });
    }
    CFD_HIP(hipGetLastError());
    const double tol_dist = kSmTolDist * plan.hmin;
    const int nbm = std::min<int>(grid_for(n), kRedBlocks);
    int sweep = 1;
    for (; sweep <= kSmSweeps; ++sweep) {
        if (order == 0) {
            k_sm_sweep_seq<<<1, 1, 0, s>>>(d, n, plan.scal.p + 2);
        } else {
            plan.dmove.zero(s);
            colour_pass([&](const int* nodes, int cnt) {
                k_sm_sweep_colour<<<grid_for(cnt), kBlock, 0, s>>>(d, nodes, cnt, plan.dmove.p);
            });
            k_sm_max<<<nbm, kBlock, 0, s>>>(plan.dmove.p, n, plan.partials.p);
            k_sm_max_fin<<<1, 256, 0, s>>>(plan.partials.p, nbm, plan.scal.p + 2);
        }
        CFD_HIP(hipGetLastError());
        double dmax;
        CFD_HIP(hipMemcpyAsync(&dmax, plan.scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
        CFD_HIP(hipStreamSynchronize(s));
        if (dmax < tol_dist) break;
    }
    return std::min(sweep, kSmSweeps);
}

}  // namespace cfd
