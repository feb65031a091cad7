// Jacobi-preconditioned CG of biconjGrad.f90:8-62 (named biCG there).
//
// Reference sequence (one solve):
//   x(fix) = x_fix; y = A x; y(fix) = 1e30 x(fix); r = b - y; r(fix) = 0
//   if r.r < tol: return
//   p = r/diag; e = r.p; y = A p (fix rows 1e30 p); alfa = e/p.y; x += alfa p
//   while |e_old| > tol and k < 1000:
//     k++; r -= alfa y; z = r/diag; e = r.z; beta = e/e_old; p = beta p + z
//     y = A p (fix rows 1e30 p); alfa = e/p.y; x += alfa p; e_old = e
//
// On the device one iteration is two full-grid kernels, each followed by a
// one-block finalize that combines the block partials and updates the
// scalars in device memory:
//   A: x += alfa p (the previous iteration's update), r -= alfa y, z = r/diag,
//      partials of r.z;                       finalize: e, beta, k.
//   B: each row forms p_new = beta p_old + z for its columns on the fly, y = A
//      p_new, stores its own p_new (double-buffered p), partials of p.y;
//                                             finalize: alfa, the loop test.
// The host polls a done flag every few iterations.  Dot products combine
// block partials in a fixed tree, so the iteration sequence is deterministic
// (not the serial summation order of the reference).
#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {
namespace {

// scal[] slots
enum { S_ALFA = 0, S_BETA, S_EOLD, S_ENEW, S_PY, S_RR, S_K, S_DONE, S_STOP, S_CONV, S_TOL, S_MAXIT };

__global__ void k_fix_init(int n, int* __restrict__ fixflag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fixflag[i] = 0;
}

__global__ void k_fix_set(int nfix, const int* __restrict__ idx, const double* __restrict__ x_fix,
                          int* __restrict__ fixflag, double* __restrict__ x) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nfix) return;
    int i = idx[q] - 1;
    fixflag[i] = 1;
    x[i] = 1.0 * x_fix[q];
}

__global__ void k_scal_init(double* __restrict__ scal, double tol, int maxit) {
    for (int q = 0; q < 16; ++q) scal[q] = 0.0;
    scal[S_TOL] = tol;
    scal[S_MAXIT] = (double)maxit;
}

// r = -1*(A x) + b with fixed rows y = 1e30 x and r(fix) = 0; partials of r.r
__global__ void __launch_bounds__(kBlock)
k_residual(int n, const double* __restrict__ A, const int* __restrict__ idx,
           const int* __restrict__ rowptr, const int* __restrict__ fixflag,
           const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
           double* __restrict__ partials) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double ri;
        if (fixflag[i]) {
            ri = 0.0;
        } else {
            double dot = 0.0;
            for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) dot = dot + A[j] * x[idx[j] - 1];
            ri = -1.0 * dot + b[i];
        }
        r[i] = ri;
        v = ri * ri;
    }
    block_partial(v, OpSum(), partials);
}

__global__ void __launch_bounds__(256)
k_fin_residual(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
    double rr = combine_partials(partials, nb, 0.0, OpSum());
    if (threadIdx.x == 0) {
        scal[S_RR] = rr;
        if (rr < scal[S_TOL]) {
            scal[S_DONE] = 1.0;
            scal[S_CONV] = 1.0;
            scal[S_EOLD] = rr;
        }
    }
}

// first preconditioned residual: z = r/diag, partials of r.z
__global__ void __launch_bounds__(kBlock)
k_first_z(int n, const double* __restrict__ r, const double* __restrict__ diag,
          double* __restrict__ z, const double* __restrict__ scal, double* __restrict__ partials) {
    if (scal[S_DONE] != 0.0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double zi = r[i] / diag[i];
        z[i] = zi;
        v = r[i] * zi;
    }
    block_partial(v, OpSum(), partials);
}

__global__ void __launch_bounds__(256)
k_fin_first_z(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
    if (scal[S_DONE] != 0.0) return;
    double e = combine_partials(partials, nb, 0.0, OpSum());
    if (threadIdx.x == 0) scal[S_ENEW] = e;
}

// Kernel A of an iteration (see the file header)
__global__ void __launch_bounds__(kBlock)
k_iter_a(int n, double* __restrict__ x, const double* __restrict__ pcur, double* __restrict__ r,
         const double* __restrict__ y, const double* __restrict__ diag, double* __restrict__ z,
         const double* __restrict__ scal, double* __restrict__ partials) {
    if (scal[S_DONE] != 0.0) return;
    const double alfa = scal[S_ALFA];
    const bool stop = scal[S_STOP] != 0.0;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (stop) {  // the loop test failed: only the last x update remains
        if (i < n) x[i] = alfa * pcur[i] + x[i];
        return;
    }
    double v = 0.0;
    if (i < n) {
        x[i] = alfa * pcur[i] + x[i];
        double ri = -alfa * y[i] + r[i];
        r[i] = ri;
        double zi = ri / diag[i];
        z[i] = zi;
        v = ri * zi;
    }
    block_partial(v, OpSum(), partials);
}

__global__ void __launch_bounds__(256)
k_fin_a(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
    if (scal[S_DONE] != 0.0) return;
    if (scal[S_STOP] != 0.0) {
        if (threadIdx.x == 0) {
            scal[S_DONE] = 1.0;
            scal[S_CONV] = fabs(scal[S_EOLD]) > scal[S_TOL] ? 0.0 : 1.0;
        }
        return;
    }
    double e = combine_partials(partials, nb, 0.0, OpSum());
    if (threadIdx.x == 0) {
        scal[S_K] = scal[S_K] + 1.0;
        scal[S_ENEW] = e;
        scal[S_BETA] = e / scal[S_EOLD];
    }
}

// Kernel B: p_new = beta p_old + z on the fly (first: p_new = z), y = A p_new
__global__ void __launch_bounds__(kBlock)
k_iter_b(int n, const double* __restrict__ A, const int* __restrict__ idx,
         const int* __restrict__ rowptr, const int* __restrict__ fixflag,
         const double* __restrict__ pold, const double* __restrict__ z, double* __restrict__ pnew,
         double* __restrict__ y, int first, const double* __restrict__ scal,
         double* __restrict__ partials) {
    if (scal[S_DONE] != 0.0 || scal[S_STOP] != 0.0) return;
    const double beta = scal[S_BETA];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double pi = first ? z[i] : beta * pold[i] + z[i];
        double yi;
        if (fixflag[i]) {
            yi = 1.0e30 * pi;
        } else {
            double dot = 0.0;
            for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) {
                int c = idx[j] - 1;
                double pc = first ? z[c] : beta * pold[c] + z[c];
                dot = dot + A[j] * pc;
            }
            yi = dot;
        }
        pnew[i] = pi;
        y[i] = yi;
        v = pi * yi;
    }
    block_partial(v, OpSum(), partials);
}

__global__ void __launch_bounds__(256)
k_fin_b(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
    if (scal[S_DONE] != 0.0 || scal[S_STOP] != 0.0) return;
    double py = combine_partials(partials, nb, 0.0, OpSum());
    if (threadIdx.x == 0) {
        double e = scal[S_ENEW];
        scal[S_PY] = py;
        scal[S_ALFA] = e / py;
        scal[S_EOLD] = e;
        bool more = fabs(e) > scal[S_TOL] && scal[S_K] < scal[S_MAXIT];
        if (!more) scal[S_STOP] = 1.0;
    }
}

__global__ void k_store_result(const double* __restrict__ scal, double* __restrict__ iters,
                               double* __restrict__ rz) {
    *iters = scal[S_K];
    *rz = scal[S_EOLD];
}

}  // namespace

void pcg_store_result(const PcgWork& w, double* iters, double* rz, hipStream_t s) {
    k_store_result<<<1, 1, 0, s>>>(w.scal.p, iters, rz);
    CFD_HIP(hipGetLastError());
}

void PcgWork::ensure(int npoin) {
    if (npoin == n && r.p) return;
    n = npoin;
    nblocks = (int)grid_for(npoin);
    r.alloc(npoin);
    z.alloc(npoin);
    y.alloc(npoin);
    pa.alloc(npoin);
    pb.alloc(npoin);
    fixflag.alloc(npoin);
    partial.alloc(nblocks);
    scal.alloc(16);
}

PcgResult pcg_solve(PcgWork& w, int npoin, const double* A, const int* idx, const int* rowptr,
                    const double* diag, double* x, const double* b, const double* x_fix,
                    const int* x_fix_idx, int nfix, double tol, int maxit, int check_every,
                    hipStream_t s) {
    w.ensure(npoin);
    const int nb = w.nblocks;
    const dim3 g(nb), blk(kBlock), one(1), fin(256);
    double* scal = w.scal.p;
    double* part = w.partial.p;
    double h[16];
    auto poll = [&]() {
        CFD_HIP(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, s));
        CFD_HIP(hipStreamSynchronize(s));
    };
    double* pbuf[2] = {w.pa.p, w.pb.p};
    auto iter_a = [&](int cur) {
        k_iter_a<<<g, blk, 0, s>>>(npoin, x, pbuf[cur], w.r.p, w.y.p, diag, w.z.p, scal, part);
        k_fin_a<<<one, fin, 0, s>>>(part, nb, scal);
    };
    auto iter_b = [&](int cur, int first) {
        k_iter_b<<<g, blk, 0, s>>>(npoin, A, idx, rowptr, w.fixflag.p, pbuf[cur], w.z.p,
                                   pbuf[cur ^ 1], w.y.p, first, scal, part);
        k_fin_b<<<one, fin, 0, s>>>(part, nb, scal);
    };
    k_scal_init<<<one, one, 0, s>>>(scal, tol, maxit);
    k_fix_init<<<g, blk, 0, s>>>(npoin, w.fixflag.p);
    if (nfix > 0)
        k_fix_set<<<grid_for(nfix), blk, 0, s>>>(nfix, x_fix_idx, x_fix, w.fixflag.p, x);
    k_residual<<<g, blk, 0, s>>>(npoin, A, idx, rowptr, w.fixflag.p, x, b, w.r.p, part);
    k_fin_residual<<<one, fin, 0, s>>>(part, nb, scal);
    k_first_z<<<g, blk, 0, s>>>(npoin, w.r.p, diag, w.z.p, scal, part);
    k_fin_first_z<<<one, fin, 0, s>>>(part, nb, scal);
    iter_b(1, 1);  // p0 = z0 into pbuf[0]
    CFD_HIP(hipGetLastError());
    poll();        // most warm-started mesh solves stop here (r.r < tol)
    int cur = 0;   // pbuf[cur] holds the current p
    const int chunk = check_every > 0 ? check_every : 8;
    for (int launched = 0; h[S_DONE] == 0.0; launched += chunk) {
        if (h[S_STOP] != 0.0) {  // one A applies the final x update and ends the solve
            iter_a(cur);
            poll();
            break;
        }
        for (int q = 0; q < chunk; ++q) {
            iter_a(cur);
            iter_b(cur, 0);
            cur ^= 1;
        }
        CFD_HIP(hipGetLastError());
        poll();
        CFD_CHECK(launched <= 4 * maxit + 64, "PCG did not terminate");
    }
    PcgResult res;
    res.iters = (int)h[S_K];
    res.rz = h[S_EOLD];
    res.converged = h[S_CONV] != 0.0;
    return res;
}

}  // namespace cfd
