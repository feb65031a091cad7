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
// scalars:
//   A: x += alfa p (the previous iteration's update), r -= alfa y, z = r/diag,
//      partials of r.z;                       finalize: e, beta, k.
//   B: each row forms p_new = beta p_old + z for its columns on the fly, y = A
//      p_new, stores its own p_new (double-buffered p), partials of p.y;
//                                             finalize: alfa, the loop test.
// The scalars live in device memory.  Kernels queued after a loop has ended
// return at once (done flag), so the host queues batches of iterations
// between polls.  Several independent systems on one matrix (the x and y mesh
// solves) run interleaved and share the polls.  Dot products combine block
// partials in a fixed tree, so the iteration sequence is deterministic (not
// the serial summation order of the reference).
//
// Two systems on one matrix (the x and y mesh solves) iterate in lockstep
// through paired kernels: one launch per kernel for both, with the matrix
// entries staged once for both products, and row blocks in XCD order
// (xcd_block: a row's neighbours' blocks mostly on the same XCD).  Each
// system's arithmetic, logical block decomposition and partial order are
// those of the single-system kernels, so the paired iterations are bit for
// bit the single ones; a system that has finished is skipped inside the
// paired kernels.
//
// Relative mode (rel = 1): the two tests use tol^2 times the r.r and r.z of
// the residual of the guess that is zero off the fixed nodes, so the stopping
// point does not depend on the initial guess (warm starts compare at equal
// tolerance).
#include <mutex>

#include "cfd_internal.hpp"
#include "reduce.hpp"

namespace cfd {
namespace {

// scal[] slots
enum {
    S_ALFA = 0, S_BETA, S_EOLD, S_ENEW, S_PY, S_RR, S_K, S_DONE, S_STOP, S_CONV, S_TOL, S_MAXIT,
    S_TOLR, S_TOLZ, S_NSCAL = 16
};

// The CSR entries of a block's rows [i0, i1) are one contiguous range
// [rowptr[i0], rowptr[i1]).  They are staged through LDS in chunks with
// coalesced loads; each row then visits its own entries in order,
// f(a_ij, column j (0-based)).  Every thread of the block must call this
// (rows without work pass an empty range b == e).
constexpr int kSpChunk = 2048;

template <class F>
__device__ inline void staged_rows(const double* __restrict__ A, const int* __restrict__ idx,
                                   int s0, int s1, int b, int e, double* sA, int* sI, F f) {
    for (int c0 = s0; c0 < s1; c0 += kSpChunk) {
        const int c1 = min(c0 + kSpChunk, s1);
        __syncthreads();
        for (int q = c0 + (int)threadIdx.x; q < c1; q += (int)blockDim.x) {
            sA[q - c0] = A[q];
            sI[q - c0] = idx[q] - 1;
        }
        __syncthreads();
        const int lo = max(b, c0), hi = min(e, c1);
        for (int j = lo; j < hi; ++j) f(sA[j - c0], sI[j - c0]);
    }
}

// fixflag = 0; scalars reset (block 0)
__global__ void k_fix_init(int n, int* __restrict__ fixflag, double* __restrict__ scal, double tol,
                           int maxit) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fixflag[i] = 0;
    if (i == 0) {
        for (int q = 0; q < S_NSCAL; ++q) scal[q] = 0.0;
        scal[S_TOL] = tol;
        scal[S_MAXIT] = (double)maxit;
    }
}

__global__ void k_fix_set(int nfix, const int* __restrict__ idx, const double* __restrict__ x_fix,
                          int* __restrict__ fixflag, double* __restrict__ x) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nfix) return;
    int i = idx[q] - 1;
    fixflag[i] = 1;
    x[i] = 1.0 * x_fix[q];
}

// r = -1*(A x) + b with fixed rows y = 1e30 x and r(fix) = 0; r.r and, in
// relative mode, rd.rd and rd.(rd/diag) for the residual rd of the guess that
// is zero off the fixed nodes; the first residual test
__global__ void __launch_bounds__(kBlock)
k_residual(int n, const double* __restrict__ A, const int* __restrict__ idx,
           const int* __restrict__ rowptr, const int* __restrict__ fixflag,
           const double* __restrict__ diag, const double* __restrict__ x,
           const double* __restrict__ b, double* __restrict__ r, int rel,
           double* __restrict__ partials) {
    __shared__ double sA[kSpChunk];
    __shared__ int sI[kSpChunk];
    const int i0 = blockIdx.x * blockDim.x, i = i0 + threadIdx.x;
    const bool act = i < n;
    const bool row = act && !fixflag[i];
    double dot = 0.0, dotd = 0.0;
    staged_rows(A, idx, rowptr[i0], rowptr[min(i0 + (int)blockDim.x, n)], row ? rowptr[i] : 0,
                row ? rowptr[i + 1] : 0, sA, sI, [&](double a, int c) {
                    const double xc = x[c];
                    dot = dot + a * xc;
                    if (rel && fixflag[c]) dotd = dotd + a * xc;
                });
    double v = 0.0, vdr = 0.0, vdz = 0.0;
    if (act) {
        double ri = 0.0;
        if (row) {
            ri = -1.0 * dot + b[i];
            if (rel) {
                const double rd = -1.0 * dotd + b[i];
                vdr = rd * rd;
                vdz = rd * (rd / diag[i]);
            }
        }
        r[i] = ri;
        v = ri * ri;
    }
    const int nb = gridDim.x;
    block_partial(v, OpSum(), partials);
    if (rel) {
        block_partial(vdr, OpSum(), partials + nb);
        block_partial(vdz, OpSum(), partials + 2 * nb);
    }
}

// the first residual test (one block)
__global__ void __launch_bounds__(256)
k_fin_residual(const double* __restrict__ partials, int nb, int rel, double* __restrict__ scal) {
    const double rr = combine_partials(partials, nb, 0.0, OpSum());
    double dr = 0.0, dz = 0.0;
    if (rel) {
        dr = combine_partials(partials + nb, nb, 0.0, OpSum());
        dz = combine_partials(partials + 2 * nb, nb, 0.0, OpSum());
    }
    if (threadIdx.x == 0) {
        const double tol = scal[S_TOL];
        const double tolr = rel ? tol * tol * dr : tol;
        const double tolz = rel ? tol * tol * dz : tol;
        scal[S_TOLR] = tolr;
        scal[S_TOLZ] = tolz;
        scal[S_RR] = rr;
        if (rr < tolr) {
            scal[S_DONE] = 1.0;
            scal[S_CONV] = 1.0;
            scal[S_EOLD] = rr;
        }
    }
}

// first preconditioned residual: z = r/diag, partials of r.z
__global__ void __launch_bounds__(kBlock)
k_first_z(int n, const double* __restrict__ r, const double* __restrict__ diag,
          double* __restrict__ z, const double* __restrict__ scal,
          double* __restrict__ partials) {
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
    const double e = combine_partials(partials, nb, 0.0, OpSum());
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
            scal[S_CONV] = fabs(scal[S_EOLD]) > scal[S_TOLZ] ? 0.0 : 1.0;
        }
        return;
    }
    const double e = combine_partials(partials, nb, 0.0, OpSum());
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
    __shared__ double sA[kSpChunk];
    __shared__ int sI[kSpChunk];
    const double beta = scal[S_BETA];
    const int i0 = blockIdx.x * blockDim.x, i = i0 + threadIdx.x;
    const bool act = i < n;
    const bool row = act && !fixflag[i];
    double dot = 0.0;
    staged_rows(A, idx, rowptr[i0], rowptr[min(i0 + (int)blockDim.x, n)], row ? rowptr[i] : 0,
                row ? rowptr[i + 1] : 0, sA, sI, [&](double a, int c) {
                    const double pc = first ? z[c] : beta * pold[c] + z[c];
                    dot = dot + a * pc;
                });
    double v = 0.0;
    if (act) {
        double pi = first ? z[i] : beta * pold[i] + z[i];
        double yi;
        if (!row) {
            yi = 1.0e30 * pi;
        } else {
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
    const double py = combine_partials(partials, nb, 0.0, OpSum());
    if (threadIdx.x == 0) {
        double e = scal[S_ENEW];
        scal[S_PY] = py;
        scal[S_ALFA] = e / py;
        scal[S_EOLD] = e;
        bool more = fabs(e) > scal[S_TOLZ] && scal[S_K] < scal[S_MAXIT];
        if (!more) scal[S_STOP] = 1.0;
    }
}

// ---- paired systems (0 and 1): the kernels above, both systems per launch
struct PairSys {
    double *x, *r, *z, *y, *partial, *scal;
    const double *pcur, *pold;
    double* pnew;
    const int* fixflag;
};

__global__ void __launch_bounds__(kBlock)
k_iter_a2(int n, PairSys s0, PairSys s1, const double* __restrict__ diag) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);  // logical row block
    const int i = lb * blockDim.x + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const PairSys& q = k ? s1 : s0;
        if (q.scal[S_DONE] != 0.0) continue;  // uniform over the grid
        const double alfa = q.scal[S_ALFA];
        if (q.scal[S_STOP] != 0.0) {
            if (i < n) q.x[i] = alfa * q.pcur[i] + q.x[i];
            continue;
        }
        double v = 0.0;
        if (i < n) {
            q.x[i] = alfa * q.pcur[i] + q.x[i];
            double ri = -alfa * q.y[i] + q.r[i];
            q.r[i] = ri;
            double zi = ri / diag[i];
            q.z[i] = zi;
            v = ri * zi;
        }
        block_partial(v, OpSum(), q.partial, lb);
    }
}

// block k finalizes system k (k_fin_a's body)
__global__ void __launch_bounds__(256)
k_fin_a2(PairSys s0, PairSys s1, int nb) {
    const PairSys& q = blockIdx.x ? s1 : s0;
    double* scal = q.scal;
    if (scal[S_DONE] != 0.0) return;
    if (scal[S_STOP] != 0.0) {
        if (threadIdx.x == 0) {
            scal[S_DONE] = 1.0;
            scal[S_CONV] = fabs(scal[S_EOLD]) > scal[S_TOLZ] ? 0.0 : 1.0;
        }
        return;
    }
    const double e = combine_partials(q.partial, nb, 0.0, OpSum());
    if (threadIdx.x == 0) {
        scal[S_K] = scal[S_K] + 1.0;
        scal[S_ENEW] = e;
        scal[S_BETA] = e / scal[S_EOLD];
    }
}

// k_iter_b (not the first iteration) for both systems, A staged once
__global__ void __launch_bounds__(kBlock)
k_iter_b2(int n, const double* __restrict__ A, const int* __restrict__ idx,
          const int* __restrict__ rowptr, PairSys s0, PairSys s1) {
    const bool on0 = s0.scal[S_DONE] == 0.0 && s0.scal[S_STOP] == 0.0;
    const bool on1 = s1.scal[S_DONE] == 0.0 && s1.scal[S_STOP] == 0.0;
    if (!on0 && !on1) return;
    __shared__ double sA[kSpChunk];
    __shared__ int sI[kSpChunk];
    const double beta0 = s0.scal[S_BETA], beta1 = s1.scal[S_BETA];
    const int lb = xcd_block(blockIdx.x, gridDim.x);  // logical row block
    const int i0 = lb * blockDim.x, i = i0 + threadIdx.x;
    const bool act = i < n;
    const bool row0 = act && !s0.fixflag[i], row1 = act && !s1.fixflag[i];
    const bool row = (on0 && row0) || (on1 && row1);
    double dot0 = 0.0, dot1 = 0.0;
    staged_rows(A, idx, rowptr[i0], rowptr[min(i0 + (int)blockDim.x, n)], row ? rowptr[i] : 0,
                row ? rowptr[i + 1] : 0, sA, sI, [&](double a, int c) {
                    if (on0 && row0) dot0 = dot0 + a * (beta0 * s0.pold[c] + s0.z[c]);
                    if (on1 && row1) dot1 = dot1 + a * (beta1 * s1.pold[c] + s1.z[c]);
                });
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const PairSys& q = k ? s1 : s0;
        if (!(k ? on1 : on0)) continue;  // uniform over the grid
        double v = 0.0;
        if (act) {
            const double pi = (k ? beta1 : beta0) * q.pold[i] + q.z[i];
            const double yi = (k ? row1 : row0) ? (k ? dot1 : dot0) : 1.0e30 * pi;
            q.pnew[i] = pi;
            q.y[i] = yi;
            v = pi * yi;
        }
        block_partial(v, OpSum(), q.partial, lb);
    }
}

// block k finalizes system k (k_fin_b's body)
__global__ void __launch_bounds__(256)
k_fin_b2(PairSys s0, PairSys s1, int nb) {
    const PairSys& q = blockIdx.x ? s1 : s0;
    double* scal = q.scal;
    if (scal[S_DONE] != 0.0 || scal[S_STOP] != 0.0) return;
    const double py = combine_partials(q.partial, nb, 0.0, OpSum());
    if (threadIdx.x == 0) {
        double e = scal[S_ENEW];
        scal[S_PY] = py;
        scal[S_ALFA] = e / py;
        scal[S_EOLD] = e;
        bool more = fabs(e) > scal[S_TOLZ] && scal[S_K] < scal[S_MAXIT];
        if (!more) scal[S_STOP] = 1.0;
    }
}

// ---- ELL-order SpMV variants of k_iter_b / k_iter_b2: the same rows, the
// same per-row summation order and block decomposition, the entries read
// coalesced from the ELL copy (no LDS staging, no barriers before the
// block partial)
template <int kW>
__device__ inline double ell_row_dot(const EllMat& m, int n, int i, int len, double beta,
                                     const double* __restrict__ pold, const double* __restrict__ z,
                                     bool first) {
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        if (k < len) {
            const size_t q = (size_t)k * n + i;
            const double a = m.val[q];
            const int c = m.col[q];
            const double pc = first ? z[c] : beta * pold[c] + z[c];
            dot = dot + a * pc;
        }
    }
    return dot;
}

template <int kW>
__global__ void __launch_bounds__(kBlock)
k_iter_b_ell(int n, EllMat m, const int* __restrict__ fixflag, const double* __restrict__ pold,
             const double* __restrict__ z, double* __restrict__ pnew, double* __restrict__ y,
             int first, const double* __restrict__ scal, double* __restrict__ partials) {
    if (scal[S_DONE] != 0.0 || scal[S_STOP] != 0.0) return;
    const double beta = scal[S_BETA];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const bool row = !fixflag[i];
        const double dot =
            row ? ell_row_dot<kW>(m, n, i, m.rowptr[i + 1] - m.rowptr[i], beta, pold, z, first)
                : 0.0;
        const double pi = first ? z[i] : beta * pold[i] + z[i];
        const double yi = row ? dot : 1.0e30 * pi;
        pnew[i] = pi;
        y[i] = yi;
        v = pi * yi;
    }
    block_partial(v, OpSum(), partials);
}

template <int kW>
__global__ void __launch_bounds__(kBlock)
k_iter_b2_ell(int n, EllMat m, PairSys s0, PairSys s1) {
    const bool on0 = s0.scal[S_DONE] == 0.0 && s0.scal[S_STOP] == 0.0;
    const bool on1 = s1.scal[S_DONE] == 0.0 && s1.scal[S_STOP] == 0.0;
    if (!on0 && !on1) return;
    const double beta0 = s0.scal[S_BETA], beta1 = s1.scal[S_BETA];
    const int lb = xcd_block(blockIdx.x, gridDim.x);  // logical row block
    const int i = lb * blockDim.x + threadIdx.x;
    const bool act = i < n;
    const bool row0 = act && !s0.fixflag[i], row1 = act && !s1.fixflag[i];
    double dot0 = 0.0, dot1 = 0.0;
    if ((on0 && row0) || (on1 && row1)) {
        const int len = m.rowptr[i + 1] - m.rowptr[i];
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            if (k < len) {
                const size_t q = (size_t)k * n + i;
                const double a = m.val[q];
                const int c = m.col[q];
                if (on0 && row0) dot0 = dot0 + a * (beta0 * s0.pold[c] + s0.z[c]);
                if (on1 && row1) dot1 = dot1 + a * (beta1 * s1.pold[c] + s1.z[c]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const PairSys& q = k ? s1 : s0;
        if (!(k ? on1 : on0)) continue;  // uniform over the grid
        double v = 0.0;
        if (act) {
            const double pi = (k ? beta1 : beta0) * q.pold[i] + q.z[i];
            const double yi = (k ? row1 : row0) ? (k ? dot1 : dot0) : 1.0e30 * pi;
            q.pnew[i] = pi;
            q.y[i] = yi;
            v = pi * yi;
        }
        block_partial(v, OpSum(), q.partial, lb);
    }
}

// ELL column layout and values (EllLap::build / refresh)
__global__ void k_ell_col(int n, int width, const int* __restrict__ rowptr,
                          const int* __restrict__ idx, int* __restrict__ col) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = rowptr[i], len = rowptr[i + 1] - b;
    for (int k = 0; k < width; ++k) col[(size_t)k * n + i] = k < len ? idx[b + k] - 1 : i;
}

__global__ void k_ell_val(int n, int width, const int* __restrict__ rowptr,
                          const double* __restrict__ A, double* __restrict__ val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = rowptr[i], len = rowptr[i + 1] - b;
    for (int k = 0; k < width; ++k) val[(size_t)k * n + i] = k < len ? A[b + k] : 0.0;
}

__global__ void k_store_result(const double* __restrict__ scal, double* __restrict__ iters,
                               double* __restrict__ rz, double* __restrict__ itsum,
                               double* __restrict__ unconv) {
    *iters = scal[S_K];
    *rz = scal[S_EOLD];
    if (itsum) *itsum = *itsum + scal[S_K];
    if (unconv && scal[S_CONV] == 0.0) *unconv = *unconv + 1.0;
}

__global__ void __launch_bounds__(256)
k_fin_sum(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    const double s = combine_partials(partials, nb, 0.0, OpSum());
    if (threadIdx.x == 0) *out = s;
}

}  // namespace

void pcg_initial_rr(PcgWork& w, int npoin, const double* A, const int* idx, const int* rowptr,
                    const double* diag, const double* x, const double* b, const double* x_fix,
                    const int* x_fix_idx, int nfix, double* rr_out, hipStream_t s) {
    w.ensure(npoin);
    const dim3 g(grid_for(npoin)), blk(kBlock);
    CFD_HIP(hipMemcpyAsync(w.z.p, x, sizeof(double) * npoin, hipMemcpyDeviceToDevice, s));
    k_fix_init<<<g, blk, 0, s>>>(npoin, w.fixflag.p, w.scal.p, 1.0, 0);
    if (nfix > 0)
        k_fix_set<<<grid_for(nfix), blk, 0, s>>>(nfix, x_fix_idx, x_fix, w.fixflag.p, w.z.p);
    k_residual<<<g, blk, 0, s>>>(npoin, A, idx, rowptr, w.fixflag.p, diag, w.z.p, b, w.r.p, 0,
                                 w.partial.p);
    k_fin_sum<<<1, 256, 0, s>>>(w.partial.p, (int)g.x, rr_out);
    CFD_HIP(hipGetLastError());
}

void pcg_store_result(const PcgWork& w, double* iters, double* rz, double* itsum, hipStream_t s,
                      double* unconv) {
    k_store_result<<<1, 1, 0, s>>>(w.scal.p, iters, rz, itsum, unconv);
    CFD_HIP(hipGetLastError());
}

bool EllLap::build(const Tables& t, hipStream_t s) {
    n = t.npoin;
    width = 0;
    if (n <= 0 || t.nnz <= 0) return false;
    std::vector<int> rp(n + 1);  // the longest row, on the host (once per session)
    t.rowptr.download(rp.data(), n + 1, s);
    CFD_HIP(hipStreamSynchronize(s));
    int w = 0;
    for (int i = 0; i < n; ++i) w = std::max(w, rp[i + 1] - rp[i]);
    if (w <= 0 || w > kEllMaxWidth || (double)w * n > 1.5 * (double)t.nnz) return false;
    width = w;
    col.alloc((size_t)w * n);
    val.alloc((size_t)w * n);
    k_ell_col<<<grid_for(n), kBlock, 0, s>>>(n, w, t.rowptr.p, t.idx.p, col.p);
    CFD_HIP(hipGetLastError());
    return true;
}

void EllLap::refresh(const Tables& t, const double* A, hipStream_t s) {
    if (!width) return;
    k_ell_val<<<grid_for(n), kBlock, 0, s>>>(n, width, t.rowptr.p, A, val.p);
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
    partial.alloc(3 * (size_t)nblocks);
    scal.alloc(S_NSCAL);
    hpin.alloc(S_NSCAL);
}

// launches the ELL SpMV kernel of the width class that holds m.width
template <class F8, class F16>
static void ell_dispatch(const EllMat& m, F8 f8, F16 f16) {
    if (m.width <= 8) f8();
    else f16();
}

void pcg_launch_multi(const PcgSystem* sys, int nsys, int npoin, const double* A, const int* idx,
                      const int* rowptr, const double* diag, double tol, int maxit, int rel,
                      hipStream_t s, const EllMat* ell) {
    CFD_CHECK(npoin > 0, "arg: PCG needs npoin > 0");
    constexpr int kMaxSys = 2;
    CFD_CHECK(nsys >= 1 && nsys <= kMaxSys, "internal: PCG systems per launch");
    const dim3 g(grid_for(npoin)), blk(kBlock);
    int cur[kMaxSys] = {0, 0};  // pbuf[cur] holds the current p
    const int nb = (int)g.x;
    const dim3 one(1), fin(256);
    auto iter_a = [&](int k) {
        PcgWork& w = *sys[k].w;
        double* pbuf[2] = {w.pa.p, w.pb.p};
        k_iter_a<<<g, blk, 0, s>>>(npoin, sys[k].x, pbuf[cur[k]], w.r.p, w.y.p, diag, w.z.p,
                                   w.scal.p, w.partial.p);
        k_fin_a<<<one, fin, 0, s>>>(w.partial.p, nb, w.scal.p);
    };
    auto iter_b = [&](int k, int first) {
        PcgWork& w = *sys[k].w;
        double* pbuf[2] = {w.pa.p, w.pb.p};
        if (ell) {
            const EllMat m = *ell;
            ell_dispatch(
                m,
                [&] {
                    k_iter_b_ell<8><<<g, blk, 0, s>>>(npoin, m, w.fixflag.p, pbuf[cur[k]], w.z.p,
                                                      pbuf[cur[k] ^ 1], w.y.p, first, w.scal.p,
                                                      w.partial.p);
                },
                [&] {
                    k_iter_b_ell<kEllMaxWidth><<<g, blk, 0, s>>>(
                        npoin, m, w.fixflag.p, pbuf[cur[k]], w.z.p, pbuf[cur[k] ^ 1], w.y.p, first,
                        w.scal.p, w.partial.p);
                });
        } else {
            k_iter_b<<<g, blk, 0, s>>>(npoin, A, idx, rowptr, w.fixflag.p, pbuf[cur[k]], w.z.p,
                                       pbuf[cur[k] ^ 1], w.y.p, first, w.scal.p, w.partial.p);
        }
        k_fin_b<<<one, fin, 0, s>>>(w.partial.p, nb, w.scal.p);
    };
    for (int k = 0; k < nsys; ++k) {
        PcgWork& w = *sys[k].w;
        w.ensure(npoin);
        k_fix_init<<<g, blk, 0, s>>>(npoin, w.fixflag.p, w.scal.p, tol, maxit);
        if (sys[k].nfix > 0)
            k_fix_set<<<grid_for(sys[k].nfix), blk, 0, s>>>(sys[k].nfix, sys[k].x_fix_idx,
                                                            sys[k].x_fix, w.fixflag.p, sys[k].x);
        k_residual<<<g, blk, 0, s>>>(npoin, A, idx, rowptr, w.fixflag.p, diag, sys[k].x, sys[k].b,
                                     w.r.p, rel, w.partial.p);
        k_fin_residual<<<one, fin, 0, s>>>(w.partial.p, nb, rel, w.scal.p);
        k_first_z<<<g, blk, 0, s>>>(npoin, w.r.p, diag, w.z.p, w.scal.p, w.partial.p);
        k_fin_first_z<<<one, fin, 0, s>>>(w.partial.p, nb, w.scal.p);
        cur[k] = 1;
        iter_b(k, 1);  // p0 = z0 into pbuf[0]
        cur[k] = 0;
    }
    CFD_HIP(hipGetLastError());
    bool done[kMaxSys] = {false, false}, stop[kMaxSys] = {false, false};
    double local[kMaxSys][S_NSCAL];  // the polls' host copy where no pinned buffer exists
    auto poll = [&]() {
        for (int k = 0; k < nsys; ++k) {
            double* h = sys[k].w->hpin.p ? sys[k].w->hpin.p : local[k];
            CFD_HIP(hipMemcpyAsync(h, sys[k].w->scal.p, S_NSCAL * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
        }
        CFD_HIP(hipStreamSynchronize(s));
        bool all = true;
        for (int k = 0; k < nsys; ++k) {
            const double* h = sys[k].w->hpin.p ? sys[k].w->hpin.p : local[k];
            done[k] = h[S_DONE] != 0.0;
            stop[k] = h[S_STOP] != 0.0;
            all = all && done[k];
        }
        return all;
    };
    // two systems: the paired kernels (both p buffers flip together; a
    // system that has stopped or finished is skipped on the device)
    auto pair = [&](int k) {
        PcgWork& w = *sys[k].w;
        double* pbuf[2] = {w.pa.p, w.pb.p};
        return PairSys{sys[k].x, w.r.p, w.z.p, w.y.p, w.partial.p, w.scal.p, pbuf[cur[k]],
                       pbuf[cur[k]], pbuf[cur[k] ^ 1], w.fixflag.p};
    };
    // most warm-started mesh solves stop at the first residual test; batches
    // grow while a solve runs on (few polls, few no-op launches)
    int chunk = 4;
    for (long launched = 0; !poll(); launched += chunk) {
        if (nsys == 2 && !done[0] && !done[1] && cur[0] == cur[1]) {
            for (int q = 0; q < chunk; ++q) {
                const PairSys p0 = pair(0), p1 = pair(1);
                k_iter_a2<<<g, blk, 0, s>>>(npoin, p0, p1, diag);
                k_fin_a2<<<dim3(2), fin, 0, s>>>(p0, p1, nb);
                if (ell) {
                    const EllMat m = *ell;
                    ell_dispatch(
                        m, [&] { k_iter_b2_ell<8><<<g, blk, 0, s>>>(npoin, m, p0, p1); },
                        [&] { k_iter_b2_ell<kEllMaxWidth><<<g, blk, 0, s>>>(npoin, m, p0, p1); });
                } else {
                    k_iter_b2<<<g, blk, 0, s>>>(npoin, A, idx, rowptr, p0, p1);
                }
                k_fin_b2<<<dim3(2), fin, 0, s>>>(p0, p1, nb);
                cur[0] ^= 1;
                cur[1] ^= 1;
            }
            CFD_HIP(hipGetLastError());
            CFD_CHECK(launched <= 4L * maxit + 64, "PCG did not terminate");
            chunk = std::min(chunk * 2, 16);
            continue;
        }
        for (int k = 0; k < nsys; ++k)
            if (!done[k] && stop[k]) iter_a(k);  // the final x update ends the solve
        for (int q = 0; q < chunk; ++q) {
            for (int k = 0; k < nsys; ++k) {
                if (done[k] || stop[k]) continue;
                iter_a(k);
                iter_b(k, 0);
                cur[k] ^= 1;
            }
        }
        CFD_HIP(hipGetLastError());
        CFD_CHECK(launched <= 4L * maxit + 64, "PCG did not terminate");
        chunk = std::min(chunk * 2, 16);
    }
}

double pcg_polled_iterations(const PcgWork& w) { return w.hpin.p ? w.hpin.p[S_K] : -1.0; }

void pcg_launch(PcgWork& w, int npoin, const double* A, const int* idx, const int* rowptr,
                const double* diag, double* x, const double* b, const double* x_fix,
                const int* x_fix_idx, int nfix, double tol, int maxit, int rel, hipStream_t s) {
    PcgSystem sys{&w, x, b, x_fix, x_fix_idx, nfix};
    pcg_launch_multi(&sys, 1, npoin, A, idx, rowptr, diag, tol, maxit, rel, s);
}

PcgResult pcg_result(const PcgWork& w, hipStream_t s) {
    double h[S_NSCAL];
    CFD_HIP(hipMemcpyAsync(h, w.scal.p, sizeof(h), hipMemcpyDeviceToHost, s));
    CFD_HIP(hipStreamSynchronize(s));
    PcgResult res;
    res.iters = (int)h[S_K];
    res.rz = h[S_EOLD];
    res.converged = h[S_CONV] != 0.0;
    return res;
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// The same solve with no allocation (cfd_bicg_ex_noalloc): every work vector
// lives in one file-scope device array, `scratch`, sized for kNoAllocRows
// rows; its address on each device comes from hipGetSymbolAddress once, with
// that device current, and only that address is handed to kernels and copies.
// Calls on one device share the array: overlapping calls there race.
// ---------------------------------------------------------------------------
namespace cfd {
namespace {

constexpr int kNoAllocRows = 1 << 19;  // 524,288 rows
constexpr int kNoAllocBlocks = (kNoAllocRows + kBlock - 1) / kBlock;
constexpr size_t kAlign = 32;  // doubles: 256-byte sub-array alignment
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
// layout (in doubles): r, z, y, pa, pb (kNoAllocRows each), partial
// (3 kNoAllocBlocks), scal (S_NSCAL), fixflag (kNoAllocRows int32), the CSR
// check flag (one int32, in its own slot)
constexpr size_t kOffVec = 0;
constexpr size_t kOffPartial = align_up(kOffVec + 5 * (size_t)kNoAllocRows);
constexpr size_t kOffScal = align_up(kOffPartial + 3 * (size_t)kNoAllocBlocks);
constexpr size_t kOffFix = align_up(kOffScal + S_NSCAL);
constexpr size_t kOffFlag = align_up(kOffFix + kNoAllocRows / 2);
constexpr size_t kScratchDoubles = align_up(kOffFlag + 1);

}  // namespace

// about 23 MB per device, held until the process ends
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

double* scratch_on_current_device() {
    int dev = 0;
    CFD_HIP(hipGetDevice(&dev));
    CFD_CHECK(dev >= 0 && dev < kMaxDevices, "device id beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        CFD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        CFD_CHECK(ptr != nullptr, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}

// The row pointer starts at 0 and never decreases, every column index and
// fixed-node index lies in 1..n: else *flag = 1 (flag cleared by the caller's
// first launch, block 0 thread 0 of k_csr_flag_clear).
__global__ void k_csr_flag_clear(int* __restrict__ flag) { *flag = 0; }

__global__ void k_csr_check(int n, const int* __restrict__ rowptr, const int* __restrict__ idx,
                            const int* __restrict__ fix_idx, int nfix, int* __restrict__ flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    if (t == 0 && rowptr[0] != 0) *flag = 1;
    for (int i = t; i < n; i += nt)
        if (rowptr[i + 1] < rowptr[i]) *flag = 1;
    const int nnz = rowptr[n];
    for (int q = t; q < nnz; q += nt) {
        const int c = idx[q];
        if (c < 1 || c > n) *flag = 1;
    }
    for (int q = t; q < nfix; q += nt) {
        const int c = fix_idx[q];
        if (c < 1 || c > n) *flag = 1;
    }
}

}  // namespace
}  // namespace cfd

extern "C" int cfd_bicg_ex_noalloc(int npoin, const double* spMtx, const int* spIdx,
                                   const int* spRowptr, const double* diagMtx, double* x,
                                   const double* b, const double* x_fix, const int* x_fixIdx,
                                   int nfix, double tol, int maxit, int rel, void* stream,
                                   int* iters, double* rz, int* converged) {
    using namespace cfd;
    CFD_TRY({
        CFD_ARG(npoin > 0 && spMtx && spIdx && spRowptr && diagMtx && x && b, "null argument");
        CFD_ARG(npoin <= kNoAllocRows, "npoin " + std::to_string(npoin) + " exceeds the " +
                                           std::to_string(kNoAllocRows) +
                                           " rows of the allocation-free scratch");
        CFD_ARG(tol > 0.0 && maxit >= 0 && (rel == 0 || rel == 1), "tol > 0, maxit >= 0, rel 0/1");
        CFD_ARG(nfix >= 0 && (nfix == 0 || (x_fix && x_fixIdx)), "fixed-node lists");
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* base = scratch_on_current_device();
        // the CSR and fixed-node indices are checked on the device first
        int* flag = reinterpret_cast<int*>(base + kOffFlag);
        k_csr_flag_clear<<<1, 1, 0, s>>>(flag);
        k_csr_check<<<1024, kBlock, 0, s>>>(npoin, spRowptr, spIdx, x_fixIdx, nfix, flag);
        CFD_HIP(hipGetLastError());
        int bad = 0;
        CFD_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        CFD_HIP(hipStreamSynchronize(s));
        CFD_ARG(bad == 0, "row pointer, column index or fixed-node index out of range");
        PcgWork w;  // views of the scratch array; nothing is allocated
        const int nb = (int)grid_for(npoin);
        w.n = npoin;
        w.nblocks = nb;
        w.r.borrow(base + kOffVec, npoin);
        w.z.borrow(base + kOffVec + (size_t)kNoAllocRows, npoin);
        w.y.borrow(base + kOffVec + 2 * (size_t)kNoAllocRows, npoin);
        w.pa.borrow(base + kOffVec + 3 * (size_t)kNoAllocRows, npoin);
        w.pb.borrow(base + kOffVec + 4 * (size_t)kNoAllocRows, npoin);
        w.partial.borrow(base + kOffPartial, 3 * (size_t)nb);
        w.scal.borrow(base + kOffScal, S_NSCAL);
        w.fixflag.borrow(reinterpret_cast<int*>(base + kOffFix), npoin);
        pcg_launch(w, npoin, spMtx, spIdx, spRowptr, diagMtx, x, b, x_fix, x_fixIdx, nfix, tol,
                   maxit, rel, s);
        PcgResult res = pcg_result(w, s);
        if (iters) *iters = res.iters;
        if (rz) *rz = res.rz;
        if (converged) *converged = res.converged;
    })
}
