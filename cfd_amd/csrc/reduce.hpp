// Deterministic grid reductions: every block reduces its values in a fixed
// tree, stores one partial, and the last block to arrive (atomic ticket)
// combines the partials in index order.  The result does not depend on block
// scheduling, so repeated runs are bitwise identical.
#pragma once

#include <hip/hip_runtime.h>

namespace cfd {

// Reducing kernels run a fixed grid of at most kRedGrid blocks with a
// grid-stride loop: few partials (one ticket atomic per block on a single
// address is the expensive part), and a partition that depends on n only.
constexpr unsigned kRedGrid = 512;
inline unsigned red_grid(size_t n, unsigned block = 256) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return g < kRedGrid ? (unsigned)g : kRedGrid;
}

template <class Op>
__device__ inline double wave_reduce(double v, Op op) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
    return v;  // valid in lane 0
}

struct OpMin {
    __device__ double operator()(double a, double b) const { return b < a ? b : a; }
};
struct OpSum {
    __device__ double operator()(double a, double b) const { return a + b; }
};

// Reduces `v` over the block into partials[blockIdx.x]; returns true in every
// thread of the last block (after an acquire fence), so that block can read
// all partials.  Block size must be a multiple of 64 and at most 1024.
template <class Op>
__device__ inline bool block_partial(double v, Op op, double* partials, unsigned* counter) {
    __shared__ double sh[16];
    __shared__ int last;
    v = wave_reduce(v, op);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) acc = op(acc, sh[w]);
        partials[blockIdx.x] = acc;
        __threadfence();
        unsigned ticket = atomicAdd(counter, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// In the last block, called by every thread: combine partials[0..n).
// Thread t folds partials t, t+blockDim, ... in index order, then the block
// tree-reduces; the assignment and the tree are fixed, so the result is the
// same on every run.  The value is valid in thread 0.
template <class Op>
__device__ inline double combine_partials(const double* partials, int n, double init, Op op) {
    __shared__ double shc[16];
    double acc = init;
    for (int b = threadIdx.x; b < n; b += blockDim.x) acc = op(acc, partials[b]);
    acc = wave_reduce(acc, op);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) shc[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        acc = shc[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) acc = op(acc, shc[w]);
    }
    return acc;
}

}  // namespace cfd
