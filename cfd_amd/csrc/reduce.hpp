// Deterministic grid reductions without atomics: a full-grid kernel reduces
// each block's values in a fixed tree and stores one partial per block; a
// one-block finalize kernel combines the partials, each thread folding a fixed
// strided subset in index order, then a fixed tree.  The result does not
// depend on scheduling, so repeated runs are bitwise identical.
#pragma once

#include <hip/hip_runtime.h>

namespace cfd {

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
struct OpMax {
    __device__ double operator()(double a, double b) const { return b > a ? b : a; }
};

// Block-wide reduction; the value is valid in thread 0.  Every thread of the
// block must call it.  Block size: a multiple of 64, at most 1024.
template <class Op>
__device__ inline double block_reduce(double v, Op op) {
    __shared__ double sh[16];
    v = wave_reduce(v, op);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();  // sh may still be read by a previous call
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = op(v, sh[w]);
    }
    return v;
}

// Full-grid pass: one partial per block (at index `at`: the block's logical
// index when the grid is remapped, else blockIdx.x).
template <class Op>
__device__ inline void block_partial(double v, Op op, double* partials, int at) {
    v = block_reduce(v, op);
    if (threadIdx.x == 0) partials[at] = v;
}
template <class Op>
__device__ inline void block_partial(double v, Op op, double* partials) {
    block_partial(v, op, partials, (int)blockIdx.x);
}

// One-block finalize: combine partials[0..n); valid in thread 0.
template <class Op>
__device__ inline double combine_partials(const double* partials, int n, double init, Op op) {
    double acc = init;
    // each thread folds partials threadIdx.x, + blockDim.x, ... in that order;
    // sixteen (then eight, four) at a time are loaded before they are folded,
    // so that large grids keep many loads in flight
    const int stride = (int)blockDim.x;
    int b = (int)threadIdx.x;
    for (; b + 15 * stride < n; b += 16 * stride) {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = partials[b + q * stride];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = op(acc, v[q]);
    }
    if (b + 7 * stride < n) {  // then eight
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = partials[b + q * stride];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc = op(acc, v[q]);
        b += 8 * stride;
    }
    if (b + 3 * stride < n) {  // then four
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = partials[b + q * stride];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = op(acc, v[q]);
        b += 4 * stride;
    }
    for (; b < n; b += stride) acc = op(acc, partials[b]);
    return block_reduce(acc, op);
}

// (A "last block combines" variant -- a device-scope fence and an atomic
// arrival count per block -- measured 2.6x slower per PCG iteration than the
// separate one-block finalize kernel on gfx950: each block's release fence
// writes back its L2.  The finalize kernels stay separate.)

namespace {
// *out = min(init, partials[0..n)), then min(*out, cap) (one block of 256)
__global__ void __launch_bounds__(256)
k_min_finalize(const double* __restrict__ partials, int n, double init, double cap,
               double* __restrict__ out) {
    double m = combine_partials(partials, n, init, OpMin());
    if (threadIdx.x == 0) *out = m > cap ? cap : m;
}
}  // namespace

}  // namespace cfd
