// A small MLP on the matrix cores (gfx950 MFMA, BF16 inputs, FP32
// accumulation): 16 -> 64 -> 64 -> 16 (tanh, tanh, linear), one row per mesh
// node.  It warm-starts the mesh-Laplace solves (see the session); this file
// holds the inference and training kernels.
//
// Layout on the matrix cores: nodes run along the N dimension, so a layer is
// Z(out x nodes) = W^T(out x in) . H(in x nodes) + b.  With
// v_mfma_f32_16x16x16bf16_1k (wave64) lane l holds
//   A[row = l%16][k = 4(l/16)+j],  B[k = 4(l/16)+j][col = l%16],
//   D[row = 4(l/16)+i][col = l%16]           (i, j = 0..3)
// so a layer's D tile (16 outputs x 16 nodes) is, element for element, the B
// operand of the next layer: activations never leave the registers.  Weight
// gradients need dZ and H with nodes along K; those tiles are transposed
// through LDS.
//
// Rounding: inputs, weights, activations fed to the next layer and the
// back-propagated dZ are rounded to BF16 (round to nearest even, done here
// bit-exactly); accumulations, the output, tanh and bias gradients are FP32.
#include "cfd_internal.hpp"

namespace cfd {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u)                  // inf / nan kept
        return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}
__device__ __forceinline__ bf16x4 pack4(const f32x4& v) { return pack4(v[0], v[1], v[2], v[3]); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand of W^T (rows = outputs, cols = inputs), tile (mt, kb), from the
// BF16 copy stored row-major [out][in] with `in` columns
__device__ __forceinline__ bf16x4 load_wt(const unsigned short* __restrict__ w, int in, int mt,
                                          int kb, int lane) {
    const unsigned short* p = w + (size_t)(mt * 16 + (lane & 15)) * in + kb * 16 + 4 * (lane >> 4);
    bf16x4 r;
    r[0] = (short)p[0];
    r[1] = (short)p[1];
    r[2] = (short)p[2];
    r[3] = (short)p[3];
    return r;
}
// A operand of W (rows = inputs, cols = outputs) from the same [out][in] copy:
// tile (kt, mb) = W[kt*16 + l%16][mb*16 + 4(l/16)+j] = Wt[mb*16 + 4(l/16)+j][kt*16 + l%16]
__device__ __forceinline__ bf16x4 load_w(const unsigned short* __restrict__ w, int in, int kt,
                                         int mb, int lane) {
    const int c = kt * 16 + (lane & 15), r0 = mb * 16 + 4 * (lane >> 4);
    bf16x4 r;
    r[0] = (short)w[(size_t)(r0 + 0) * in + c];
    r[1] = (short)w[(size_t)(r0 + 1) * in + c];
    r[2] = (short)w[(size_t)(r0 + 2) * in + c];
    r[3] = (short)w[(size_t)(r0 + 3) * in + c];
    return r;
}
__device__ __forceinline__ f32x4 bias4(const float* __restrict__ b, int mt, int lane) {
    const float* p = b + mt * 16 + 4 * (lane >> 4);
    f32x4 r;
    r[0] = p[0];
    r[1] = p[1];
    r[2] = p[2];
    r[3] = p[3];
    return r;
}
__device__ __forceinline__ f32x4 tanh4(f32x4 z) {
    f32x4 h;
    for (int i = 0; i < 4; ++i) h[i] = tanhf(z[i]);
    return h;
}

struct MlpDev {
    const unsigned short *w1, *w2, *w3;  // BF16 [out][in]
    const float *b1, *b2, *b3;           // FP32
};

// forward of one tile of 16 nodes; x4 = this lane's B operand of the input
__device__ __forceinline__ void mlp_tile_forward(const MlpDev& m, bf16x4 x4, int lane, f32x4 (&z1)[4],
                                                 f32x4 (&h1)[4], f32x4 (&z2)[4], f32x4 (&h2)[4],
                                                 f32x4& z3) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        z1[mt] = mfma(load_wt(m.w1, kMlpIn, mt, 0, lane), x4, bias4(m.b1, mt, lane));
        h1[mt] = tanh4(z1[mt]);
    }
    bf16x4 hb[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) hb[kb] = pack4(h1[kb]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = bias4(m.b2, mt, lane);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(load_wt(m.w2, kMlpHid, mt, kb, lane), hb[kb], acc);
        z2[mt] = acc;
        h2[mt] = tanh4(acc);
    }
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) hb[kb] = pack4(h2[kb]);
    f32x4 acc = bias4(m.b3, 0, lane);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(load_wt(m.w3, kMlpHid, 0, kb, lane), hb[kb], acc);
    z3 = acc;
}

// this lane's B operand of the input tile: features X[node][16] (BF16)
// (sample q = node0 + l%16 is node q * stride; samples q >= n are zero)
__device__ __forceinline__ bf16x4 load_x(const unsigned short* __restrict__ X, long long node0, int n,
                                         int lane, int stride = 1) {
    const long long node = node0 + (lane & 15);
    bf16x4 r = {0, 0, 0, 0};
    if (node < n) {
        const unsigned short* p = X + node * stride * kMlpIn + 4 * (lane >> 4);
        r[0] = (short)p[0];
        r[1] = (short)p[1];
        r[2] = (short)p[2];
        r[3] = (short)p[3];
    }
    return r;
}

// Inference: out[node][0..1] (FP32) for every node; one wave per 16 nodes
__global__ void __launch_bounds__(64) k_mlp_forward(MlpDev m, const unsigned short* __restrict__ X,
                                                     int n, float* __restrict__ out) {
    const int lane = threadIdx.x;
    const long long node0 = (long long)blockIdx.x * 16;
    f32x4 z1[4], h1[4], z2[4], h2[4], z3;
    mlp_tile_forward(m, load_x(X, node0, n, lane), lane, z1, h1, z2, h2, z3);
    const long long node = node0 + (lane & 15);
    if (lane < 16 && node < n) {  // rows 0 and 1 of the output tile
        out[2 * node] = z3[0];
        out[2 * node + 1] = z3[1];
    }
}

// 16x16 tile transpose through LDS: in D layout (lane l holds [4(l/16)+i][l%16]),
// out in A layout ([l%16][4(l/16)+j]), rounded to BF16
__device__ __forceinline__ bf16x4 transpose_tile(float* sm, const f32x4& d, int lane) {
#pragma unroll
    for (int i = 0; i < 4; ++i) sm[(4 * (lane >> 4) + i) * 17 + (lane & 15)] = d[i];
    __syncthreads();
    bf16x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = (short)f2bf(sm[(lane & 15) * 17 + 4 * (lane >> 4) + j]);
    __syncthreads();
    return r;
}

// Training: loss sum((y - t)^2)/n over outputs 0..1 of the n sampled nodes
// q * stride (q < n); one wave per block, `tiles` tiles of 16 samples per block.
// Writes this block's gradient partial (kMlpParams floats, parameter layout
// of the master copy) and its loss partial.
__global__ void __launch_bounds__(64) k_mlp_grad(MlpDev m, const unsigned short* __restrict__ X,
                                                  const float* __restrict__ T, int n, int stride,
                                                  int tiles, float inv_n, float* __restrict__ gpart,
                                                  float* __restrict__ lpart) {
    __shared__ float sm[16 * 17];
    const int lane = threadIdx.x;
    const int r4 = 4 * (lane >> 4), c = lane & 15;
    // weight-gradient accumulators, D layout: dW1t 4 tiles (64x16), dW2t 16
    // tiles (64x64: [mt][kb]), dW3t 4 tiles (16x64: [kb]); bias sums per row
    f32x4 g1[4], g2[16], g3[4];
    f32x4 gb1[4], gb2[4], gb3;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) { g1[t] = zero; g3[t] = zero; gb1[t] = zero; gb2[t] = zero; }
#pragma unroll
    for (int t = 0; t < 16; ++t) g2[t] = zero;
    gb3 = zero;
    float loss = 0.f;
    for (int tt = 0; tt < tiles; ++tt) {
        const long long node0 = ((long long)blockIdx.x * tiles + tt) * 16;
        if (node0 >= n) break;  // uniform over the wave
        const bf16x4 x4 = load_x(X, node0, n, lane, stride);
        f32x4 z1[4], h1[4], z2[4], h2[4], z3;
        mlp_tile_forward(m, x4, lane, z1, h1, z2, h2, z3);
        // dZ3 (D layout): rows 0, 1 carry the loss; padded rows and samples 0
        const long long q = node0 + c, node = q * stride;
        f32x4 dz3 = zero;
        if (q < n && r4 == 0) {
            const float d0 = z3[0] - T[2 * node], d1 = z3[1] - T[2 * node + 1];
            loss += d0 * d0 + d1 * d1;
            dz3[0] = 2.f * d0 * inv_n;
            dz3[1] = 2.f * d1 * inv_n;
        }
        // dH2 = W3 dZ3 (64 x 16 nodes), dZ2 = dH2 (1 - h2^2)
        const bf16x4 dz3b = pack4(dz3);
        f32x4 dz2[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            const f32x4 dh = mfma(load_w(m.w3, kMlpHid, kt, 0, lane), dz3b, zero);
#pragma unroll
            for (int i = 0; i < 4; ++i) dz2[kt][i] = dh[i] * (1.f - h2[kt][i] * h2[kt][i]);
        }
        // dH1 = W2 dZ2, dZ1 = dH1 (1 - h1^2)
        bf16x4 dz2b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) dz2b[t] = pack4(dz2[t]);
        f32x4 dz1[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            f32x4 dh = zero;
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) dh = mfma(load_w(m.w2, kMlpHid, kt, mb, lane), dz2b[mb], dh);
#pragma unroll
            for (int i = 0; i < 4; ++i) dz1[kt][i] = dh[i] * (1.f - h1[kt][i] * h1[kt][i]);
        }
        // bias gradients: row sums over the 16 nodes (reduced at the end)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gb1[t] += dz1[t];
            gb2[t] += dz2[t];
        }
        gb3 += dz3;
        // weight gradients dWt(out x in) += dZ(out x nodes) H^T(nodes x in):
        // A = dZ with nodes along K, B = H^T; both transposed through LDS
        const bf16x4 a3 = transpose_tile(sm, dz3, lane);
        bf16x4 a2[4], a1[4], bh2[4], bh1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            a2[t] = transpose_tile(sm, dz2[t], lane);
            a1[t] = transpose_tile(sm, dz1[t], lane);
            bh2[t] = transpose_tile(sm, h2[t], lane);  // B[node][in] = H[in][node]
            bh1[t] = transpose_tile(sm, h1[t], lane);
        }
        // input tile with nodes along K: B[node = 4(l/16)+j][in = l%16]
        bf16x4 bx;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long nd = node0 + r4 + j;
            bx[j] = nd < n ? (short)X[nd * stride * kMlpIn + c] : (short)0;
        }
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) g3[kb] = mfma(a3, bh2[kb], g3[kb]);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) g2[mt * 4 + kb] = mfma(a2[mt], bh1[kb], g2[mt * 4 + kb]);
            g1[mt] = mfma(a1[mt], bx, g1[mt]);
        }
    }
    // store this block's partial in the master layout
    float* gp = gpart + (size_t)blockIdx.x * kMlpParams;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            gp[kMlpW1 + (mt * 16 + r4 + i) * kMlpIn + c] = g1[mt][i];
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
                gp[kMlpW2 + (mt * 16 + r4 + i) * kMlpHid + kb * 16 + c] = g2[mt * 4 + kb][i];
        }
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int i = 0; i < 4; ++i) gp[kMlpW3 + (r4 + i) * kMlpHid + kb * 16 + c] = g3[kb][i];
    // bias: sum over the 16 node columns (lanes with the same l/16), fixed order
    auto rowsum = [&](const f32x4& v, float* dst) {
#pragma unroll
        for (int i = 0; i < 4; ++i) sm[(r4 + i) * 17 + c] = v[i];
        __syncthreads();
        if (lane < 16) {
            float s = 0.f;
            for (int q = 0; q < 16; ++q) s += sm[lane * 17 + q];
            dst[lane] = s;
        }
        __syncthreads();
    };
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        rowsum(gb1[t], gp + kMlpB1 + t * 16);
        rowsum(gb2[t], gp + kMlpB2 + t * 16);
    }
    rowsum(gb3, gp + kMlpB3);
    // loss partial
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (lane == 0) lpart[blockIdx.x] = loss;
}

// grad[p] = sum over blocks of gpart[b][p]; the loss likewise from lpart (as
// column kMlpParams).  A block sums 64 columns; each of its kRedWaves waves
// sums one contiguous range of the partial blocks, and the waves' sums are
// added in wave order: a fixed order for a given block count.
constexpr int kRedWaves = 16;
__global__ void __launch_bounds__(64 * kRedWaves)
k_mlp_reduce(const float* __restrict__ gpart, const float* __restrict__ lpart, int nblocks,
             float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float sm[kRedWaves][65];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const int b0 = (int)((long long)nblocks * wv / kRedWaves);
    const int b1 = (int)((long long)nblocks * (wv + 1) / kRedWaves);
    float s = 0.f;
    if (p < kMlpParams) {
#pragma unroll 8
        for (int b = b0; b < b1; ++b) s += gpart[(size_t)b * kMlpParams + p];
    } else if (p == kMlpParams) {
#pragma unroll 8
        for (int b = b0; b < b1; ++b) s += lpart[b];
    }
    sm[wv][lane] = s;
    __syncthreads();
    if (wv == 0) {
        float t = 0.f;
#pragma unroll
        for (int w = 0; w < kRedWaves; ++w) t += sm[w][lane];
        if (p < kMlpParams) grad[p] = t;
        else if (p == kMlpParams) *loss = t;
    }
}

// Adam on the FP32 master copy; rows of W3 and b3 beyond the used outputs
// stay zero.  Then the BF16 copies of the weights.
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ mom, float* __restrict__ vel,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps,
                           float c1, float c2, int nout, unsigned short* __restrict__ wb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const bool used = !((p >= kMlpW3 && p < kMlpB3 && (p - kMlpW3) / kMlpHid >= nout) ||
                        (p >= kMlpB3 && p - kMlpB3 >= nout));
    if (used) {
        const float g = grad[p];
        const float mm = b1 * mom[p] + (1.f - b1) * g;
        const float vv = b2 * vel[p] + (1.f - b2) * g * g;
        mom[p] = mm;
        vel[p] = vv;
        w[p] = w[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    }
    if (p < kMlpB3 && !(p >= kMlpB1 && p < kMlpW2) && !(p >= kMlpB2 && p < kMlpW3))
        wb[p] = f2bf(w[p]);
}

__global__ void k_mlp_to_bf16(const float* __restrict__ w, unsigned short* __restrict__ wb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) wb[p] = f2bf(w[p]);
}

MlpDev mlp_dev(const MlpWeights& w) {
    MlpDev m;
    m.w1 = w.wb.p + kMlpW1;
    m.w2 = w.wb.p + kMlpW2;
    m.w3 = w.wb.p + kMlpW3;
    m.b1 = w.w.p + kMlpB1;
    m.b2 = w.w.p + kMlpB2;
    m.b3 = w.w.p + kMlpB3;
    return m;
}

}  // namespace

void MlpWeights::init(const float* host_params, hipStream_t s) {
    w.alloc(kMlpParams);
    m.alloc(kMlpParams);
    v.alloc(kMlpParams);
    wb.alloc(kMlpParams);
    grad.alloc(kMlpParams);
    loss.alloc(1);
    w.upload(host_params, kMlpParams, s);
    m.zero(s);
    v.zero(s);
    step = 0;
    k_mlp_to_bf16<<<grid_for(kMlpParams), kBlock, 0, s>>>(w.p, wb.p);
    CFD_HIP(hipGetLastError());
}

void mlp_forward(const MlpWeights& w, const unsigned short* X, int n, float* out, hipStream_t s) {
    if (n <= 0) return;
    k_mlp_forward<<<(n + 15) / 16, 64, 0, s>>>(mlp_dev(w), X, n, out);
    CFD_HIP(hipGetLastError());
}

int mlp_gradient(MlpWeights& w, const unsigned short* X, const float* T, int n, int tiles_per_block,
                 hipStream_t s, int stride) {
    CFD_CHECK(n > 0 && tiles_per_block > 0 && stride > 0, "arg: MLP training needs nodes");
    const int ns = (n + stride - 1) / stride;  // samples 0, stride, 2 stride, ... < n
    const int nb = (ns + 16 * tiles_per_block - 1) / (16 * tiles_per_block);
    if ((size_t)nb * kMlpParams > w.gpart.n) w.gpart.alloc((size_t)nb * kMlpParams);
    if ((size_t)nb > w.lpart.n) w.lpart.alloc(nb);
    k_mlp_grad<<<nb, 64, 0, s>>>(mlp_dev(w), X, T, ns, stride, tiles_per_block, 1.0f / (float)ns,
                                 w.gpart.p, w.lpart.p);
    k_mlp_reduce<<<kMlpParams / 64 + 1, 64 * kRedWaves, 0, s>>>(w.gpart.p, w.lpart.p, nb,
                                                                  w.grad.p, w.loss.p);
    CFD_HIP(hipGetLastError());
    return ns;
}

void mlp_adam(MlpWeights& w, float lr, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++w.step;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)w.step));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)w.step));
    k_mlp_adam<<<grid_for(kMlpParams), kBlock, 0, s>>>(w.w.p, w.m.p, w.v.p, w.grad.p, lr, b1, b2, eps,
                                                       c1, c2, kMlpOutUsed, w.wb.p);
    CFD_HIP(hipGetLastError());
}

}  // namespace cfd
