// mlp.hip - forward and backward of the warm-start network on the gfx950 matrix
// cores.  One wavefront handles 16 samples at a time with v_mfma_f32_16x16x16_bf16
// (BF16 inputs, FP32 accumulation).  Layers are computed transposed, Z^T = W X^T, so
// the accumulator layout of one layer (lane l: sample l%16, features 4(l/16)..+3 of
// each 16-wide tile) is exactly the B-operand layout of the next: activations never
// leave registers in the forward pass.  The weight gradients reduce over samples,
// which needs samples in the K role; those operands are transposed through LDS.
//
// Rounding points (the CPU check in oracle/mlp_oracle.py applies the same ones):
// inputs, weights, H1 and the dZ fed to the matrix cores are BF16; biases,
// activations, the output layer and every accumulation are FP32.
#include "mlp.h"

namespace pcisph {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ unsigned short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = sigmoid(z);
    return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct Weights {
    unsigned short w1[kMlpHidden][16];  // K padded 8 -> 16 with zeros
    unsigned short w2[kMlpHidden][kMlpHidden];
    float b1[kMlpHidden], b2[kMlpHidden], w3[kMlpHidden];
    float b3;
};

__device__ void stage_weights(Weights& s, const float* __restrict__ P) {
    for (int i = threadIdx.x; i < kMlpHidden * 16; i += blockDim.x) {
        const int n = i / 16, k = i % 16;
        s.w1[n][k] = k < kMlpIn ? f2bf(P[kMlpW1 + n * kMlpIn + k]) : (unsigned short)0;
    }
    for (int i = threadIdx.x; i < kMlpHidden * kMlpHidden; i += blockDim.x)
        s.w2[i / kMlpHidden][i % kMlpHidden] = f2bf(P[kMlpW2 + i]);
    for (int i = threadIdx.x; i < kMlpHidden; i += blockDim.x) {
        s.b1[i] = P[kMlpB1 + i];
        s.b2[i] = P[kMlpB2 + i];
        s.w3[i] = bf2f(f2bf(P[kMlpW3 + i]));
    }
    if (threadIdx.x == 0) s.b3 = P[kMlpB3];
}

__device__ __forceinline__ v4s load4(const unsigned short* p) {
    v4s v;
    v[0] = (short)p[0];
    v[1] = (short)p[1];
    v[2] = (short)p[2];
    v[3] = (short)p[3];
    return v;
}

// Forward of the 16 rows starting at row0.  Lane l: r = l%16, g = l/16.
// z1[t][j] = Z1[r][16t+4g+j], z2[u][j] = Z2[r][16u+4g+j] (pre-activation, bias
// included); returns y[r] in every lane.  xb = bf16 inputs X[r][4g..4g+3].
__device__ __forceinline__ float tile_forward(const Weights& s, v4s xb, v4f z1[4], v4f z2[4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    (void)r;
    v4s h1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const v4s a = load4(&s.w1[16 * t + (lane & 15)][4 * g]);
        z1[t] = mfma(a, xb, v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float z = z1[t][j] + s.b1[16 * t + 4 * g + j];
            z1[t][j] = z;
            h1[t][j] = (short)f2bf(silu(z));
        }
    }
    float part = 0.f;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; ++t) acc = mfma(load4(&s.w2[16 * u + (lane & 15)][16 * t + 4 * g]), h1[t], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float z = acc[j] + s.b2[16 * u + 4 * g + j];
            acc[j] = z;
            part += s.w3[16 * u + 4 * g + j] * silu(z);
        }
        z2[u] = acc;
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + s.b3;
}

__device__ __forceinline__ v4s load_x(const float* __restrict__ x, uint32_t row, uint32_t n, int g) {
    v4s xb = {0, 0, 0, 0};
    if (g < 2 && row < n) {
        const float4 v = *reinterpret_cast<const float4*>(x + (size_t)row * kMlpIn + 4 * g);
        xb[0] = (short)f2bf(v.x);
        xb[1] = (short)f2bf(v.y);
        xb[2] = (short)f2bf(v.z);
        xb[3] = (short)f2bf(v.w);
    }
    return xb;
}

__global__ __launch_bounds__(kThreads) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ x,
                                                          float* __restrict__ y, uint32_t n) {
    __shared__ Weights s;
    stage_weights(s, P);
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * kWaves;
    for (uint32_t row0 = wave * 16; row0 < n; row0 += waves * 16) {
        const uint32_t row = row0 + (lane & 15);
        v4f z1[4], z2[4];
        const float out = tile_forward(s, load_x(x, row, n, g), z1, z2);
        if (g == 0 && row < n) y[row] = out;
    }
}

// Per-wave LDS scratch for the transposes of the weight-gradient products.
struct WaveScratch {
    unsigned short h1[16][kMlpHidden];   // H1 (bf16), row = sample
    unsigned short dz2[16][kMlpHidden];  // dZ2 (bf16)
    unsigned short dz1[16][kMlpHidden];  // dZ1 (bf16)
    unsigned short x[16][16];            // inputs (bf16), K padded
};

__global__ __launch_bounds__(kThreads) void k_mlp_backward(const float* __restrict__ P, const float* __restrict__ x,
                                                           const float* __restrict__ t, uint32_t n,
                                                           float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ Weights s;
    __shared__ WaveScratch ws[kWaves];
    __shared__ float sgrad[kMlpParams];
    stage_weights(s, P);
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sgrad[i] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4, w = threadIdx.x >> 6;
    WaveScratch& sc = ws[w];
    const uint32_t wave = blockIdx.x * kWaves + w;
    const uint32_t waves = gridDim.x * kWaves;
    const float inv_n = 1.f / (float)n;

    v4f gw2[4][4], gw1[4];
    float gw3[4][4], gb2[4][4], gb1[4][4], gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw1[a] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gw2[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
            gw3[a][b] = gb2[a][b] = gb1[a][b] = 0.f;
        }
    }

    for (uint32_t row0 = wave * 16; row0 < n; row0 += waves * 16) {
        const uint32_t row = row0 + r;
        const bool valid = row < n;
        const v4s xb = load_x(x, row, n, g);
        v4f z1[4], z2[4];
        const float y = tile_forward(s, xb, z1, z2);
        const float dy = valid ? (y - t[row]) * inv_n : 0.f;
        if (g == 0 && valid) {
            lsum += 0.5f * (y - t[row]) * (y - t[row]);
            gb3 += dy;
        }
        // output layer and dZ2
        v4s dz2b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int nn = 16 * u + 4 * g + j;
                const float z = z2[u][j];
                gw3[u][j] += dy * silu(z);
                const float d = dy * s.w3[nn] * silu_grad(z);
                gb2[u][j] += d;
                const unsigned short db = f2bf(d);
                dz2b[u][j] = (short)db;
                sc.dz2[r][nn] = db;
                sc.h1[r][nn] = f2bf(silu(z1[u][j]));
            }
        // dH1^T = W2^T dZ2^T, then dZ1
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v4s a;
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = (short)s.w2[16 * u + 4 * g + j][16 * tt + r];
                acc = mfma(a, dz2b[u], acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * tt + 4 * g + j;
                const float d = acc[j] * silu_grad(z1[tt][j]);
                gb1[tt][j] += d;
                sc.dz1[r][k] = f2bf(d);
            }
        }
        if (g < 4) {
#pragma unroll
            for (int j = 0; j < 4; ++j) sc.x[r][4 * g + j] = (unsigned short)xb[j];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // dW2[n2][k] += sum_r dZ2[r][n2] H1[r][k]  (M = n2, N = k, K = r)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v4s a;
#pragma unroll
            for (int j = 0; j < 4; ++j) a[j] = (short)sc.dz2[4 * g + j][16 * u + r];
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) {
                v4s b;
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = (short)sc.h1[4 * g + j][16 * tt + r];
                gw2[u][tt] = mfma(a, b, gw2[u][tt]);
            }
        }
        // dW1[n][f] += sum_r dZ1[r][n] X[r][f]  (M = n, N = f, K = r)
        v4s bx;
#pragma unroll
        for (int j = 0; j < 4; ++j) bx[j] = (short)sc.x[4 * g + j][r];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            v4s a;
#pragma unroll
            for (int j = 0; j < 4; ++j) a[j] = (short)sc.dz1[4 * g + j][16 * v + r];
            gw1[v] = mfma(a, bx, gw1[v]);
        }
        __builtin_amdgcn_wave_barrier();
    }

    // reduce the per-lane partial sums of this wave into the block's LDS copy
    // dW2: lane holds dW2[16u + 4g + j][16tt + r]
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                atomicAdd(&sgrad[kMlpW2 + (16 * u + 4 * g + j) * kMlpHidden + 16 * tt + r], gw2[u][tt][j]);
    // dW1: lane holds dW1[16v + 4g + j][f = r], f < 8 real
    if (r < kMlpIn) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int j = 0; j < 4; ++j) atomicAdd(&sgrad[kMlpW1 + (16 * v + 4 * g + j) * kMlpIn + r], gw1[v][j]);
    }
    // vectors indexed by 16u + 4g + j, summed over r (the 16 lanes of a group)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float a = gw3[u][j], b = gb2[u][j], c = gb1[u][j];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                a += __shfl_xor(a, off);
                b += __shfl_xor(b, off);
                c += __shfl_xor(c, off);
            }
            if (r == 0) {
                atomicAdd(&sgrad[kMlpW3 + 16 * u + 4 * g + j], a);
                atomicAdd(&sgrad[kMlpB2 + 16 * u + 4 * g + j], b);
                atomicAdd(&sgrad[kMlpB1 + 16 * u + 4 * g + j], c);
            }
        }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off);
        lsum += __shfl_xor(lsum, off);
    }
    if (lane == 0) {
        atomicAdd(&sgrad[kMlpB3], gb3);
        atomicAdd(loss, lsum);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x)
        if (sgrad[i] != 0.f) atomicAdd(&grad[i], sgrad[i]);
}

__global__ void k_zero(float* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Inference over the row ranges of `rows` (layout in mlp.h): each wavefront takes a
// contiguous run of 16-row tiles, finds the range of its first tile by bisection and
// walks on from there; rows past a range's end are zero inputs and not stored.
// Workgroups with no tile return before staging the weights.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_mlp_warm(const float* __restrict__ P, const float* __restrict__ x,
                                                       const uint32_t* __restrict__ rows, R* __restrict__ y) {
    __shared__ Weights s;
    const uint32_t count = rows[0], tiles = rows[1];
    const uint32_t waves = gridDim.x * kWaves;
    const uint32_t per = (tiles + waves - 1) / waves;
    if ((uint64_t)blockIdx.x * kWaves * per >= tiles) return;  // whole workgroup idle
    stage_weights(s, P);
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
    uint32_t q = wave * per;
    const uint32_t qe = q + per < tiles ? q + per : tiles;
    if (q >= qe) return;
    const uint32_t* r = rows + 2;  // {begin, end, first tile} per range, first tiles increasing
    uint32_t lo = 0, hi = count;   // r[3 lo + 2] <= q < r[3 hi + 2] (hi == count: past the last)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (r[3 * mid + 2] <= q) lo = mid;
        else hi = mid;
    }
    uint32_t k = lo, b = r[3 * k], e = r[3 * k + 1], t0 = r[3 * k + 2];
    for (; q < qe; ++q) {
        while (q >= t0 + (e - b + 15) / 16) {
            ++k;
            b = r[3 * k];
            e = r[3 * k + 1];
            t0 = r[3 * k + 2];
        }
        const uint32_t row = b + 16 * (q - t0) + (lane & 15);
        v4f z1[4], z2[4];
        const float out = tile_forward(s, load_x(x, row, e, g), z1, z2);
        if (g == 0 && row < e) y[row] = (R)out;
    }
}

template <typename R>
__global__ void k_collect(const float* __restrict__ feat, const R* __restrict__ p, uint32_t n, uint32_t stride,
                          uint32_t phase, float p_ref, float* __restrict__ sx, float* __restrict__ st, uint64_t base,
                          uint64_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i % stride != phase) return;
    const uint64_t slot = base + i / stride;
    if (slot >= cap) return;
    const float4* src = reinterpret_cast<const float4*>(feat + (size_t)i * kMlpIn);
    float4* dst = reinterpret_cast<float4*>(sx + slot * kMlpIn);
    dst[0] = src[0];
    dst[1] = src[1];
    st[slot] = (float)p[i] / p_ref - feat[(size_t)i * kMlpIn + 2];  // change from the previous pressure
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ void k_gather_batch(const float* __restrict__ sx, const float* __restrict__ stt, uint64_t count,
                               uint64_t seed, uint32_t step, uint32_t batch, float* __restrict__ x,
                               float* __restrict__ t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= batch) return;
    const uint64_t j = mix64(seed ^ mix64(((uint64_t)step << 32) | i)) % count;
    const float4* src = reinterpret_cast<const float4*>(sx + j * kMlpIn);
    float4* dst = reinterpret_cast<float4*>(x + (size_t)i * kMlpIn);
    dst[0] = src[0];
    dst[1] = src[1];
    t[i] = stt[j];
}

uint32_t blocks_for(uint32_t n) {
    const uint32_t tiles = (n + 15) / 16;
    uint32_t b = (tiles + kWaves - 1) / kWaves;
    return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

}  // namespace

void mlp_forward(const MlpLaunch& L, const float* params, const float* x, float* y, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(n)), dim3(kThreads), 0, L.stream, params, x, y, n);
}

void mlp_gradients(const MlpLaunch& L, const float* params, const float* x, const float* t, uint32_t n,
                   float* grad, float* loss) {
    hipLaunchKernelGGL(k_zero, dim3((kMlpParams + 255) / 256), dim3(256), 0, L.stream, grad, (uint32_t)kMlpParams);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, L.stream, loss, 1u);
    if (!n) return;
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks_for(n)), dim3(kThreads), 0, L.stream, params, x, t, n, grad,
                       loss);
}

template <typename R>
void mlp_warm_start(const MlpLaunch& L, const float* params, const float* feat, const uint32_t* rows,
                    uint32_t max_rows, R* y) {
    if (!max_rows) return;
    hipLaunchKernelGGL((k_mlp_warm<R>), dim3(blocks_for(max_rows)), dim3(kThreads), 0, L.stream, params, feat, rows,
                       y);
}
template void mlp_warm_start<float>(const MlpLaunch&, const float*, const float*, const uint32_t*, uint32_t, float*);
template void mlp_warm_start<double>(const MlpLaunch&, const float*, const float*, const uint32_t*, uint32_t,
                                     double*);

template <typename R>
void mlp_collect(const MlpLaunch& L, const float* feat, const R* p, uint32_t n, uint32_t stride, uint32_t phase,
                 float p_ref, float* store_x, float* store_t, uint64_t base, uint64_t cap) {
    if (!n || !stride) return;
    hipLaunchKernelGGL((k_collect<R>), dim3((n + 255) / 256), dim3(256), 0, L.stream, feat, p, n, stride, phase,
                       p_ref, store_x, store_t, base, cap);
}
template void mlp_collect<float>(const MlpLaunch&, const float*, const float*, uint32_t, uint32_t, uint32_t, float,
                                 float*, float*, uint64_t, uint64_t);
template void mlp_collect<double>(const MlpLaunch&, const float*, const double*, uint32_t, uint32_t, uint32_t,
                                  float, float*, float*, uint64_t, uint64_t);

void mlp_gather_batch(const MlpLaunch& L, const float* store_x, const float* store_t, uint64_t count, uint64_t seed,
                      uint32_t step, uint32_t batch, float* x, float* t) {
    if (!batch || !count) return;
    hipLaunchKernelGGL(k_gather_batch, dim3((batch + 255) / 256), dim3(256), 0, L.stream, store_x, store_t, count,
                       seed, step, batch, x, t);
}

void mlp_adam(const MlpLaunch& L, float* params, const float* grad, float* m, float* v, uint32_t step, float lr,
              float beta1, float beta2, float eps) {
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, L.stream, params, grad, m, v,
                       (uint32_t)kMlpParams, lr, beta1, beta2, eps, c1, c2);
}

}  // namespace pcisph
