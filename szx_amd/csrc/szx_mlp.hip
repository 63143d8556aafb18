// szx_mlp.hip -- the fixed-rate rate model: a 24 -> 64 -> 64 -> 1 MLP with
// GELU activations, trained and evaluated with BF16 matrix-core (MFMA)
// kernels accumulating in FP32.
//
// Parameters live in one FP32 array (the master copy Adam updates):
//   W1 [64][24] | b1 [64] | W2 [64][64] | b2 [64] | w3 [64] | b3      (5825)
// A workgroup rounds W1 (zero-padded to [64][32]) and W2 to BF16 into LDS
// once; the output layer (64 -> 1) and all biases stay FP32.
//
// Each wave owns 16 samples at a time and runs the two hidden layers as
// v_mfma_f32_16x16x16_bf16 tiles (D = A * B with A 16xK, B Kx16, K = 16):
//   A fragment: lane l holds A[l % 16][4 (l / 16) + t], t = 0..3
//   B fragment: lane l holds B[4 (l / 16) + t][l % 16]
//   D tile:     lane l, register r holds D[4 (l / 16) + r][l % 16]
// Activations move between the D layout and the next A layout through LDS.
//
// BF16 rounding points (the CPU oracle of the tests applies the same ones):
// inputs X, weights W1 and W2, the layer-1 activations G1, and in the
// backward pass dZ2 and dZ1.  Pre-activations Z1, Z2, activations G2, the
// output layer and every accumulation are FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szx_mlp.hpp"

namespace szxamd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int IN_PAD = 32;
constexpr int H = MLP_HIDDEN;

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
__device__ inline short to_bf16(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (short)((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline float from_bf16(short b) { return __builtin_bit_cast(float, (uint32_t)(uint16_t)b << 16); }

// GELU, tanh form, and its derivative.
constexpr float GELU_C = 0.7978845608028654f;   // sqrt(2 / pi)
constexpr float GELU_A = 0.044715f;
__device__ inline float gelu(float z)
{
    const float t = tanhf(GELU_C * (z + GELU_A * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ inline float gelu_grad(float z)
{
    const float t = tanhf(GELU_C * (z + GELU_A * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * GELU_C * (1.0f + 3.0f * GELU_A * z * z);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct MlpShared {
    short w1[H][IN_PAD];     // BF16 W1, zero-padded inputs
    short w2[H][H];          // BF16 W2 [out][in]
    short x[4][16][IN_PAD];  // per wave: BF16 inputs
    short g1[4][16][H];      // per wave: BF16 layer-1 activations
    short dz[4][16][H];      // per wave: BF16 dZ2, then dZ1
};

__device__ inline void load_weights(const float *P, MlpShared &s)
{
    for (int e = threadIdx.x; e < H * IN_PAD; e += blockDim.x) {
        const int j = e / IN_PAD, k = e % IN_PAD;
        s.w1[j][k] = k < MLP_IN ? to_bf16(P[MLP_W1 + j * MLP_IN + k]) : (short)0;
    }
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) s.w2[e / H][e % H] = to_bf16(P[MLP_W2 + e]);
}

// Forward pass of one 16-sample tile for the wave; rows past `rows` are zero.
// Leaves BF16 X and G1 in LDS, returns Z1, Z2 (D layout, [n-tile][r]) and y
// for rows 4 (l / 16) + r in yr[r].
__device__ inline void forward_tile(const float *P, const float *X, uint64_t row0, int rows,
                                    MlpShared &s, int wv, int l, f32x4 (&z1)[4], f32x4 (&z2)[4],
                                    float (&yr)[4])
{
    // inputs -> LDS (BF16, padded)
    for (int e = l; e < 16 * IN_PAD; e += 64) {
        const int i = e / IN_PAD, k = e % IN_PAD;
        s.x[wv][i][k] = (i < rows && k < MLP_IN) ? to_bf16(X[(row0 + i) * MLP_IN + k]) : (short)0;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const int lr = l % 16, lg = l / 16;
    // layer 1: Z1 = X W1^T + b1   (K = 32: two steps)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                a[t] = s.x[wv][lr][16 * ks + 4 * lg + t];
                b[t] = s.w1[16 * nt + lr][16 * ks + 4 * lg + t];
            }
            acc = mfma(a, b, acc);
        }
        const float bias = P[MLP_B1 + 16 * nt + lr];
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] += bias;
        z1[nt] = acc;
    }
    // G1 -> LDS as BF16 (row 4 lg + r, column 16 nt + lr)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int r = 0; r < 4; r++) s.g1[wv][4 * lg + r][16 * nt + lr] = to_bf16(gelu(z1[nt][r]));
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 2: Z2 = G1 W2^T + b2   (K = 64: four steps)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                a[t] = s.g1[wv][lr][16 * ks + 4 * lg + t];
                b[t] = s.w2[16 * nt + lr][16 * ks + 4 * lg + t];
            }
            acc = mfma(a, b, acc);
        }
        const float bias = P[MLP_B2 + 16 * nt + lr];
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] += bias;
        z2[nt] = acc;
    }
    // output layer (FP32): y_i = sum_j gelu(Z2[i][j]) w3[j] + b3
    float part[4] = {0, 0, 0, 0};
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const float w = P[MLP_W3 + 16 * nt + lr];
#pragma unroll
        for (int r = 0; r < 4; r++) part[r] += gelu(z2[nt][r]) * w;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) part[r] += __shfl_xor(part[r], m, 64);
        yr[r] = part[r] + P[MLP_B3];
    }
}

// y[i] for n samples of X ([n][24]).
__global__ __launch_bounds__(256) void k_mlp_forward(const float *P, const float *X, uint64_t n, float *y)
{
    __shared__ MlpShared s;
    load_weights(P, s);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint64_t ntiles = (n + 15) / 16;
    for (uint64_t tile = blockIdx.x * 4ull + wv; tile < ntiles; tile += gridDim.x * 4ull) {
        const uint64_t row0 = tile * 16;
        const int rows = (int)(n - row0 < 16 ? n - row0 : 16);
        f32x4 z1[4], z2[4];
        float yr[4];
        forward_tile(P, X, row0, rows, s, wv, l, z1, z2, yr);
        if (l % 16 == 0)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 4 * (l / 16) + r;
                if (i < rows) y[row0 + i] = yr[r];
            }
    }
}

// Gradient of the mean squared error sum_i (y_i - t_i)^2 / n over n samples:
// per workgroup partial gradients part[wg][MLP_PARAMS] and partial loss
// loss_part[wg] (sum of squared errors / n).
__global__ __launch_bounds__(256) void k_mlp_grad(const float *P, const float *X, const float *T,
                                                  uint64_t n, float *part, float *loss_part)
{
    __shared__ MlpShared s;
    __shared__ float s_grad[MLP_PARAMS];
    __shared__ float s_loss[4];
    load_weights(P, s);
    for (int e = threadIdx.x; e < MLP_PARAMS; e += blockDim.x) s_grad[e] = 0.0f;
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lr = l % 16, lg = l / 16;
    const float inv_n = 1.0f / (float)n;

    // per-lane gradient accumulators
    f32x4 gw2[4][4];   // [mt][nt]: dW2 rows 16 mt + 4 lg + r, cols 16 nt + lr
    f32x4 gw1[4][2];   // [mt][nt]: dW1 rows 16 mt + 4 lg + r, cols 16 nt + lr
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int b = 0; b < 4; b++) gw2[a][b] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 2; b++) gw1[a][b] = f32x4{0, 0, 0, 0};
    }
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0};
    float gb3 = 0.0f, lsum = 0.0f;

    const uint64_t ntiles = (n + 15) / 16;
    for (uint64_t tile = blockIdx.x * 4ull + wv; tile < ntiles; tile += gridDim.x * 4ull) {
        const uint64_t row0 = tile * 16;
        const int rows = (int)(n - row0 < 16 ? n - row0 : 16);
        f32x4 z1[4], z2[4];
        float yr[4];
        forward_tile(P, X, row0, rows, s, wv, l, z1, z2, yr);
        // dy_i = 2 (y_i - t_i) / n for the lane's rows 4 lg + r
        float dy[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 4 * lg + r;
            const float e = i < rows ? yr[r] - T[row0 + i] : 0.0f;
            dy[r] = 2.0f * e * inv_n;
            if (lr == 0) lsum += e * e * inv_n;
        }
        if (lr == 0) gb3 += dy[0] + dy[1] + dy[2] + dy[3];
        // output layer and dZ2 (D layout)
        f32x4 dz2[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const float w = P[MLP_W3 + 16 * nt + lr];
            float g3 = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                g3 += dy[r] * gelu(z2[nt][r]);
                dz2[nt][r] = dy[r] * w * gelu_grad(z2[nt][r]);
            }
            gw3[nt] += g3;
            gb2[nt] += dz2[nt][0] + dz2[nt][1] + dz2[nt][2] + dz2[nt][3];
        }
        // dZ2 -> LDS (BF16)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int r = 0; r < 4; r++) s.dz[wv][4 * lg + r][16 * nt + lr] = to_bf16(dz2[nt][r]);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW2 += dZ2^T G1   (M = out 64, N = in 64, K = the 16 samples)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            bf16x4 a;
#pragma unroll
            for (int t = 0; t < 4; t++) a[t] = s.dz[wv][4 * lg + t][16 * mt + lr];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                bf16x4 b;
#pragma unroll
                for (int t = 0; t < 4; t++) b[t] = s.g1[wv][4 * lg + t][16 * nt + lr];
                gw2[mt][nt] = mfma(a, b, gw2[mt][nt]);
            }
        }
        // dG1 = dZ2 W2   (M = 16 samples, N = in 64, K = out 64)
        f32x4 dz1[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                bf16x4 a, b;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    a[t] = s.dz[wv][lr][16 * ks + 4 * lg + t];
                    b[t] = s.w2[16 * ks + 4 * lg + t][16 * nt + lr];
                }
                acc = mfma(a, b, acc);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) acc[r] *= gelu_grad(z1[nt][r]);
            dz1[nt] = acc;
            gb1[nt] += acc[0] + acc[1] + acc[2] + acc[3];
        }
        __builtin_amdgcn_wave_barrier();   // every lane has read dZ2
        // dZ1 -> LDS (BF16), then dW1 += dZ1^T X   (M = 64, N = 32, K = 16)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int r = 0; r < 4; r++) s.dz[wv][4 * lg + r][16 * nt + lr] = to_bf16(dz1[nt][r]);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            bf16x4 a;
#pragma unroll
            for (int t = 0; t < 4; t++) a[t] = s.dz[wv][4 * lg + t][16 * mt + lr];
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                bf16x4 b;
#pragma unroll
                for (int t = 0; t < 4; t++) b[t] = s.x[wv][4 * lg + t][16 * nt + lr];
                gw1[mt][nt] = mfma(a, b, gw1[mt][nt]);
            }
        }
        __builtin_amdgcn_wave_barrier();   // LDS tiles are rewritten by the next tile
    }

    // fold the wave's accumulators into the workgroup's LDS gradient
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * mt + 4 * lg + r;
#pragma unroll
            for (int nt = 0; nt < 4; nt++) atomicAdd(&s_grad[MLP_W2 + row * H + 16 * nt + lr], gw2[mt][nt][r]);
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int col = 16 * nt + lr;
                if (col < MLP_IN) atomicAdd(&s_grad[MLP_W1 + row * MLP_IN + col], gw1[mt][nt][r]);
            }
        }
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        float b1 = gb1[nt], b2 = gb2[nt], w3 = gw3[nt];
        b1 += __shfl_xor(b1, 16, 64); b1 += __shfl_xor(b1, 32, 64);
        b2 += __shfl_xor(b2, 16, 64); b2 += __shfl_xor(b2, 32, 64);
        w3 += __shfl_xor(w3, 16, 64); w3 += __shfl_xor(w3, 32, 64);
        if (lg == 0) {
            atomicAdd(&s_grad[MLP_B1 + 16 * nt + lr], b1);
            atomicAdd(&s_grad[MLP_B2 + 16 * nt + lr], b2);
            atomicAdd(&s_grad[MLP_W3 + 16 * nt + lr], w3);
        }
    }
    gb3 += __shfl_xor(gb3, 16, 64);
    gb3 += __shfl_xor(gb3, 32, 64);
    lsum += __shfl_xor(lsum, 16, 64);
    lsum += __shfl_xor(lsum, 32, 64);
    if (l == 0) {
        atomicAdd(&s_grad[MLP_B3], gb3);
        s_loss[wv] = lsum;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < MLP_PARAMS; e += blockDim.x)
        part[(uint64_t)blockIdx.x * MLP_PARAMS + e] = s_grad[e];
    if (threadIdx.x == 0) loss_part[blockIdx.x] = s_loss[0] + s_loss[1] + s_loss[2] + s_loss[3];
}

// grad[p] = sum over workgroups of part[wg][p]; loss = sum of loss_part.
__global__ void k_mlp_fold(const float *part, const float *loss_part, int nparts, float *grad, float *loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_PARAMS) {
        float g = 0.0f;
        for (int w = 0; w < nparts; w++) g += part[(uint64_t)w * MLP_PARAMS + p];
        grad[p] = g;
    }
    if (p == 0) {
        float s = 0.0f;
        for (int w = 0; w < nparts; w++) s += loss_part[w];
        *loss = s;
    }
}

// Adam (bias-corrected) on the FP32 master parameters.
__global__ void k_mlp_adam(float *P, const float *grad, float *m, float *v, float lr, float b1,
                           float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static unsigned mlp_grid(uint64_t n)
{
    uint64_t tiles = (n + 15) / 16, g = (tiles + 3) / 4;
    if (g < 1) g = 1;
    if (g > MLP_MAX_WG) g = MLP_MAX_WG;
    return (unsigned)g;
}

uint64_t mlp_grad_workspace_bytes() { return (uint64_t)MLP_MAX_WG * (MLP_PARAMS + 1) * 4 + 256; }

void mlp_forward_launch(const float *P, const float *X, uint64_t n, float *y, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_grid(n)), dim3(256), 0, st, P, X, n, y);
}

void mlp_grad_launch(const float *P, const float *X, const float *T, uint64_t n, float *grad,
                     float *loss, uint8_t *ws, hipStream_t st)
{
    const unsigned g = mlp_grid(n);
    float *part = reinterpret_cast<float *>(ws);
    float *lpart = part + (uint64_t)MLP_MAX_WG * MLP_PARAMS;
    hipLaunchKernelGGL(k_mlp_grad, dim3(g), dim3(256), 0, st, P, X, T, n, part, lpart);
    hipLaunchKernelGGL(k_mlp_fold, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, st, part, lpart,
                       (int)g, grad, loss);
}

void mlp_adam_launch(float *P, const float *grad, float *m, float *v, float lr, float b1, float b2,
                     float eps, int step, hipStream_t st)
{
    float c1 = 1.0f, c2 = 1.0f;
    for (int i = 0; i < step; i++) {
        c1 *= b1;
        c2 *= b2;
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, st, P, grad, m, v,
                       lr, b1, b2, eps, 1.0f - c1, 1.0f - c2);
}

}  // namespace szxamd
