// The learned initial guess of the rate-and-state Newton solve: a
// 6 -> 64 -> 64 -> 64 -> 1 GELU network evaluated by one wavefront for 64
// fault nodes (or training samples) at a time.
//
// Numerics: weights and activations enter the matrix products as bfloat16
// (round to nearest even), products accumulate in float32 on the matrix cores
// (v_mfma_f32_32x32x16_bf16); biases, GELU (erf form) and the loss are
// float32.  The 64x64 hidden layers run on MFMA; layer 1 runs on MFMA with
// its K = 6 padded to 16; the 64 -> 1 output layer is a float32 dot product
// per lane.
//
// MFMA 32x32x16 operand layout (lane l, r = l % 32, h = l / 32):
//   A (32 x 16): lane holds A[r][8h .. 8h+7]
//   B (16 x 32): lane holds B[8h .. 8h+7][r]
//   C (32 x 32): lane holds C[8(i/4) + 4h + i%4][r] in accumulator i = 0..15
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eqd {

constexpr int kMlpIn = 6, kMlpH = 64, kMlpK1 = 16;
// parameter vector (float32): W1[64][6] b1 W2[64][64] b2 W3[64][64] b3 w4[64] b4
// in_shift[6] in_scale[6] out_shift out_scale
constexpr int kOffW1 = 0, kOffB1 = 384, kOffW2 = 448, kOffB2 = 4544, kOffW3 = 4608, kOffB3 = 8704, kOffW4 = 8768,
              kOffB4 = 8832, kMlpNTrain = 8833, kOffNorm = 8833, kMlpNParams = 8847;
// row stride (bf16 elements) of a 64-wide activation tile in LDS: 64 + 8 keeps
// rows 16-byte aligned and spreads the lanes of one read over the banks
constexpr int kLd = 72;
constexpr int kLd0 = 24;  // layer-1 input tile, K = 16 + 8

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Device view of one network.  w1 is [64][16] (K padded with zeros), w2/w3
// are [out][in], w2t/w3t their transposes; w4r holds the bfloat16-rounded
// output weights as float; norm = in_shift[6], in_scale[6], out_shift, out_scale.
struct MlpNet {
    const __bf16 *w1, *w2, *w3, *w2t, *w3t;
    const float *b1, *b2, *b3, *w4r, *b4, *norm;
};

__device__ __forceinline__ __bf16 to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}

__device__ __forceinline__ float from_bf16(__bf16 b)
{
    return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// acc[mt][nt] += A * B over K, for the 64 x 64 result of one wave:
// A is 64 x K row-major in memory (stride lda), B[k][n] = W[n][k] with W
// row-major (stride ldw).  K is a multiple of 16.
__device__ __forceinline__ void wave_mm(const __bf16 *A, int lda, const __bf16 *W, int ldw, int K,
                                        f32x16 acc[2][2])
{
    const int l = threadIdx.x & 63, r = l & 31, h = l >> 5;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            a[t] = *(const bf16x8 *)(A + (t * 32 + r) * lda + k0 + 8 * h);
            b[t] = *(const bf16x8 *)(W + (t * 32 + r) * ldw + k0 + 8 * h);
        }
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
}

__device__ __forceinline__ void zero_acc(f32x16 acc[2][2])
{
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[mt][nt][i] = 0.0f;
}

// row m and column n of accumulator i of tile (mt, nt) for this lane
__device__ __forceinline__ int acc_row(int mt, int i) { return mt * 32 + 8 * (i >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (i & 3); }
__device__ __forceinline__ int acc_col(int nt) { return nt * 32 + (threadIdx.x & 31); }

// out[m][n] = bf16(gelu(acc + bias[n])); when zsave is given it receives gelu'(z) (float, stride 64)
__device__ __forceinline__ void epilogue_gelu(const f32x16 acc[2][2], const float *bias, __bf16 *out, float *gsave)
{
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            const int n = acc_col(nt);
            const float b = bias[n];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = acc_row(mt, i);
                const float z = acc[mt][nt][i] + b;
                out[m * kLd + n] = to_bf16(gelu(z));
                if (gsave) gsave[m * 64 + n] = gelu_grad(z);
            }
        }
}

// Normalised, bfloat16-rounded layer-1 input row of this lane (sample = lane).
__device__ __forceinline__ void write_input_row(const MlpNet &M, const float x[kMlpIn], __bf16 *a0)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) a0[l * kLd0 + k] = to_bf16((x[k] - M.norm[k]) * M.norm[kMlpIn + k]);
#pragma unroll
    for (int k = kMlpIn; k < kMlpK1; k++) a0[l * kLd0 + k] = to_bf16(0.0f);
}

// Forward pass of one wave: lane l holds sample l's raw inputs x and gets
// its network output (normalised; ln V = y * out_scale + out_shift).  buf0 is
// a [64][kLd0] and buf a [64][kLd] bf16 LDS tile private to this wave.
// Every lane of every wave of the workgroup must call it (it synchronises
// the workgroup).
__device__ __forceinline__ float mlp_wave_forward(const MlpNet &M, const float x[kMlpIn], __bf16 *buf0, __bf16 *buf)
{
    const int l = threadIdx.x & 63;
    write_input_row(M, x, buf0);
    __syncthreads();
    f32x16 acc[2][2];
    zero_acc(acc);
    wave_mm(buf0, kLd0, M.w1, kMlpK1, kMlpK1, acc);
    epilogue_gelu(acc, M.b1, buf, nullptr);
    __syncthreads();
    zero_acc(acc);
    wave_mm(buf, kLd, M.w2, 64, 64, acc);
    __syncthreads();
    epilogue_gelu(acc, M.b2, buf, nullptr);
    __syncthreads();
    zero_acc(acc);
    wave_mm(buf, kLd, M.w3, 64, 64, acc);
    __syncthreads();
    epilogue_gelu(acc, M.b3, buf, nullptr);
    __syncthreads();
    float y = M.b4[0];
    for (int n = 0; n < 64; n++) y = y + from_bf16(buf[l * kLd + n]) * M.w4r[n];
    __syncthreads();
    return y;
}

}  // namespace eqd
