// Device view of the learned predictor (mlp.hip) shared with the reactor kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace pelelm {

constexpr int MLP_HIDDEN = 128;

// round-to-nearest-even FP32 -> bf16 (NaN kept quiet), and back
__host__ __device__ inline unsigned short f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct MlpDev {
  int n_in, n_out, nparams;
  int fan_in[4], fan_out[4], kp[4], np[4], off_w[4], off_b[4];
  const unsigned short* wt[4];   // [np][kp] bf16
  const unsigned short* wk[4];   // [kp][np] bf16
  const float* b[4];             // [np]
  unsigned short* wt_rw[4];
  unsigned short* wk_rw[4];
  float* bias_rw[4];
  const float* norm;             // in_mean[64], in_std[64], out_mean[64], out_std[64]
};

// ---------------------------------------------------------------------------
// device building blocks shared by the trainer (mlp.hip) and the reactor's predictor
// kernel.  A workgroup of MLP_THREADS (four waves) owns MLP_ROWS samples.
// gfx950's MFMA 16x16x32 bf16 (v_mfma_f32_16x16x32_bf16, twice the K of CDNA3's
// 16x16x16 form per instruction) operand layout, lane l of 64 (checked by a GPU unit
// test on raw registers):
//   A (16 x 32, M x K): A[l % 16][8 (l / 16) + i], i = 0..7
//   B (32 x 16, K x N): B[8 (l / 16) + i][l % 16], i = 0..7
//   D (16 x 16, M x N): D[4 (l / 16) + v][l % 16], v = 0..3
// ---------------------------------------------------------------------------
constexpr int MLP_ROWS = 64;              // samples per workgroup
constexpr int MLP_THREADS = 256;          // four waves, 16 rows each in the row-tiled products
constexpr int MLP_LDW = MLP_HIDDEN + 8;   // LDS row stride of an activation buffer (bf16)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float mlp_gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float mlp_gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ v4f mlp_mfma(bf16x8 a, bf16x8 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// eight consecutive bf16 words (16-byte aligned) as one operand
__device__ __forceinline__ bf16x8 mlp_ld8(const unsigned short* p) {
  bf16x8 v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}
// eight bf16 words gathered one by one
__device__ __forceinline__ bf16x8 mlp_pack8(const unsigned short (&w)[8]) {
  bf16x8 v;
  __builtin_memcpy(&v, w, sizeof(v));
  return v;
}

// one dense layer on the 64 rows of this workgroup: Hin [64][KP] (LDS) x W -> NP outputs;
// z = acc + b is handed to emit(row, col, z) for every (row, col < NP)
template <int KP, int NP, class Emit>
__device__ __forceinline__ void mlp_dense(const unsigned short* Hin, const unsigned short* __restrict__ wt,
                                      const float* __restrict__ bias, Emit emit) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = wave * 16;
  const int lr = lane & 15, lq = lane >> 4;
#pragma unroll 1
  for (int nt = 0; nt < NP / 16; ++nt) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KP / 32; ++kt) {
      const bf16x8 a = mlp_ld8(&Hin[(m0 + lr) * MLP_LDW + kt * 32 + 8 * lq]);
      const bf16x8 b = mlp_ld8(&wt[(long)(nt * 16 + lr) * KP + kt * 32 + 8 * lq]);
      acc = mlp_mfma(a, b, acc);
    }
    const int col = nt * 16 + lr;
    const float bc = bias[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) emit(m0 + 4 * lq + i, col, acc[i] + bc);
  }
}


hipError_t mlp_launch_fwd(const MlpDev& w, const float* X, long B, float* Y, float* Z, hipStream_t s);
hipError_t mlp_launch_bwd(const MlpDev& w, const float* X, const float* Z, const float* Y, const float* Yt, long B,
                          float* part, float* loss_part, hipStream_t s);
// the device view of a handle from mi_mlp_create (nullptr for nullptr)
const MlpDev* mlp_device(void* h);

}  // namespace pelelm
