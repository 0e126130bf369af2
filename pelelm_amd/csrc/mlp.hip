// Learned Newton predictor: a 4-layer perceptron (n_in -> 128 -> 128 -> 128 -> n_out,
// GELU on the hidden layers) with BF16 weights and activations multiplied on the
// matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), and its training:
// forward with stored pre-activations, backward (MSE loss), gradient reduction and
// Adam on FP32 master weights.
//
// The reference has no counterpart: this is the learned part of the chosen reactor
// form.  Inputs per cell are (T, log p, log10 Y_k, log dt) and outputs (dT, dY_k),
// both normalised by the caller's means and deviations.
//
// Layouts (row = one sample / cell):
//   master parameters (FP32), per layer l: W_l[n_out_l][n_in_l] row-major, then b_l[n_out_l]
//   device copies per layer: wt[NP][KP] (bf16, output-major: k contiguous, feeds the
//     forward B operand), wk[KP][NP] (bf16, input-major: feeds the backward dH = G W),
//     bias[NP] (FP32); KP / NP are the padded (multiple of 16) fan-in / fan-out
//   activations in LDS: [64 rows][KP + 8] bf16 per buffer
//
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pelelm_mlp.h"
#include "mlp_internal.h"

namespace pelelm {
namespace {

constexpr int ROWS = MLP_ROWS;
constexpr int THREADS = MLP_THREADS;
constexpr int LDW = MLP_LDW;

// X rows [row0, row0 + 64) -> bf16 LDS (zero beyond n_in and beyond B)
template <int KIN>
__device__ __forceinline__ void load_input(unsigned short* H, const float* __restrict__ X, int n_in, long row0, long B) {
  for (int e = threadIdx.x; e < ROWS * KIN; e += THREADS) {
    const int r = e / KIN, k = e % KIN;
    const long row = row0 + r;
    const float v = (row < B && k < n_in) ? X[row * n_in + k] : 0.f;
    H[r * LDW + k] = f2bf(v);
  }
}

// forward over 64 rows.  Z (nullable): pre-activations of the hidden layers, [3][B][128].
template <int KIN, int NOUT>
__global__ void __launch_bounds__(THREADS) k_mlp_fwd(const MlpDev w, const float* __restrict__ X, long B,
                                                    float* __restrict__ Y, float* __restrict__ Z) {
  __shared__ unsigned short h0[ROWS * LDW], h1[ROWS * LDW];
  const long row0 = (long)blockIdx.x * ROWS;
  load_input<KIN>(h0, X, w.n_in, row0, B);
  __syncthreads();
  unsigned short* hin = h0;
  unsigned short* hout = h1;
#pragma unroll 1
  for (int l = 0; l < 3; ++l) {
    auto emit = [&](int r, int c, float z) {
      const long row = row0 + r;
      if (Z && row < B) Z[((long)l * B + row) * MLP_HIDDEN + c] = z;
      hout[r * LDW + c] = f2bf(mlp_gelu(z));
    };
    if (l == 0)
      mlp_dense<KIN, MLP_HIDDEN>(hin, w.wt[0], w.b[0], emit);
    else
      mlp_dense<MLP_HIDDEN, MLP_HIDDEN>(hin, w.wt[l], w.b[l], emit);
    __syncthreads();
    unsigned short* t = hin;
    hin = hout;
    hout = t;
  }
  const int n_out = w.n_out;
  mlp_dense<MLP_HIDDEN, NOUT>(hin, w.wt[3], w.b[3], [&](int r, int c, float z) {
    const long row = row0 + r;
    if (row < B && c < n_out) Y[row * n_out + c] = z;
  });
}

// Backward of the mean-squared error L = sum (Y - Yt)^2 / (B n_out) over 64 rows; writes
// this workgroup's parameter gradients (master layout) to part[blockIdx.x][nparams]
// and its loss contribution to loss_part[blockIdx.x].
template <int KIN, int NOUT>
__global__ void __launch_bounds__(THREADS) k_mlp_bwd(const MlpDev w, const float* __restrict__ X,
                                                    const float* __restrict__ Z, const float* __restrict__ Y,
                                                    const float* __restrict__ Yt, long B, float* __restrict__ part,
                                                    float* __restrict__ loss_part) {
  __shared__ unsigned short g[ROWS * LDW], hin[ROWS * LDW], gn[ROWS * LDW];
  __shared__ float red[THREADS];
  const long row0 = (long)blockIdx.x * ROWS;
  const int n_out = w.n_out;
  const float scale = 2.0f / ((float)B * (float)n_out);
  float* P = part + (long)blockIdx.x * w.nparams;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;

  // dL/dY and the loss
  float lsum = 0.f;
  for (int e = threadIdx.x; e < ROWS * NOUT; e += THREADS) {
    const int r = e / NOUT, c = e % NOUT;
    const long row = row0 + r;
    float d = 0.f;
    if (row < B && c < n_out) {
      d = Y[row * n_out + c] - Yt[row * n_out + c];
      lsum += d * d;
    }
    g[r * LDW + c] = f2bf(d * scale);
  }
  red[threadIdx.x] = lsum;
  __syncthreads();
  for (int s = THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] / ((float)B * (float)n_out);

  // input activations of layer l (0-based): X for l = 0, else bf16(gelu(Z[l-1]))
  auto load_hin = [&](int l, int KP) {
    for (int e = threadIdx.x; e < ROWS * KP; e += THREADS) {
      const int r = e / KP, k = e % KP;
      const long row = row0 + r;
      float v = 0.f;
      if (row < B) {
        if (l == 0)
          v = (k < w.n_in) ? X[row * w.n_in + k] : 0.f;
        else
          v = mlp_gelu(Z[((long)(l - 1) * B + row) * MLP_HIDDEN + k]);
      }
      hin[r * LDW + k] = f2bf(v);
    }
  };

#pragma unroll 1
  for (int l = 3; l >= 0; --l) {
    const int KP = (l == 0) ? KIN : MLP_HIDDEN;
    const int NP = (l == 3) ? NOUT : MLP_HIDDEN;
    const int fan_in = (l == 0) ? w.n_in : MLP_HIDDEN;
    const int fan_out = (l == 3) ? n_out : MLP_HIDDEN;
    load_hin(l, KP);
    __syncthreads();
    // dW[n][k] = sum_r g[r][n] hin[r][k]: tiles (nt, kt) over the four waves
    const int ntiles = (NP / 16) * (KP / 16);
    float* Pw = P + w.off_w[l];
#pragma unroll 1
    for (int t = wave; t < ntiles; t += 4) {
      const int nt = t / (KP / 16), kt = t % (KP / 16);
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int rs = 0; rs < ROWS / 32; ++rs) {
        unsigned short a[8], b[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int r = rs * 32 + 8 * lq + i;
          a[i] = g[r * LDW + nt * 16 + lr];
          b[i] = hin[r * LDW + kt * 16 + lr];
        }
        acc = mlp_mfma(mlp_pack8(a), mlp_pack8(b), acc);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int n = nt * 16 + 4 * lq + i, k = kt * 16 + lr;
        if (n < fan_out && k < fan_in) Pw[n * fan_in + k] = acc[i];
      }
    }
    // db[n] = sum_r g[r][n]
    for (int n = threadIdx.x; n < fan_out; n += THREADS) {
      float s = 0.f;
      for (int r = 0; r < ROWS; ++r) s += bf2f(g[r * LDW + n]);
      P[w.off_b[l] + n] = s;
    }
    if (l == 0) break;
    // gn[r][k] = (sum_n g[r][n] W[n][k]) * gelu'(Z[l-1][r][k]); wave w owns rows 16w..16w+15
    const unsigned short* wk = w.wk[l];
#pragma unroll 1
    for (int ct = 0; ct < MLP_HIDDEN / 16; ++ct) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < NP / 32; ++kt) {
        const bf16x8 a = mlp_ld8(&g[(wave * 16 + lr) * LDW + kt * 32 + 8 * lq]);
        const bf16x8 b = mlp_ld8(&wk[(long)(ct * 16 + lr) * NP + kt * 32 + 8 * lq]);
        acc = mlp_mfma(a, b, acc);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wave * 16 + 4 * lq + i, k = ct * 16 + lr;
        const long row = row0 + r;
        const float d = (row < B) ? acc[i] * mlp_gelu_d(Z[((long)(l - 1) * B + row) * MLP_HIDDEN + k]) : 0.f;
        gn[r * LDW + k] = f2bf(d);
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ROWS * MLP_HIDDEN; e += THREADS) {
      const int r = e / MLP_HIDDEN, k = e % MLP_HIDDEN;
      g[r * LDW + k] = gn[r * LDW + k];
    }
    __syncthreads();
  }
}

// grad[p] = sum over blocks of part[b][p]; loss = sum of loss_part
__global__ void k_mlp_reduce(const float* __restrict__ part, int nblk, int nparams, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nparams) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(long)b * nparams + p];
  grad[p] = s;
}

// Adam on the master parameters (bias-corrected), then the bf16 / FP32 device copies
__global__ void k_mlp_adam(const MlpDev w, float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.nparams) return;
  const float gr = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * gr;
  const float v = b2 * m2[p] + (1.f - b2) * gr * gr;
  m1[p] = a;
  m2[p] = v;
  master[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// master -> device copies (bf16 weights in both orientations, FP32 biases); pads stay zero
__global__ void k_mlp_pack(MlpDev w, const float* __restrict__ master) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.nparams) return;
  for (int l = 3; l >= 0; --l) {
    if (p >= w.off_b[l]) {
      if (p < w.off_b[l] + w.fan_out[l]) w.bias_rw[l][p - w.off_b[l]] = master[p];
      return;
    }
    if (p >= w.off_w[l]) {
      const int q = p - w.off_w[l];
      const int n = q / w.fan_in[l], k = q % w.fan_in[l];
      const unsigned short h = f2bf(master[p]);
      w.wt_rw[l][(long)n * w.kp[l] + k] = h;
      w.wk_rw[l][(long)k * w.np[l] + n] = h;
      return;
    }
  }
}

template <int KIN, int NOUT>
hipError_t launch_fwd_t(const MlpDev& w, const float* X, long B, float* Y, float* Z, hipStream_t s) {
  hipLaunchKernelGGL((k_mlp_fwd<KIN, NOUT>), dim3((unsigned)((B + ROWS - 1) / ROWS)), dim3(THREADS), 0, s, w, X, B,
                     Y, Z);
  return hipGetLastError();
}
template <int KIN, int NOUT>
hipError_t launch_bwd_t(const MlpDev& w, const float* X, const float* Z, const float* Y, const float* Yt, long B,
                        float* part, float* loss_part, hipStream_t s) {
  hipLaunchKernelGGL((k_mlp_bwd<KIN, NOUT>), dim3((unsigned)((B + ROWS - 1) / ROWS)), dim3(THREADS), 0, s, w, X, Z, Y,
                     Yt, B, part, loss_part);
  return hipGetLastError();
}

}  // namespace

hipError_t mlp_launch_fwd(const MlpDev& w, const float* X, long B, float* Y, float* Z, hipStream_t s) {
  if (w.kp[0] == 32 && w.np[3] == 32) return launch_fwd_t<32, 32>(w, X, B, Y, Z, s);
  if (w.kp[0] == 32 && w.np[3] == 64) return launch_fwd_t<32, 64>(w, X, B, Y, Z, s);
  if (w.kp[0] == 64 && w.np[3] == 32) return launch_fwd_t<64, 32>(w, X, B, Y, Z, s);
  if (w.kp[0] == 64 && w.np[3] == 64) return launch_fwd_t<64, 64>(w, X, B, Y, Z, s);
  return hipErrorInvalidValue;
}
hipError_t mlp_launch_bwd(const MlpDev& w, const float* X, const float* Z, const float* Y, const float* Yt, long B,
                          float* part, float* loss_part, hipStream_t s) {
  if (w.kp[0] == 32 && w.np[3] == 32) return launch_bwd_t<32, 32>(w, X, Z, Y, Yt, B, part, loss_part, s);
  if (w.kp[0] == 32 && w.np[3] == 64) return launch_bwd_t<32, 64>(w, X, Z, Y, Yt, B, part, loss_part, s);
  if (w.kp[0] == 64 && w.np[3] == 32) return launch_bwd_t<64, 32>(w, X, Z, Y, Yt, B, part, loss_part, s);
  if (w.kp[0] == 64 && w.np[3] == 64) return launch_bwd_t<64, 64>(w, X, Z, Y, Yt, B, part, loss_part, s);
  return hipErrorInvalidValue;
}

}  // namespace pelelm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace pelelm;

namespace {

struct Mlp {
  MlpDev d{};
  int n_in = 0, n_out = 0, nparams = 0;
  unsigned short* wbuf = nullptr;       // bf16 copies of every layer (both orientations)
  float* fbuf = nullptr;                 // biases, master, Adam moments, gradient, normalisation
  float* master = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* grad = nullptr;
  float* norm = nullptr;                 // in_mean, in_std, out_mean, out_std (FP32, padded to 64 each)
  long cap = 0;                          // rows the scratch below holds
  float* Z = nullptr;
  float* Yb = nullptr;
  float* part = nullptr;
  float* loss_part = nullptr;
  long step = 0;
  std::string err;
};

int set_err(Mlp* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}

int pad_io(int n) { return n <= 32 ? 32 : 64; }

#define MLP_HIP(call)                                                                  \
  do {                                                                                 \
    const hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) return set_err(m, MI_MLP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int ensure_rows(Mlp* m, long B) {
  if (B <= m->cap) return MI_MLP_OK;
  (void)hipFree(m->Z);
  (void)hipFree(m->Yb);
  (void)hipFree(m->part);
  (void)hipFree(m->loss_part);
  m->Z = m->Yb = m->part = m->loss_part = nullptr;
  m->cap = 0;
  const long nblk = (B + ROWS - 1) / ROWS;
  MLP_HIP(hipMalloc(&m->Z, sizeof(float) * 3 * B * MLP_HIDDEN));
  MLP_HIP(hipMalloc(&m->Yb, sizeof(float) * B * m->n_out));
  MLP_HIP(hipMalloc(&m->part, sizeof(float) * nblk * m->nparams));
  MLP_HIP(hipMalloc(&m->loss_part, sizeof(float) * nblk));
  m->cap = B;
  return MI_MLP_OK;
}

int repack(Mlp* m, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((m->nparams + 255) / 256), dim3(256), 0, s, m->d, m->master);
  MLP_HIP(hipGetLastError());
  return MI_MLP_OK;
}

// forward with stored pre-activations, backward, block reduction into m->grad; loss to host
int forward_backward(Mlp* m, const float* X, const float* Yt, long B, float* loss, hipStream_t s) {
  int rc = ensure_rows(m, B);
  if (rc != MI_MLP_OK) return rc;
  MLP_HIP(mlp_launch_fwd(m->d, X, B, m->Yb, m->Z, s));
  MLP_HIP(mlp_launch_bwd(m->d, X, m->Z, m->Yb, Yt, B, m->part, m->loss_part, s));
  const int nblk = (int)((B + ROWS - 1) / ROWS);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((m->nparams + 255) / 256), dim3(256), 0, s, m->part, nblk, m->nparams,
                     m->grad);
  MLP_HIP(hipGetLastError());
  if (loss) {
    std::vector<float> lp(nblk);
    MLP_HIP(hipMemcpyAsync(lp.data(), m->loss_part, sizeof(float) * nblk, hipMemcpyDeviceToHost, s));
    MLP_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (float v : lp) t += v;
    *loss = (float)t;
  }
  return MI_MLP_OK;
}

}  // namespace

namespace pelelm {
const MlpDev* mlp_device(void* h) { return h ? &static_cast<Mlp*>(h)->d : nullptr; }
}  // namespace pelelm

extern "C" {

int mi_mlp_create(int n_in, int n_out, unsigned long long seed, void** out) {
  if (!out) return MI_MLP_EINVAL;
  *out = nullptr;
  if (n_in < 1 || n_in > 64 || n_out < 1 || n_out > 64) return MI_MLP_EINVAL;
  Mlp* m = new Mlp();
  m->n_in = n_in;
  m->n_out = n_out;
  const int fi[4] = {n_in, MLP_HIDDEN, MLP_HIDDEN, MLP_HIDDEN};
  const int fo[4] = {MLP_HIDDEN, MLP_HIDDEN, MLP_HIDDEN, n_out};
  const int kp[4] = {pad_io(n_in), MLP_HIDDEN, MLP_HIDDEN, MLP_HIDDEN};
  const int np[4] = {MLP_HIDDEN, MLP_HIDDEN, MLP_HIDDEN, pad_io(n_out)};
  int off = 0;
  long wsz = 0, bsz = 0;
  for (int l = 0; l < 4; ++l) {
    m->d.fan_in[l] = fi[l];
    m->d.fan_out[l] = fo[l];
    m->d.kp[l] = kp[l];
    m->d.np[l] = np[l];
    m->d.off_w[l] = off;
    off += fi[l] * fo[l];
    m->d.off_b[l] = off;
    off += fo[l];
    wsz += 2L * kp[l] * np[l];
    bsz += np[l];
  }
  m->nparams = off;
  m->d.nparams = off;
  m->d.n_in = n_in;
  m->d.n_out = n_out;
  const long fsz = bsz + 4L * off + 4 * 64;
  if (hipMalloc(&m->wbuf, sizeof(unsigned short) * wsz) != hipSuccess ||
      hipMalloc(&m->fbuf, sizeof(float) * fsz) != hipSuccess) {
    (void)hipFree(m->wbuf);
    (void)hipFree(m->fbuf);
    delete m;
    return MI_MLP_ENOMEM;
  }
  if (hipMemset(m->wbuf, 0, sizeof(unsigned short) * wsz) != hipSuccess ||
      hipMemset(m->fbuf, 0, sizeof(float) * fsz) != hipSuccess) {
    (void)hipFree(m->wbuf);
    (void)hipFree(m->fbuf);
    delete m;
    return MI_MLP_EHIP;
  }
  unsigned short* wp = m->wbuf;
  float* fp = m->fbuf;
  for (int l = 0; l < 4; ++l) {
    m->d.wt_rw[l] = wp;
    m->d.wt[l] = wp;
    wp += (long)kp[l] * np[l];
    m->d.wk_rw[l] = wp;
    m->d.wk[l] = wp;
    wp += (long)kp[l] * np[l];
    m->d.bias_rw[l] = fp;
    m->d.b[l] = fp;
    fp += np[l];
  }
  m->master = fp;
  fp += off;
  m->m1 = fp;
  fp += off;
  m->m2 = fp;
  fp += off;
  m->grad = fp;
  fp += off;
  m->norm = fp;
  m->d.norm = fp;
  // He-uniform initial weights from a splitmix64 stream of `seed`, zero biases
  std::vector<float> h(off, 0.f);
  unsigned long long x = seed;
  auto next = [&x]() {
    unsigned long long z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  for (int l = 0; l < 4; ++l) {
    const float lim = std::sqrt(6.0f / (float)fi[l]);
    for (int e = 0; e < fi[l] * fo[l]; ++e) {
      const double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
      h[m->d.off_w[l] + e] = (float)((2.0 * u - 1.0) * lim);
    }
  }
  std::vector<float> nrm(4 * 64, 0.f);
  for (int i = 0; i < 64; ++i) {
    nrm[64 + i] = 1.f;
    nrm[192 + i] = 1.f;
  }
  if (hipMemcpy(m->master, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->norm, nrm.data(), sizeof(float) * 256, hipMemcpyHostToDevice) != hipSuccess ||
      repack(m, nullptr) != MI_MLP_OK || hipDeviceSynchronize() != hipSuccess) {
    mi_mlp_close(m);
    return MI_MLP_EHIP;
  }
  *out = m;
  return MI_MLP_OK;
}

int mi_mlp_nparams(void* h) { return h ? static_cast<Mlp*>(h)->nparams : MI_MLP_EINVAL; }

int mi_mlp_get_params(void* h, float* host, int n) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !host || n != m->nparams) return set_err(m, MI_MLP_EINVAL, "parameter count mismatch");
  MLP_HIP(hipMemcpy(host, m->master, sizeof(float) * n, hipMemcpyDeviceToHost));
  return MI_MLP_OK;
}

int mi_mlp_set_params(void* h, const float* host, int n) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !host || n != m->nparams) return set_err(m, MI_MLP_EINVAL, "parameter count mismatch");
  MLP_HIP(hipMemcpy(m->master, host, sizeof(float) * n, hipMemcpyHostToDevice));
  int rc = repack(m, nullptr);
  if (rc != MI_MLP_OK) return rc;
  MLP_HIP(hipDeviceSynchronize());
  return MI_MLP_OK;
}

int mi_mlp_set_normalisation(void* h, const double* in_mean, const double* in_std, const double* out_mean,
                             const double* out_std) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !in_mean || !in_std || !out_mean || !out_std) return set_err(m, MI_MLP_EINVAL, "null argument");
  std::vector<float> v(256, 0.f);
  for (int i = 0; i < m->n_in; ++i) {
    if (!(in_std[i] > 0.0)) return set_err(m, MI_MLP_EINVAL, "input deviations must be positive");
    v[i] = (float)in_mean[i];
    v[64 + i] = (float)in_std[i];
  }
  for (int i = 0; i < m->n_out; ++i) {
    if (!(out_std[i] > 0.0)) return set_err(m, MI_MLP_EINVAL, "output deviations must be positive");
    v[128 + i] = (float)out_mean[i];
    v[192 + i] = (float)out_std[i];
  }
  MLP_HIP(hipMemcpy(m->norm, v.data(), sizeof(float) * 256, hipMemcpyHostToDevice));
  return MI_MLP_OK;
}

int mi_mlp_forward(void* h, const float* X, float* Y, long B, void* stream) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !X || !Y || B < 0) return set_err(m, MI_MLP_EINVAL, "bad forward arguments");
  if (B == 0) return MI_MLP_OK;
  MLP_HIP(mlp_launch_fwd(m->d, X, B, Y, nullptr, (hipStream_t)stream));
  return MI_MLP_OK;
}

int mi_mlp_grad(void* h, const float* X, const float* Yt, long B, float* grad, float* loss, void* stream) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !X || !Yt || !grad || B < 1) return set_err(m, MI_MLP_EINVAL, "bad gradient arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc = forward_backward(m, X, Yt, B, loss, s);
  if (rc != MI_MLP_OK) return rc;
  MLP_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * m->nparams, hipMemcpyDeviceToDevice, s));
  return MI_MLP_OK;
}

int mi_mlp_train_step(void* h, const float* X, const float* Yt, long B, float lr, float* loss, void* stream) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !X || !Yt || B < 1 || !(lr > 0.f)) return set_err(m, MI_MLP_EINVAL, "bad training arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc = forward_backward(m, X, Yt, B, loss, s);
  if (rc != MI_MLP_OK) return rc;
  ++m->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((m->nparams + 255) / 256), dim3(256), 0, s, m->d, m->master, m->m1, m->m2,
                     m->grad, lr, b1, b2, 1e-8f, c1, c2);
  MLP_HIP(hipGetLastError());
  return repack(m, s);
}

const char* mi_mlp_last_error(void* h) { return h ? static_cast<Mlp*>(h)->err.c_str() : "null handle"; }

int mi_mlp_close(void* h) {
  Mlp* m = static_cast<Mlp*>(h);
  if (!m) return MI_MLP_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(m->wbuf);
  (void)hipFree(m->fbuf);
  (void)hipFree(m->Z);
  (void)hipFree(m->Yb);
  (void)hipFree(m->part);
  (void)hipFree(m->loss_part);
  delete m;
  return MI_MLP_OK;
}

}  // extern "C"
