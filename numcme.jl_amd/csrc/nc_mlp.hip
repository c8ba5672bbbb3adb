// Learned space adapter network: an 8 -> 64 -> 64 -> 1 MLP with tanh-form GELU hidden
// layers and a keep/add logit, computed with BF16 MFMA (v_mfma_f32_16x16x16_bf16) and FP32
// accumulation. Training (binary cross-entropy on the logit, Adam on FP32 master weights)
// runs on the device with the forward and backward kernels below. No reference
// counterpart: NumCME.jl's adapters are rule-based.
//
// Parameters (FP32 master copy, NP = 4801 floats): W1 [8][64] (input-major), b1 [64],
// W2 [64][64] (input-major), b2 [64], w3 [64], b3. Every MFMA operand is rounded to BF16
// (round to nearest even) when it is loaded into registers.
//
// Workgroups have 4 waves; a wave works on 16 samples at a time. For
// v_mfma_f32_16x16x16_bf16 a lane l holds A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j][l % 16]
// and D[4 (l / 16) + j][l % 16], j = 0..3.
#include <algorithm>
#include <cmath>
#include <memory>
#include <random>

#include "nc_mlp.h"

namespace nc {
namespace {

constexpr int kF = 8, kH = 64;
constexpr int OW1 = 0, OB1 = OW1 + kF * kH, OW2 = OB1 + kH, OB2 = OW2 + kH * kH, OW3 = OB2 + kH, OB3 = OW3 + kH;
constexpr int kThreads = 256;
constexpr int kSamples = 64;  // per training group (one workgroup at a time)
constexpr unsigned kFwdMaxBlocks = 4096, kTrainMaxBlocks = 512;

// forward: 4 waves per workgroup, 16 samples per wave and tile
unsigned fwd_grid(int64_t n) {
  const int64_t tiles = (n + 15) / 16;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, kFwdMaxBlocks));
}
// training: one 64-sample group per workgroup and iteration
unsigned train_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kSamples - 1) / kSamples, kTrainMaxBlocks));
}

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// round to nearest even with the hardware conversion (v_cvt_pk_bf16_f32)
__device__ inline short bf16(float x) { return __builtin_bit_cast(short, (__bf16)x); }
__device__ inline v4s bf16x4(v4f x) { return __builtin_bit_cast(v4s, __builtin_convertvector(x, v4bf)); }

__device__ inline v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
// tanh-form GELU, 0.5 x (1 + tanh(u)) with u = c (x + 0.044715 x^3), written as
// x * sigmoid(2u) (the same function: 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u))) on the
// hardware exponential
// sigmoid(2u) = 1 / (1 + 2^(-2 u log2(e))) on v_exp_f32 / v_rcp_f32 (about 1 ulp each)
constexpr float kM2CLog2e = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
__device__ inline float sig2u(float x) {
  const float t = __builtin_fmaf(0.044715f * x * x, x, x);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(kM2CLog2e * t));
}
__device__ inline float gelu(float x) { return x * sig2u(x); }
// d/dx [x sigmoid(2u)] = sigmoid(2u) + 2 x sigmoid(2u) (1 - sigmoid(2u)) u'
__device__ inline float gelu_d(float x) {
  const float sg = sig2u(x);
  return sg + 2.0f * x * sg * (1.0f - sg) * kGeluC * (1.0f + 3.0f * 0.044715f * x * x);
}

// ---- forward: register-resident, transposed orientation --------------------------------
// Each layer computes Z^T = W^T H^T (rows = units, columns = 16 samples). The MFMA result
// D[4 (l/16) + j][l % 16] is then exactly the next layer's B operand B[k][n] (k = unit,
// n = sample), so activations never leave registers; the weights are A operands kept in
// VGPRs for the life of the wave, which walks over 16-sample tiles.
__global__ __launch_bounds__(kThreads) void k_mlp_fwd(const float* __restrict__ X, int64_t n,
                                                      const float* __restrict__ P, float* __restrict__ logits) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  v4s a1[4], a2[4][4], a3[4];
  float b1r[4][4], b2r[4][4];
  for (int t = 0; t < 4; ++t)
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * lk + j;
      a1[t][j] = k < kF ? bf16(P[OW1 + k * kH + 16 * t + li]) : (short)0;
      b1r[t][j] = P[OB1 + 16 * t + k];
      b2r[t][j] = P[OB2 + 16 * t + k];
      a3[t][j] = li == 0 ? bf16(P[OW3 + 16 * t + k]) : (short)0;
      for (int q = 0; q < 4; ++q) a2[t][q][j] = bf16(P[OW2 + (16 * q + k) * kH + 16 * t + li]);
    }
  const float b3 = P[OB3];
  const int64_t ntiles = (n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
    const int64_t sidx = tile * 16 + li;
    v4s xb{0, 0, 0, 0};
    if (lk < 2 && sidx < n) {
      const float4 v = *reinterpret_cast<const float4*>(X + sidx * kF + 4 * lk);
      xb = bf16x4(v4f{v.x, v.y, v.z, v.w});
    }
    v4s h1[4], h2[4];
    for (int t = 0; t < 4; ++t) {
      v4f d = mfma(a1[t], xb, v4f{0, 0, 0, 0});
      for (int j = 0; j < 4; ++j) d[j] = gelu(d[j] + b1r[t][j]);
      h1[t] = bf16x4(d);
    }
    for (int t = 0; t < 4; ++t) {
      v4f d{0, 0, 0, 0};
      for (int q = 0; q < 4; ++q) d = mfma(a2[t][q], h1[q], d);
      for (int j = 0; j < 4; ++j) d[j] = gelu(d[j] + b2r[t][j]);
      h2[t] = bf16x4(d);
    }
    v4f d3{0, 0, 0, 0};
    for (int q = 0; q < 4; ++q) d3 = mfma(a3[q], h2[q], d3);
    if (lk == 0 && sidx < n) logits[sidx] = d3[0] + b3;
  }
}

// ---- training: forward and backward of 64-sample groups ---------------------------------
struct Lds {
  float w1[kF][kH];
  float w2[kH][kH];
  float b1[kH], b2[kH], w3[kH];
  float b3;
  float x[4][16][kF];
  float z1[4][16][kH];  // pre-activation, later the hidden-1 gradient
  float h1[4][16][kH];
  float z2[4][16][kH];  // pre-activation, later the hidden-2 gradient
  float h2[4][16][kH];
  float d3[4][16];
  float acc[kH][kH];    // one group's weight gradient (W2, then W1)
  float vec[4][kH];     // per-wave vector gradients
  float lred[kThreads];
};

// One workgroup = 4 waves; a wave owns 16 samples of the current 64-sample group. The
// workgroup walks over groups blockIdx.x, blockIdx.x + gridDim.x, ... and keeps the
// gradient sums in registers (thread tid owns W2 entries tid + 256 r, W1 entries
// tid + 256 r, and with tid < 64 entry tid of b1, b2 and w3), added group by group in
// the same order on every run. It writes its gradient sum (unscaled) to
// grad_part[blockIdx.x][NP] and its summed loss to loss_part[blockIdx.x].
__global__ __launch_bounds__(kThreads) void k_mlp_train(const float* __restrict__ X, const float* __restrict__ Y,
                                                        int64_t n, const float* __restrict__ P,
                                                        float* __restrict__ grad_part, float* __restrict__ loss_part) {
  __shared__ Lds s;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;  // MFMA lane coordinates
  for (int i = tid; i < kF * kH; i += kThreads) s.w1[i / kH][i % kH] = P[OW1 + i];
  for (int i = tid; i < kH * kH; i += kThreads) s.w2[i / kH][i % kH] = P[OW2 + i];
  if (tid < kH) {
    s.b1[tid] = P[OB1 + tid];
    s.b2[tid] = P[OB2 + tid];
    s.w3[tid] = P[OW3 + tid];
  }
  if (tid == 0) s.b3 = P[OB3];
  float acc_w2[kH * kH / kThreads] = {};
  float acc_w1[kF * kH / kThreads] = {};
  float acc_b1 = 0.0f, acc_b2 = 0.0f, acc_w3 = 0.0f, acc_b3 = 0.0f, lsum = 0.0f;
  const int64_t ngroups = (n + kSamples - 1) / kSamples;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    __syncthreads();  // the previous group's reads of x, z, h, acc and vec are done
    const int64_t s0 = grp * kSamples + wave * 16;
    for (int i = lane; i < 16 * kF; i += 64) {
      const int64_t m = s0 + i / kF;
      s.x[wave][i / kF][i % kF] = m < n ? X[m * kF + i % kF] : 0.0f;
    }
    __syncthreads();
    // ---- layer 1: Z1 = X W1 + b1 (K = 8 padded to 16) ----
    v4s ax;
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * lk + j;
      ax[j] = k < kF ? bf16(s.x[wave][li][k]) : (short)0;
    }
    for (int t = 0; t < 4; ++t) {
      v4s bw;
      for (int j = 0; j < 4; ++j) {
        const int k = 4 * lk + j;
        bw[j] = k < kF ? bf16(s.w1[k][16 * t + li]) : (short)0;
      }
      const v4f d = mfma(ax, bw, v4f{0, 0, 0, 0});
      for (int j = 0; j < 4; ++j) {
        const int m = 4 * lk + j, c = 16 * t + li;
        const float z = d[j] + s.b1[c];
        s.z1[wave][m][c] = z;
        s.h1[wave][m][c] = gelu(z);
      }
    }
    __syncthreads();
    // ---- layer 2: Z2 = H1 W2 + b2 ----
    for (int t = 0; t < 4; ++t) {
      v4f d{0, 0, 0, 0};
      for (int q = 0; q < 4; ++q) {
        v4s a, b;
        for (int j = 0; j < 4; ++j) {
          a[j] = bf16(s.h1[wave][li][16 * q + 4 * lk + j]);
          b[j] = bf16(s.w2[16 * q + 4 * lk + j][16 * t + li]);
        }
        d = mfma(a, b, d);
      }
      for (int j = 0; j < 4; ++j) {
        const int m = 4 * lk + j, c = 16 * t + li;
        const float z = d[j] + s.b2[c];
        s.z2[wave][m][c] = z;
        s.h2[wave][m][c] = gelu(z);
      }
    }
    __syncthreads();
    // ---- layer 3: z3 = H2 w3 + b3 (N = 1 padded to 16: column 0 carries w3) ----
    v4f d3{0, 0, 0, 0};
    for (int q = 0; q < 4; ++q) {
      v4s a, b;
      for (int j = 0; j < 4; ++j) {
        a[j] = bf16(s.h2[wave][li][16 * q + 4 * lk + j]);
        b[j] = li == 0 ? bf16(s.w3[16 * q + 4 * lk + j]) : (short)0;
      }
      d3 = mfma(a, b, d3);
    }
    if (li == 0) {
      for (int j = 0; j < 4; ++j) {
        const int m = 4 * lk + j;
        const int64_t g = s0 + m;
        const float z = d3[j] + s.b3;
        float dz = 0.0f;
        if (g < n) {
          const float y = Y[g];
          dz = 1.0f / (1.0f + expf(-z)) - y;
          lsum += fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
        }
        s.d3[wave][m] = dz;
      }
    }
    __syncthreads();
    // ---- output layer gradients: w3, b3 ----
    {
      float g = 0.0f;
      for (int m = 0; m < 16; ++m) g += s.h2[wave][m][lane] * s.d3[wave][m];
      s.vec[wave][lane] = g;
    }
    __syncthreads();
    if (tid < kH) acc_w3 += s.vec[0][tid] + s.vec[1][tid] + s.vec[2][tid] + s.vec[3][tid];
    if (tid == 0) {
      float g = 0.0f;
      for (int w = 0; w < 4; ++w)
        for (int m = 0; m < 16; ++m) g += s.d3[w][m];
      acc_b3 += g;
    }
    // ---- hidden-2 gradient G2 = (d3 w3^T) * gelu'(Z2), kept in z2 ----
    for (int i = lane; i < 16 * kH; i += 64) {
      const int m = i / kH, c = i % kH;
      s.z2[wave][m][c] = s.d3[wave][m] * s.w3[c] * gelu_d(s.z2[wave][m][c]);
    }
    for (int i = tid; i < kH * kH; i += kThreads) s.acc[i / kH][i % kH] = 0.0f;
    __syncthreads();
    // dW2 = H1^T G2 (M = hidden-1 unit, K = samples, N = hidden-2 unit), waves added in order
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
        for (int tm = 0; tm < 4; ++tm)
          for (int tn = 0; tn < 4; ++tn) {
            v4s a, b;
            for (int j = 0; j < 4; ++j) {
              a[j] = bf16(s.h1[w][4 * lk + j][16 * tm + li]);
              b[j] = bf16(s.z2[w][4 * lk + j][16 * tn + li]);
            }
            const v4f d = mfma(a, b, v4f{0, 0, 0, 0});
            for (int j = 0; j < 4; ++j) s.acc[16 * tm + 4 * lk + j][16 * tn + li] += d[j];
          }
      }
      __syncthreads();
    }
    for (int r = 0; r < kH * kH / kThreads; ++r) {
      const int i = tid + kThreads * r;
      acc_w2[r] += s.acc[i / kH][i % kH];
    }
    {
      float g = 0.0f;
      for (int m = 0; m < 16; ++m) g += s.z2[wave][m][lane];
      s.vec[wave][lane] = g;
    }
    __syncthreads();
    if (tid < kH) acc_b2 += s.vec[0][tid] + s.vec[1][tid] + s.vec[2][tid] + s.vec[3][tid];
    // ---- hidden-1 gradient G1 = (G2 W2^T) * gelu'(Z1), kept in z1 ----
    for (int t = 0; t < 4; ++t) {
      v4f d{0, 0, 0, 0};
      for (int q = 0; q < 4; ++q) {
        v4s a, b;
        for (int j = 0; j < 4; ++j) {
          a[j] = bf16(s.z2[wave][li][16 * q + 4 * lk + j]);
          b[j] = bf16(s.w2[16 * t + li][16 * q + 4 * lk + j]);  // W2^T
        }
        d = mfma(a, b, d);
      }
      for (int j = 0; j < 4; ++j) {
        const int m = 4 * lk + j, c = 16 * t + li;
        s.z1[wave][m][c] = d[j] * gelu_d(s.z1[wave][m][c]);
      }
    }
    __syncthreads();  // every thread has read acc (W2) and vec; z1 holds G1
    // dW1 = X^T G1 (M = input feature, padded to 16), waves added in order
    for (int i = tid; i < kF * kH; i += kThreads) s.acc[i / kH][i % kH] = 0.0f;
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
        for (int tn = 0; tn < 4; ++tn) {
          v4s a, b;
          for (int j = 0; j < 4; ++j) {
            a[j] = li < kF ? bf16(s.x[w][4 * lk + j][li]) : (short)0;
            b[j] = bf16(s.z1[w][4 * lk + j][16 * tn + li]);
          }
          const v4f d = mfma(a, b, v4f{0, 0, 0, 0});
          for (int j = 0; j < 4; ++j) {
            const int r = 4 * lk + j;
            if (r < kF) s.acc[r][16 * tn + li] += d[j];
          }
        }
      }
      __syncthreads();
    }
    for (int r = 0; r < kF * kH / kThreads; ++r) {
      const int i = tid + kThreads * r;
      acc_w1[r] += s.acc[i / kH][i % kH];
    }
    {
      float g = 0.0f;
      for (int m = 0; m < 16; ++m) g += s.z1[wave][m][lane];
      s.vec[wave][lane] = g;
    }
    __syncthreads();
    if (tid < kH) acc_b1 += s.vec[0][tid] + s.vec[1][tid] + s.vec[2][tid] + s.vec[3][tid];
  }
  float* gp = grad_part + (int64_t)blockIdx.x * kNP;
  for (int r = 0; r < kH * kH / kThreads; ++r) gp[OW2 + tid + kThreads * r] = acc_w2[r];
  for (int r = 0; r < kF * kH / kThreads; ++r) gp[OW1 + tid + kThreads * r] = acc_w1[r];
  if (tid < kH) {
    gp[OB1 + tid] = acc_b1;
    gp[OB2 + tid] = acc_b2;
    gp[OW3 + tid] = acc_w3;
  }
  if (tid == 0) gp[OB3] = acc_b3;
  // the loss of the workgroup: lanes 0, 16, 32, 48 of every wave hold partial sums
  __syncthreads();
  s.lred[tid] = lsum;
  __syncthreads();
  if (tid == 0) {
    float l = 0.0f;
    for (int i = 0; i < kThreads; ++i) l += s.lred[i];
    loss_part[blockIdx.x] = l;
  }
}

// grad[p] = sum_b part[b][p] / count (fixed order); loss likewise
__global__ void k_mlp_fold(int nblocks, const float* part, const float* loss_part, float inv_count, float* grad,
                           float* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNP) {
    float g = 0.0f;
    for (int b = 0; b < nblocks; ++b) g += part[(int64_t)b * kNP + p];
    grad[p] = g * inv_count;
  }
  if (p == 0) {
    float l = 0.0f;
    for (int b = 0; b < nblocks; ++b) l += loss_part[b];
    *loss = l * inv_count;
  }
}

__global__ void k_adam(float* P, float* M, float* V, const float* G, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.0f - b1) * g;
  const float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_gather_rows(const float* X, const float* Y, const int32_t* idx, int64_t n, float* Xo, float* Yo) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = idx[i];
  for (int f = 0; f < kF; ++f) Xo[i * kF + f] = X[j * kF + f];
  Yo[i] = Y[j];
}

}  // namespace

void mlp_forward_device(nc_mlp_s* M, int64_t n, const float* d_x, float* d_logits, hipStream_t s) {
  if (n <= 0) return;
  k_mlp_fwd<<<fwd_grid(n), kThreads, 0, s>>>(d_x, n, M->params.p, d_logits);
  NC_CHECK_HIP(hipGetLastError());
}

// Gradient of the mean loss over n device samples into M->grad (and the loss on the host)
double mlp_grad_device(nc_mlp_s* M, int64_t n, const float* d_x, const float* d_y, hipStream_t s) {
  const int nb = (int)train_grid(n);
  M->part.reserve((size_t)std::max(nb, 1) * kNP);
  M->loss_part.reserve(std::max(nb, 1));
  k_mlp_train<<<nb, kThreads, 0, s>>>(d_x, d_y, n, M->params.p, M->part.p, M->loss_part.p);
  k_mlp_fold<<<grid_for(kNP, 256), 256, 0, s>>>(nb, M->part.p, M->loss_part.p, 1.0f / (float)n, M->grad.p,
                                               M->loss.p);
  NC_CHECK_HIP(hipGetLastError());
  float l = 0.0f;
  NC_CHECK_HIP(hipMemcpyAsync(&l, M->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  NC_CHECK_HIP(hipStreamSynchronize(s));
  return l;
}

}  // namespace nc

using namespace nc;

extern "C" {

int nc_mlp_create(uint64_t seed, nc_mlp* out) {
  return guard([&] {
    NC_REQUIRE(out, NC_ERR_ARG, "null argument");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
      throw Error(NC_ERR_NO_DEVICE, "no HIP device available");
    auto M = std::make_unique<nc_mlp_s>();
    std::vector<float> p(kNP, 0.0f);
    std::mt19937_64 rng(seed);
    auto init = [&](int off, int fan_in, int fan_out) {
      const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));  // Glorot uniform
      std::uniform_real_distribution<float> u(-a, a);
      for (int i = 0; i < fan_in * fan_out; ++i) p[off + i] = u(rng);
    };
    init(OW1, kF, kH);
    init(OW2, kH, kH);
    init(OW3, kH, 1);
    M->params.alloc(kNP);
    M->m.alloc(kNP);
    M->v.alloc(kNP);
    M->grad.alloc(kNP);
    M->loss.alloc(1);
    hipStream_t s = M->stream.s;
    NC_CHECK_HIP(hipMemcpyAsync(M->params.p, p.data(), sizeof(float) * kNP, hipMemcpyHostToDevice, s));
    NC_CHECK_HIP(hipMemsetAsync(M->m.p, 0, sizeof(float) * kNP, s));
    NC_CHECK_HIP(hipMemsetAsync(M->v.p, 0, sizeof(float) * kNP, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    *out = M.release();
  });
}

int nc_mlp_destroy(nc_mlp M) {
  return guard([&] { delete M; });
}

int nc_mlp_get_params(nc_mlp M, float* out) {
  return guard([&] {
    NC_REQUIRE(M && out, NC_ERR_ARG, "null argument");
    NC_CHECK_HIP(hipMemcpy(out, M->params.p, sizeof(float) * kNP, hipMemcpyDeviceToHost));
  });
}

int nc_mlp_set_params(nc_mlp M, const float* in) {
  return guard([&] {
    NC_REQUIRE(M && in, NC_ERR_ARG, "null argument");
    NC_CHECK_HIP(hipMemcpy(M->params.p, in, sizeof(float) * kNP, hipMemcpyHostToDevice));
    NC_CHECK_HIP(hipMemset(M->m.p, 0, sizeof(float) * kNP));
    NC_CHECK_HIP(hipMemset(M->v.p, 0, sizeof(float) * kNP));
    M->step = 0;
  });
}

int nc_mlp_forward(nc_mlp M, int64_t n, const float* x, float* logits) {
  return guard([&] {
    NC_REQUIRE(M && n >= 0 && (n == 0 || (x && logits)), NC_ERR_ARG, "bad arguments");
    if (n == 0) return;
    hipStream_t s = M->stream.s;
    DevBuf<float> dx((size_t)n * kF), dl((size_t)n);
    NC_CHECK_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kF, hipMemcpyHostToDevice, s));
    mlp_forward_device(M, n, dx.p, dl.p, s);
    NC_CHECK_HIP(hipMemcpyAsync(logits, dl.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

int nc_mlp_grad(nc_mlp M, int64_t n, const float* x, const float* y, float* grad, double* loss) {
  return guard([&] {
    NC_REQUIRE(M && n >= 1 && x && y && grad, NC_ERR_ARG, "bad arguments");
    hipStream_t s = M->stream.s;
    DevBuf<float> dx((size_t)n * kF), dy((size_t)n);
    NC_CHECK_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kF, hipMemcpyHostToDevice, s));
    NC_CHECK_HIP(hipMemcpyAsync(dy.p, y, sizeof(float) * n, hipMemcpyHostToDevice, s));
    const double l = mlp_grad_device(M, n, dx.p, dy.p, s);
    NC_CHECK_HIP(hipMemcpyAsync(grad, M->grad.p, sizeof(float) * kNP, hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
  });
}

int nc_mlp_train(nc_mlp M, int64_t n, const float* x, const float* y, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, double* final_loss) {
  return guard([&] {
    NC_REQUIRE(M && n >= 1 && x && y && epochs >= 1 && batch >= 1, NC_ERR_ARG, "bad arguments");
    hipStream_t s = M->stream.s;
    DevBuf<float> dx((size_t)n * kF), dy((size_t)n), bx((size_t)batch * kF), by((size_t)batch);
    DevBuf<int32_t> didx((size_t)n);
    NC_CHECK_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kF, hipMemcpyHostToDevice, s));
    NC_CHECK_HIP(hipMemcpyAsync(dy.p, y, sizeof(float) * n, hipMemcpyHostToDevice, s));
    std::vector<int32_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double loss = 0.0;
    for (int e = 0; e < epochs; ++e) {
      std::shuffle(idx.begin(), idx.end(), rng);
      NC_CHECK_HIP(hipMemcpyAsync(didx.p, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
      double esum = 0.0;
      int64_t ecount = 0;
      for (int64_t b0 = 0; b0 < n; b0 += batch) {
        const int64_t nb = std::min<int64_t>(batch, n - b0);
        k_gather_rows<<<grid_for(nb, 256), 256, 0, s>>>(dx.p, dy.p, didx.p + b0, nb, bx.p, by.p);
        const int blocks = (int)train_grid(nb);
        M->part.reserve((size_t)blocks * kNP);
        M->loss_part.reserve(blocks);
        k_mlp_train<<<blocks, kThreads, 0, s>>>(bx.p, by.p, nb, M->params.p, M->part.p, M->loss_part.p);
        k_mlp_fold<<<grid_for(kNP, 256), 256, 0, s>>>(blocks, M->part.p, M->loss_part.p, 1.0f / (float)nb, M->grad.p,
                                                     M->loss.p);
        ++M->step;
        const float c1 = 1.0f - std::pow(b1, (float)M->step), c2 = 1.0f - std::pow(b2, (float)M->step);
        k_adam<<<grid_for(kNP, 256), 256, 0, s>>>(M->params.p, M->m.p, M->v.p, M->grad.p, lr, b1, b2, eps, c1, c2);
        NC_CHECK_HIP(hipGetLastError());
        if (e == epochs - 1) {
          float l = 0.0f;
          NC_CHECK_HIP(hipMemcpyAsync(&l, M->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
          NC_CHECK_HIP(hipStreamSynchronize(s));
          esum += (double)l * nb;
          ecount += nb;
        }
      }
      if (e == epochs - 1) loss = esum / (double)std::max<int64_t>(ecount, 1);
    }
    NC_CHECK_HIP(hipStreamSynchronize(s));
    if (final_loss) *final_loss = loss;
  });
}

int nc_mlp_bench(nc_mlp M, int64_t n, int32_t reps, int32_t train, double* ms_per_launch) {
  return guard([&] {
    NC_REQUIRE(M && n >= 1 && reps >= 1 && ms_per_launch, NC_ERR_ARG, "bad arguments");
    hipStream_t s = M->stream.s;
    // features uniform in [-1, 1), labels 0/1 (a fixed 64-bit LCG, so runs are repeatable)
    std::vector<float> hx((size_t)n * kF), hy((size_t)n);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto next = [&] {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      return (float)((st >> 40) * (1.0 / 16777216.0));
    };
    for (auto& v : hx) v = 2.0f * next() - 1.0f;
    for (auto& v : hy) v = next() < 0.3f ? 1.0f : 0.0f;
    DevBuf<float> dx(hx.size()), dy(hy.size()), dl((size_t)n);
    NC_CHECK_HIP(hipMemcpyAsync(dx.p, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice, s));
    NC_CHECK_HIP(hipMemcpyAsync(dy.p, hy.data(), sizeof(float) * hy.size(), hipMemcpyHostToDevice, s));
    const unsigned g = train ? train_grid(n) : fwd_grid(n);
    if (train) {
      M->part.reserve((size_t)g * kNP);
      M->loss_part.reserve(g);
    }
    auto launch = [&] {
      if (train)
        k_mlp_train<<<g, kThreads, 0, s>>>(dx.p, dy.p, n, M->params.p, M->part.p, M->loss_part.p);
      else
        k_mlp_fwd<<<g, kThreads, 0, s>>>(dx.p, n, M->params.p, dl.p);
    };
    launch();  // warm-up
    hipEvent_t e0, e1;
    NC_CHECK_HIP(hipEventCreate(&e0));
    NC_CHECK_HIP(hipEventCreate(&e1));
    NC_CHECK_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch();
    NC_CHECK_HIP(hipEventRecord(e1, s));
    NC_CHECK_HIP(hipGetLastError());
    NC_CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    NC_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_launch = ms / reps;
  });
}

}  // extern "C"
