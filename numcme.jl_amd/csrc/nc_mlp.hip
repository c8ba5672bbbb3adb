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
// One workgroup = 4 waves = 64 samples; a wave owns 16 samples. For
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
constexpr int kSamples = 64;  // per workgroup

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline short bf16(float x) {
  // round to nearest even (inputs here are finite)
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ inline v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
__device__ inline float gelu(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}
__device__ inline float gelu_d(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  const float th = tanhf(u);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * kGeluC * (1.0f + 3.0f * 0.044715f * x * x);
}

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
  float acc[kH][kH];    // gradient accumulator of the workgroup (W2, then W1)
  float vec[4][kH];     // per-wave vector gradients
};

// Forward (and with TRAIN the backward) pass of 64 samples per workgroup. TRAIN writes the
// workgroup's gradient (sum over its samples, unscaled) to grad_part[blockIdx.x][NP] and
// its summed loss to loss_part[blockIdx.x].
template <bool TRAIN>
__global__ __launch_bounds__(kThreads) void k_mlp(const float* __restrict__ X, const float* __restrict__ Y, int64_t n,
                                                  const float* __restrict__ P, float* __restrict__ logits,
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
  const int64_t s0 = (int64_t)blockIdx.x * kSamples + wave * 16;
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
    v4f d = mfma(ax, bw, v4f{0, 0, 0, 0});
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
  float lsum = 0.0f;
  if (li == 0) {
    for (int j = 0; j < 4; ++j) {
      const int m = 4 * lk + j;
      const int64_t g = s0 + m;
      const float z = d3[j] + s.b3;
      if (g < n && logits) logits[g] = z;
      if (TRAIN) {
        float dz = 0.0f;
        if (g < n) {
          const float y = Y[g];
          dz = 1.0f / (1.0f + expf(-z)) - y;
          lsum += fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
        }
        s.d3[wave][m] = dz;
      }
    }
  }
  if (!TRAIN) return;
  __syncthreads();
  float* gp = grad_part + (int64_t)blockIdx.x * kNP;
  // ---- output layer gradients: w3, b3 (plain sums, fixed order) ----
  {
    float g = 0.0f;
    for (int m = 0; m < 16; ++m) g += s.h2[wave][m][lane] * s.d3[wave][m];
    s.vec[wave][lane] = g;
  }
  __syncthreads();
  if (tid < kH) gp[OW3 + tid] = s.vec[0][tid] + s.vec[1][tid] + s.vec[2][tid] + s.vec[3][tid];
  if (tid == 0) {
    float g = 0.0f, l = 0.0f;
    for (int w = 0; w < 4; ++w)
      for (int m = 0; m < 16; ++m) g += s.d3[w][m];
    gp[OB3] = g;
    (void)l;
  }
  // ---- hidden-2 gradient G2 = (d3 w3^T) * gelu'(Z2), kept in z2 ----
  for (int i = lane; i < 16 * kH; i += 64) {
    const int m = i / kH, c = i % kH;
    s.z2[wave][m][c] = s.d3[wave][m] * s.w3[c] * gelu_d(s.z2[wave][m][c]);
  }
  __syncthreads();
  // dW2 = H1^T G2 (M = hidden-1 unit, K = samples, N = hidden-2 unit), waves added in order
  for (int i = tid; i < kH * kH; i += kThreads) s.acc[i / kH][i % kH] = 0.0f;
  __syncthreads();
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
  for (int i = tid; i < kH * kH; i += kThreads) gp[OW2 + i] = s.acc[i / kH][i % kH];
  {
    float g = 0.0f;
    for (int m = 0; m < 16; ++m) g += s.z2[wave][m][lane];
    s.vec[wave][lane] = g;
  }
  __syncthreads();
  if (tid < kH) gp[OB2 + tid] = s.vec[0][tid] + s.vec[1][tid] + s.vec[2][tid] + s.vec[3][tid];
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
  __syncthreads();
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
  for (int i = tid; i < kF * kH; i += kThreads) gp[OW1 + i] = s.acc[i / kH][i % kH];
  {
    float g = 0.0f;
    for (int m = 0; m < 16; ++m) g += s.z1[wave][m][lane];
    s.vec[wave][lane] = g;
  }
  // the loss of the workgroup: lanes 0, 16, 32, 48 of every wave hold partial sums
  __shared__ float lred[kThreads];
  lred[tid] = lsum;
  __syncthreads();
  if (tid < kH) gp[OB1 + tid] = s.vec[0][tid] + s.vec[1][tid] + s.vec[2][tid] + s.vec[3][tid];
  if (tid == 0) {
    float l = 0.0f;
    for (int i = 0; i < kThreads; ++i) l += lred[i];
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
  const unsigned g = (unsigned)((n + kSamples - 1) / kSamples);
  k_mlp<false><<<g, kThreads, 0, s>>>(d_x, nullptr, n, M->params.p, d_logits, nullptr, nullptr);
  NC_CHECK_HIP(hipGetLastError());
}

// Gradient of the mean loss over n device samples into M->grad (and the loss on the host)
double mlp_grad_device(nc_mlp_s* M, int64_t n, const float* d_x, const float* d_y, hipStream_t s) {
  const int nb = (int)((n + kSamples - 1) / kSamples);
  M->part.reserve((size_t)std::max(nb, 1) * kNP);
  M->loss_part.reserve(std::max(nb, 1));
  k_mlp<true><<<nb, kThreads, 0, s>>>(d_x, d_y, n, M->params.p, nullptr, M->part.p, M->loss_part.p);
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
        const int blocks = (int)((nb + kSamples - 1) / kSamples);
        M->part.reserve((size_t)blocks * kNP);
        M->loss_part.reserve(blocks);
        k_mlp<true><<<blocks, kThreads, 0, s>>>(bx.p, by.p, nb, M->params.p, nullptr, M->part.p, M->loss_part.p);
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
    const unsigned g = (unsigned)((n + kSamples - 1) / kSamples);
    if (train) {
      M->part.reserve((size_t)g * kNP);
      M->loss_part.reserve(g);
    }
    auto launch = [&] {
      if (train)
        k_mlp<true><<<g, kThreads, 0, s>>>(dx.p, dy.p, n, M->params.p, nullptr, M->part.p, M->loss_part.p);
      else
        k_mlp<false><<<g, kThreads, 0, s>>>(dx.p, nullptr, n, M->params.p, dl.p, nullptr, nullptr);
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
