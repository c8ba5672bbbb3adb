// Learned space adapter network: an 8 -> 64 -> 64 -> 1 MLP with tanh-form GELU hidden
// layers and a keep/add logit, computed with BF16 MFMA and FP32 accumulation: gfx950's
// v_mfma_f32_16x16x32_bf16 for every product with a contraction of 32 or more, and
// v_mfma_f32_16x16x16_bf16 for the first layer (8 input features, padded to 16). Training
// (binary cross-entropy on the logit, Adam on FP32 master weights) runs on the device with
// the forward and backward kernels below. No reference counterpart: NumCME.jl's adapters
// are rule-based.
//
// Parameters (FP32 master copy, NP = 4801 floats): W1 [8][64] (input-major), b1 [64],
// W2 [64][64] (input-major), b2 [64], w3 [64], b3. Every MFMA operand is rounded to BF16
// (round to nearest even) when it is loaded into registers.
//
// Lane layouts (lane l, g = l / 16, i = l % 16):
//   16x16x16: A[i][4 g + j], B[4 g + j][i], j = 0..3;
//   16x16x32: A[i][k(g, j)], B[k(g, j)][i], j = 0..7, with the same K slot k(g, j) in A and
//             B (which K index a slot stands for is the program's choice: a contraction
//             only needs A and B to agree);
//   both:     D[4 g + j][i], j = 0..3.
// Activations are kept transposed (units x samples): a layer's D tile, lane (g, sample i)
// holding units 16 t + 4 g + j, is the next layer's B operand with no data movement, the
// K slots of a 32-unit chunk q being k(g, j) = 32 q + 16 (j / 4) + 4 g + j % 4, i.e. the
// D tiles 2q and 2q+1 side by side.
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
constexpr unsigned kFwdMaxBlocks = 4096, kTrainMaxBlocks = 512;

// forward: 4 waves per workgroup, 16 samples per wave and tile
unsigned fwd_grid(int64_t n) {
  const int64_t tiles = (n + 15) / 16;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, kFwdMaxBlocks));
}
// training: 32-sample chunks, one per wave and iteration, 4 waves per workgroup
unsigned train_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 127) / 128, kTrainMaxBlocks));
}

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// round to nearest even with the hardware conversion (v_cvt_pk_bf16_f32)
__device__ inline short bf16(float x) { return __builtin_bit_cast(short, (__bf16)x); }
__device__ inline v4s bf16x4(v4f x) { return __builtin_bit_cast(v4s, __builtin_convertvector(x, v4bf)); }

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ inline v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ inline v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}
// the K = 32 operand of two 16-row D tiles (see the lane layouts above)
__device__ inline v8s cat(v4s lo, v4s hi) { return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7); }
// K unit of slot (g, j) of chunk q
__device__ inline int kslot(int q, int g, int j) { return 32 * q + 16 * (j >> 2) + 4 * g + (j & 3); }

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
// tanh-form GELU, 0.5 x (1 + tanh(u)) with u = c (x + 0.044715 x^3), written as
// x * sigmoid(2u) (the same function: 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u))) on the
// hardware exponential and reciprocal (about 1 ulp each):
// sigmoid(2u) = 1 / (1 + 2^(x (k + 0.044715 k x^2))), k = -2 c log2(e)
constexpr float kM2CLog2e = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
__device__ inline float sig2u(float x) {
  const float e = x * __builtin_fmaf(0.044715f * kM2CLog2e * x, x, kM2CLog2e);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(e));
}
__device__ inline float gelu(float x) { return x * sig2u(x); }
// d/dx [x sigmoid(2u)] = sigmoid(2u) + 2 x sigmoid(2u) (1 - sigmoid(2u)) u', from the
// forward pass's sigmoid sg
__device__ inline float gelu_d(float x, float sg) {
  return sg + 2.0f * x * sg * (1.0f - sg) * kGeluC * __builtin_fmaf(3.0f * 0.044715f * x, x, 1.0f);
}

// ---- forward: register-resident, transposed orientation --------------------------------
// Each layer computes Z^T = W^T H^T (rows = units, columns = 16 samples); the weights are
// A operands kept in VGPRs for the life of the wave, which walks over 16-sample tiles.
__global__ __launch_bounds__(kThreads) void k_mlp_fwd(const float* __restrict__ X, int64_t n,
                                                      const float* __restrict__ P, float* __restrict__ logits) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  v4s a1[4];
  v8s a2[4][2], a3[2];
  float b1r[4][4], b2r[4][4];
  for (int t = 0; t < 4; ++t)
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * lk + j;
      a1[t][j] = k < kF ? bf16(P[OW1 + k * kH + 16 * t + li]) : (short)0;
      b1r[t][j] = P[OB1 + 16 * t + k];
      b2r[t][j] = P[OB2 + 16 * t + k];
    }
  for (int q = 0; q < 2; ++q)
    for (int j = 0; j < 8; ++j) {
      const int k = kslot(q, lk, j);
      a3[q][j] = li == 0 ? bf16(P[OW3 + k]) : (short)0;
      for (int t = 0; t < 4; ++t) a2[t][q][j] = bf16(P[OW2 + k * kH + 16 * t + li]);
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
      for (int q = 0; q < 2; ++q) d = mfma32(a2[t][q], cat(h1[2 * q], h1[2 * q + 1]), d);
      for (int j = 0; j < 4; ++j) d[j] = gelu(d[j] + b2r[t][j]);
      h2[t] = bf16x4(d);
    }
    v4f d3{0, 0, 0, 0};
    for (int q = 0; q < 2; ++q) d3 = mfma32(a3[q], cat(h2[2 * q], h2[2 * q + 1]), d3);
    if (lk == 0 && sidx < n) logits[sidx] = d3[0] + b3;
  }
}

// ---- training: forward and backward, register-resident -------------------------------
// A wave takes 32-sample chunks (chunk w, w + waves, ... of the grid's waves, in order).
// For each 16-sample half it runs the forward pass in the transposed orientation above,
// the loss gradient dz, the output-layer and hidden-2 gradients G2^T = dz w3 * gelu'(Z2^T)
// and the hidden-1 gradient G1^T = (W2 G2^T) * gelu'(Z1^T) (W2 as the A operand, G2^T's D
// tiles as B), all in registers. The weight gradients contract over samples, which the D
// tiles hold across lanes: H1, G2 and G1 are written to the wave's own LDS rows
// [unit][sample] (BF16) and read back with lane = unit, K slots = the chunk's 32 samples,
// so that dW2 = H1^T G2 and dW1 = X^T G1 are one 16x16x32 MFMA per output tile and chunk,
// accumulated in registers. Only the wave's own lanes exchange data through its LDS rows,
// so no workgroup barrier runs inside the loop. At the end the four waves' sums go
// through LDS and are added in wave order into the workgroup's partial
// grad_part[blockIdx.x][NP] (unscaled), its summed loss into loss_part[blockIdx.x].
constexpr int kTP = 36;                    // LDS row pitch (BF16) of a transposed activation
constexpr int kTWave = 3 * kH * kTP;       // H1, G2, G1 rows of one wave
constexpr int kRed = kNP + 4;              // floats of one wave's partial sums (+ loss)
constexpr size_t kTrainLds = std::max<size_t>(4 * kTWave * sizeof(short), 4 * kRed * sizeof(float));

// orders a wave's LDS stores before its own lanes' later loads of them
__device__ inline void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kThreads) void k_mlp_train(const float* __restrict__ X, const float* __restrict__ Y,
                                                        int64_t n, const float* __restrict__ P,
                                                        float* __restrict__ grad_part, float* __restrict__ loss_part) {
  __shared__ __align__(16) unsigned char lds[kTrainLds];
  __shared__ __align__(16) float s_b1[kH], s_b2[kH], s_w3[kH];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  short* tH1 = reinterpret_cast<short*>(lds) + wave * kTWave;
  short* tG2 = tH1 + kH * kTP;
  short* tG1 = tG2 + kH * kTP;
  if (tid < kH) {
    s_b1[tid] = P[OB1 + tid];
    s_b2[tid] = P[OB2 + tid];
    s_w3[tid] = P[OW3 + tid];
  }
  // weights as A operands: W1^T (16x16x16), W2^T (forward), W2 (G1), the w3 row
  v4s a1[4];
  v8s a2[4][2], a2b[4][2], a3[2];
  for (int t = 0; t < 4; ++t)
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * lk + j;
      a1[t][j] = k < kF ? bf16(P[OW1 + k * kH + 16 * t + li]) : (short)0;
    }
  for (int q = 0; q < 2; ++q)
    for (int j = 0; j < 8; ++j) {
      const int k = kslot(q, lk, j);
      a3[q][j] = li == 0 ? bf16(P[OW3 + k]) : (short)0;
      for (int t = 0; t < 4; ++t) {
        a2[t][q][j] = bf16(P[OW2 + k * kH + 16 * t + li]);   // A[c][k] = W2[k][c]
        a2b[t][q][j] = bf16(P[OW2 + (16 * t + li) * kH + k]);  // A[k'][c] = W2[k'][c]
      }
    }
  const float b3 = P[OB3];
  __syncthreads();
  v4f gw2[4][4], gw1[4];
  float gb1[4][4], gb2[4][4], gw3[4][4];
  for (int t = 0; t < 4; ++t) {
    gw1[t] = v4f{0, 0, 0, 0};
    for (int u = 0; u < 4; ++u) gw2[t][u] = v4f{0, 0, 0, 0};
    for (int j = 0; j < 4; ++j) gb1[t][j] = gb2[t][j] = gw3[t][j] = 0.0f;
  }
  float gb3 = 0.0f, lsum = 0.0f;
  const int64_t nchunks = (n + 31) / 32;
  for (int64_t ch = (int64_t)blockIdx.x * 4 + wave; ch < nchunks; ch += (int64_t)gridDim.x * 4) {
    for (int hf = 0; hf < 2; ++hf) {
      const int64_t sidx = ch * 32 + hf * 16 + li;
      const int col = hf * 16 + li;  // this lane's sample column in the LDS rows
      v4s xb{0, 0, 0, 0};
      if (lk < 2 && sidx < n) {
        const float4 v = *reinterpret_cast<const float4*>(X + sidx * kF + 4 * lk);
        xb = bf16x4(v4f{v.x, v.y, v.z, v.w});
      }
      // layer 1: Z1^T, kept with its GELU sigmoid; H1^T to LDS
      float z1[4][4], s1[4][4];
      v4s h1[4];
      for (int t = 0; t < 4; ++t) {
        const v4f d = mfma(a1[t], xb, v4f{0, 0, 0, 0});
        const float4 b = *reinterpret_cast<const float4*>(s_b1 + 16 * t + 4 * lk);
        const float bb[4] = {b.x, b.y, b.z, b.w};
        v4f h;
        for (int j = 0; j < 4; ++j) {
          z1[t][j] = d[j] + bb[j];
          s1[t][j] = sig2u(z1[t][j]);
          h[j] = z1[t][j] * s1[t][j];
        }
        h1[t] = bf16x4(h);
        for (int j = 0; j < 4; ++j) tH1[(16 * t + 4 * lk + j) * kTP + col] = h1[t][j];
      }
      // layer 2
      float z2[4][4], s2[4][4];
      v4s h2[4];
      for (int t = 0; t < 4; ++t) {
        v4f d{0, 0, 0, 0};
        for (int q = 0; q < 2; ++q) d = mfma32(a2[t][q], cat(h1[2 * q], h1[2 * q + 1]), d);
        const float4 b = *reinterpret_cast<const float4*>(s_b2 + 16 * t + 4 * lk);
        const float bb[4] = {b.x, b.y, b.z, b.w};
        v4f h;
        for (int j = 0; j < 4; ++j) {
          z2[t][j] = d[j] + bb[j];
          s2[t][j] = sig2u(z2[t][j]);
          h[j] = z2[t][j] * s2[t][j];
        }
        h2[t] = bf16x4(h);
      }
      // layer 3 and the loss gradient dz (row 0 of D: lanes with lk == 0, sample li)
      v4f d3{0, 0, 0, 0};
      for (int q = 0; q < 2; ++q) d3 = mfma32(a3[q], cat(h2[2 * q], h2[2 * q + 1]), d3);
      float dz = 0.0f;
      if (lk == 0 && sidx < n) {
        const float z = d3[0] + b3, y = Y[sidx];
        dz = 1.0f / (1.0f + expf(-z)) - y;
        lsum += fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
        gb3 += dz;
      }
      dz = __shfl(dz, li, 64);  // every lane: dz of its sample
      // output layer (w3 with the unrounded H2) and hidden-2 gradient; G2^T to LDS
      v4s g2[4];
      for (int t = 0; t < 4; ++t) {
        const float4 w = *reinterpret_cast<const float4*>(s_w3 + 16 * t + 4 * lk);
        const float ww[4] = {w.x, w.y, w.z, w.w};
        v4f g;
        for (int j = 0; j < 4; ++j) {
          gw3[t][j] += z2[t][j] * s2[t][j] * dz;
          g[j] = dz * ww[j] * gelu_d(z2[t][j], s2[t][j]);
          gb2[t][j] += g[j];
        }
        g2[t] = bf16x4(g);
        for (int j = 0; j < 4; ++j) tG2[(16 * t + 4 * lk + j) * kTP + col] = g2[t][j];
      }
      // hidden-1 gradient G1^T = (W2 G2^T) * gelu'(Z1^T); G1^T to LDS
      for (int t = 0; t < 4; ++t) {
        v4f d{0, 0, 0, 0};
        for (int q = 0; q < 2; ++q) d = mfma32(a2b[t][q], cat(g2[2 * q], g2[2 * q + 1]), d);
        v4f g;
        for (int j = 0; j < 4; ++j) {
          g[j] = d[j] * gelu_d(z1[t][j], s1[t][j]);
          gb1[t][j] += g[j];
        }
        const v4s gb = bf16x4(g);
        for (int j = 0; j < 4; ++j) tG1[(16 * t + 4 * lk + j) * kTP + col] = gb[j];
      }
    }
    wave_lds_fence();
    // weight gradients over the chunk: operands with lane = unit 16 tu + li and the K slots
    // k(g, j) -> sample 16 (j / 4) + 4 g + j % 4 of the chunk
    auto rows = [&](const short* T, int tu) {
      const short* r = T + (16 * tu + li) * kTP + 4 * lk;
      return cat(*reinterpret_cast<const v4s*>(r), *reinterpret_cast<const v4s*>(r + 16));
    };
    v8s hT[4], gT[4];
    for (int u = 0; u < 4; ++u) {
      hT[u] = rows(tH1, u);
      gT[u] = rows(tG2, u);
    }
    for (int tk = 0; tk < 4; ++tk)
      for (int tc = 0; tc < 4; ++tc) gw2[tk][tc] = mfma32(hT[tk], gT[tc], gw2[tk][tc]);  // dW2[k][c]
    v8s xa{0, 0, 0, 0, 0, 0, 0, 0};  // A[f][s] = X[s][f], f = li < 8
    if (li < kF)
      for (int j = 0; j < 8; ++j) {
        const int64_t sm = ch * 32 + 16 * (j >> 2) + 4 * lk + (j & 3);
        xa[j] = sm < n ? bf16(X[sm * kF + li]) : (short)0;
      }
    for (int u = 0; u < 4; ++u) gw1[u] = mfma32(xa, rows(tG1, u), gw1[u]);  // dW1[f][u]
    wave_lds_fence();  // the next chunk's stores follow these loads
  }
  // the waves' sums into LDS (the transposed rows are no longer needed), then added in
  // wave order
  __syncthreads();
  float* red = reinterpret_cast<float*>(lds) + wave * kRed;
  for (int tk = 0; tk < 4; ++tk)
    for (int tc = 0; tc < 4; ++tc)
      for (int j = 0; j < 4; ++j) red[OW2 + (16 * tk + 4 * lk + j) * kH + 16 * tc + li] = gw2[tk][tc][j];
  for (int u = 0; u < 4; ++u)
    for (int j = 0; j < 4; ++j) {
      const int f = 4 * lk + j;
      if (f < kF) red[OW1 + f * kH + 16 * u + li] = gw1[u][j];
    }
  // bias and w3 sums: add over the 16 sample lanes of each lane group
  for (int t = 0; t < 4; ++t)
    for (int j = 0; j < 4; ++j) {
      float v1 = gb1[t][j], v2 = gb2[t][j], v3 = gw3[t][j];
      for (int off = 8; off > 0; off >>= 1) {
        v1 += __shfl_xor(v1, off, 16);
        v2 += __shfl_xor(v2, off, 16);
        v3 += __shfl_xor(v3, off, 16);
      }
      if (li == 0) {
        red[OB1 + 16 * t + 4 * lk + j] = v1;
        red[OB2 + 16 * t + 4 * lk + j] = v2;
        red[OW3 + 16 * t + 4 * lk + j] = v3;
      }
    }
  {
    float v3b = gb3, vl = lsum;
    for (int off = 32; off > 0; off >>= 1) {
      v3b += __shfl_xor(v3b, off, 64);
      vl += __shfl_xor(vl, off, 64);
    }
    if (lane == 0) {
      red[OB3] = v3b;
      red[kNP] = vl;
    }
  }
  __syncthreads();
  const float* r0 = reinterpret_cast<const float*>(lds);
  float* gp = grad_part + (int64_t)blockIdx.x * kNP;
  for (int p = tid; p < kNP; p += kThreads) gp[p] = ((r0[p] + r0[kRed + p]) + r0[2 * kRed + p]) + r0[3 * kRed + p];
  if (tid == 0) loss_part[blockIdx.x] = ((r0[kNP] + r0[kRed + kNP]) + r0[2 * kRed + kNP]) + r0[3 * kRed + kNP];
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
