// Projection sums of one Arnoldi column over 512-row blocks (internal, device code).
//
// Column j of the Krylov process needs, over the state rows, V_i . q and V_i . V_j for
// i <= j (q = A V_j) and q . q. They are summed per block of 512 consecutive rows, one
// workgroup of 256 threads per block, and the per-block sums are folded in block order
// by a separate launch. Two kernels produce the block sums with the same arithmetic:
// the mat-vec itself, which holds q and V_j for its block's rows in registers
// (k_matvec with the projection epilogue), and a block pass over the stored q (the
// Krylov solver's k_dual_blocks) for the generators whose mat-vec has no such epilogue.
// Either way every block sum is formed the same way from the same values, so a solve
// gives the same result bit for bit whichever kernel ran. The sink rows (past the
// states) are added by the fold.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nc {

constexpr int kDualRows = 512;           // rows per block: 256 threads x 2 rows
constexpr int64_t kDualMaxBlocks = 4096;  // more blocks: the grid-stride passes instead

template <typename T, int R>
struct alignas(sizeof(T) * R) DualPack {
  T v[R];
};

__device__ inline double dual_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block sums of the block starting at row `base`. Lane l of wave w holds x (q) and y
// (V_last) at rows base + 4l .. base + 4l + 3 (x[0..3], y[0..3]) and base + 256 + 4l ..
// (x[4..7], y[4..7]), zero at rows past the states. Wave w takes basis vectors w, w + 4,
// ... (MVW of them, the index clamped to nv - 1 so every load is unconditional; clamped
// duplicates are not stored), their loads issued four vectors at a time: with all of
// them in flight at once the epilogue kernel held 118 registers (FP32) and 180 (FP64),
// four and two waves per SIMD, and ran 1% and 11% slower over a Krylov step's mix of
// basis sizes than with 91 and 115. Writes part[i][blk] = sc_i sum V_i x, part[nv + i][blk] =
// sum V_i y and part[2 nv][blk] = sum x^2, value-major with nblk blocks per value. The
// basis rows of a block lie below the padded vector length (a multiple of 1024), so the
// loads need no bound.
template <typename BT, int MVW>
__device__ inline void dual_block_sums(const BT* __restrict__ V, int64_t ld, int nv, int64_t base,
                                       const double (&x)[8], const double (&y)[8], const double* sc,
                                       double* __restrict__ part, int64_t nblk, int64_t blk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r0 = base + 4 * lane, r1 = r0 + 256;
#pragma unroll
  for (int k0 = 0; k0 < MVW; k0 += 4) {
    constexpr int C = MVW < 4 ? MVW : 4;
    DualPack<BT, 4> a[C], b[C];
#pragma unroll
    for (int u = 0; u < C; ++u) {
      if (k0 + u < MVW) {
        const BT* vk = V + (int64_t)min(4 * (k0 + u) + wave, nv - 1) * ld;
        a[u] = *reinterpret_cast<const DualPack<BT, 4>*>(vk + r0);
        b[u] = *reinterpret_cast<const DualPack<BT, 4>*>(vk + r1);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < C; ++u) {
      if (k0 + u >= MVW) break;
      double d = 0.0, g = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        d = __fma_rn((double)a[u].v[r], x[r], d);
        g = __fma_rn((double)a[u].v[r], y[r], g);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        d = __fma_rn((double)b[u].v[r], x[4 + r], d);
        g = __fma_rn((double)b[u].v[r], y[4 + r], g);
      }
      d = dual_wave_sum(d);
      g = dual_wave_sum(g);
      const int i = 4 * (k0 + u) + wave;
      if (lane == 0 && i < nv) {
        part[(int64_t)i * nblk + blk] = d * (sc ? sc[i] : 1.0);
        part[(int64_t)(nv + i) * nblk + blk] = g;
      }
    }
  }
  if (wave == 0) {
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) s = __fma_rn(x[r], x[r], s);
    s = dual_wave_sum(s);
    if (lane == 0) part[(int64_t)(2 * nv) * nblk + blk] = s;
  }
}

// nv in 1..32: the per-wave vector count picks the exact-load variant (wave-uniform)
template <typename BT>
__device__ inline void dual_block(const BT* __restrict__ V, int64_t ld, int nv, int64_t base, const double (&x)[8],
                                  const double (&y)[8], const double* sc, double* __restrict__ part, int64_t nblk,
                                  int64_t blk) {
  switch ((nv + 3) >> 2) {
    case 1: dual_block_sums<BT, 1>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    case 2: dual_block_sums<BT, 2>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    case 3: dual_block_sums<BT, 3>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    case 4: dual_block_sums<BT, 4>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    case 5: dual_block_sums<BT, 5>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    case 6: dual_block_sums<BT, 6>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    case 7: dual_block_sums<BT, 7>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
    default: dual_block_sums<BT, 8>(V, ld, nv, base, x, y, sc, part, nblk, blk); break;
  }
}

}  // namespace nc
