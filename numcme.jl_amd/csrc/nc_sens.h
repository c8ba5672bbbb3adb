// Forward-sensitivity FSP operator (internal).
//
// The block operator of sensfspmatrixsparse.jl:97-142 on u = [v_0; v_1; ...; v_P]:
//   out_0 = A(t) v_0,   out_p = A(t) v_p + dA_p(t) v_0.
// A is the ordinary generator (nc_matrix_s, owned). dA_p is stored per derivative pair
// (r, p) of the gradient sparsity pattern, in the layout of a time-varying slot:
//   ddiag  VT [ld]   -d_p sf_r(x_i)     (the diagonal of reaction r's derivative matrix)
//   dpred  VT [ld]   +d_p sf_r(x_{i-s_r}) or 0 when the predecessor is absent
//   dcol   int32 [ld] predecessor index (or i), one array per reaction used by a pair
//   dsink  VT [E_r]  d_p sf_r at reaction r's sink entries, aligned with A's sink list
// sf is the whole propensity for time-invariant reactions, the state factor for
// separable ones (whose time-factor derivative reuses A's slot of the reaction) and the
// value at the current time for joint ones (refreshed like A's joint slots).
#pragma once

#include "nc_matrix.h"

namespace nc {
constexpr int kMaxSensBlocks = 32;  // parameters + 1
constexpr int kMaxPairs = 128;

struct SensPair {
  int32_t r, p, kind;
  int32_t arr;    // index of the pair's ddiag/dpred arrays, or -1 (no state-factor derivative)
  int32_t cslot;  // index of reaction r's predecessor-column array
  int32_t q;      // A's time-varying slot of reaction r (separable), or -1
  int64_t sink_off;  // offset of the pair's dsink values (arr >= 0)
};
}  // namespace nc

struct nc_sensmatrix_s {
  nc_matrix_s* A = nullptr;  // owned
  nc_model_s* model = nullptr;  // borrowed
  int P = 0;
  std::vector<nc::SensPair> pairs;
  std::vector<int32_t> pstart;  // pairs of block b (b >= 1, parameter b-1): pstart[b] .. pstart[b+1)
  int n_arr = 0, n_cslots = 0;
  bool has_joint = false;
  nc::DevBuf<unsigned char> meta_dev;  // per-pair metadata read by the kernel
  nc::DevBuf<unsigned char> terms_dev;  // derivative terms of the mat-vec's row part, by wave
  int32_t tstart[5] = {0, 0, 0, 0, 0};  // terms of wave w: tstart[w] .. tstart[w+1)
  int32_t bstart[5] = {0, 0, 0, 0, 0};  // vector blocks of wave w: bstart[w] .. bstart[w+1)
  int32_t fold = 1;                     // row-tile fold of the mat-vec's launch
  nc::DevBuf<unsigned char> ddiag, dpred, dsink;
  nc::DevBuf<int32_t> dcol;
  nc::DevBuf<int32_t> states, state_conn;  // kept for joint refreshes
  double t_cache = -INFINITY;
  double norm1 = 0.0;
  ~nc_sensmatrix_s();
};

namespace nc {
nc_sensmatrix_s* sensmatrix_create(nc_space_s* sp, nc_model_s* m, int precision);
// out_b = (A v_b) for block 0 and A v_b + dA_{b-1} v_0 for the others; v holds P+1 VT
// blocks with stride vstride, out P+1 FP64 blocks with stride ostride.
// With epi set (vstride == ostride, the stage vectors in the same block layout), the
// launch also forms the next DP5(4) stage input in its epilogue (see RkEpilogue).
void sens_matvec_launch(nc_sensmatrix_s* SA, double t, const void* v, int64_t vstride, double* out, int64_t ostride,
                        hipStream_t s, const RkEpilogue* epi = nullptr);
// Whether the epilogue kernels cover this operator (at most 16 blocks).
bool sens_rk_fusable(const nc_sensmatrix_s* SA);
}  // namespace nc
