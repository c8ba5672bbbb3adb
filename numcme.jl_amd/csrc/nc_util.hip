// Library-wide utilities: error state, version/device queries, the rocPRIM prefix sum
// used by the state space, and the host dense matrix exponential used by the Krylov
// integrator.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>

#include "nc_internal.h"

namespace nc {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, DevBuf<unsigned char>& tmp, hipStream_t s,
                        uint32_t* total_host) {
  if (n <= 0) {
    if (total_host) *total_host = 0;
    return;
  }
  size_t bytes = 0;
  NC_CHECK_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
  tmp.reserve(bytes + 16);
  NC_CHECK_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
  if (total_host) {
    uint32_t* h = static_cast<uint32_t*>(host_stage(0));
    NC_CHECK_HIP(hipMemcpyAsync(h, in + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipMemcpyAsync(h + 1, out + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NC_CHECK_HIP(hipStreamSynchronize(s));
    *total_host = h[0] + h[1];
  }
}

// ---- dense matrix exponential (host, FP64) --------------------------------------------
// Scaling and squaring with the [13/13] Pade approximant and the theta_13 threshold of
// N. J. Higham, "The scaling and squaring method for the matrix exponential revisited",
// SIAM J. Matrix Anal. Appl. 26(4), 2005. Matrices here are at most (m+2)x(m+2) Krylov
// Hessenberg matrices, so plain triple loops suffice.
namespace {

void matmul(int n, const double* a, const double* b, double* c) {
  std::fill(c, c + (size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      double aik = a[i * n + k];
      if (aik == 0.0) continue;
      const double* brow = b + k * n;
      double* crow = c + i * n;
      for (int j = 0; j < n; ++j) crow[j] += aik * brow[j];
    }
}

// Solve Q X = P in place (X overwrites P); Gaussian elimination with partial pivoting.
void solve_inplace(int n, std::vector<double>& q, std::vector<double>& p) {
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = std::fabs(q[col * n + col]);
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(q[r * n + col]) > best) best = std::fabs(q[r * n + col]), piv = r;
    if (best == 0.0) throw Error(NC_ERR_NUMERIC, "singular Pade denominator in expm");
    if (piv != col) {
      for (int j = 0; j < n; ++j) std::swap(q[col * n + j], q[piv * n + j]), std::swap(p[col * n + j], p[piv * n + j]);
    }
    double inv = 1.0 / q[col * n + col];
    for (int r = col + 1; r < n; ++r) {
      double f = q[r * n + col] * inv;
      if (f == 0.0) continue;
      for (int j = col; j < n; ++j) q[r * n + j] -= f * q[col * n + j];
      for (int j = 0; j < n; ++j) p[r * n + j] -= f * p[col * n + j];
    }
  }
  for (int col = n - 1; col >= 0; --col) {
    double inv = 1.0 / q[col * n + col];
    for (int j = 0; j < n; ++j) p[col * n + j] *= inv;
    for (int r = 0; r < col; ++r) {
      double f = q[r * n + col];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) p[r * n + j] -= f * p[col * n + j];
    }
  }
}

}  // namespace

void expm_dense(int n, const double* a, double* out) {
  static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                               129060195264000.0,   10559470521600.0,    670442572800.0,     33522128640.0,
                               1323241920.0,        40840800.0,          960960.0,           16380.0,
                               182.0,               1.0};
  const double theta13 = 5.371920351148152;
  const size_t nn = (size_t)n * n;
  double norm1 = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(a[i * n + j]);
    norm1 = std::max(norm1, s);
  }
  int sq = 0;
  if (norm1 > theta13) sq = std::max(0, (int)std::ceil(std::log2(norm1 / theta13)));
  const double scale = std::ldexp(1.0, -sq);
  std::vector<double> A(nn), A2(nn), A4(nn), A6(nn), U(nn), V(nn), T(nn), W(nn);
  for (size_t k = 0; k < nn; ++k) A[k] = a[k] * scale;
  matmul(n, A.data(), A.data(), A2.data());
  matmul(n, A2.data(), A2.data(), A4.data());
  matmul(n, A4.data(), A2.data(), A6.data());
  // U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I]
  for (size_t k = 0; k < nn; ++k) T[k] = b[13] * A6[k] + b[11] * A4[k] + b[9] * A2[k];
  matmul(n, A6.data(), T.data(), W.data());
  for (size_t k = 0; k < nn; ++k) W[k] += b[7] * A6[k] + b[5] * A4[k] + b[3] * A2[k];
  for (int i = 0; i < n; ++i) W[(size_t)i * n + i] += b[1];
  matmul(n, A.data(), W.data(), U.data());
  // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
  for (size_t k = 0; k < nn; ++k) T[k] = b[12] * A6[k] + b[10] * A4[k] + b[8] * A2[k];
  matmul(n, A6.data(), T.data(), V.data());
  for (size_t k = 0; k < nn; ++k) V[k] += b[6] * A6[k] + b[4] * A4[k] + b[2] * A2[k];
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] += b[0];
  // (V - U) X = (V + U)
  std::vector<double> Q(nn), P(nn);
  for (size_t k = 0; k < nn; ++k) Q[k] = V[k] - U[k], P[k] = V[k] + U[k];
  solve_inplace(n, Q, P);
  for (int k = 0; k < sq; ++k) {
    matmul(n, P.data(), P.data(), T.data());
    std::swap(P, T);
  }
  std::copy(P.begin(), P.end(), out);
}

}  // namespace nc

using namespace nc;

extern "C" {

int nc_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return NC_ERR_ARG;
  size_t k = std::min(len - 1, g_last_error.size());
  std::memcpy(buf, g_last_error.data(), k);
  buf[k] = '\0';
  return NC_OK;
}

int nc_version(int32_t* major, int32_t* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return NC_OK;
}

int nc_device_name(char* buf, size_t len) {
  return guard([&] {
    NC_REQUIRE(buf && len, NC_ERR_ARG, "null buffer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(NC_ERR_NO_DEVICE, "no HIP device available");
    int dev = 0;
    NC_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NC_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    std::string s = std::string(prop.name) + " (" + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) + " CUs)";
    size_t k = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
  });
}

int nc_expm_dense(int32_t n, const double* a, double* out) {
  return guard([&] {
    NC_REQUIRE(n > 0 && a && out, NC_ERR_ARG, "bad expm arguments");
    expm_dense(n, a, out);
  });
}

}  // extern "C"
