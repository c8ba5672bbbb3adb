/*
 * CPU restatement of the reference's hot loop, for the timed CPU baseline only
 * (TEST INFRASTRUCTURE: imported by bench.py's cpu_baseline leg, never by the product).
 *
 * NumCME.jl applies its generator with SparseArrays' CSC mul! (src/fspmatrix/sparse/
 * fspsparsematrix.jl:200-214): out is cleared, then each column j scatters
 * nzval[k] * v[j] into out[rowval[k]]. csc_matvec restates that loop, single-threaded.
 * arnoldi_cgs2 and combine restate, in plain C and FP64, the Krylov expv work the GPU
 * integrator does per step (Arnoldi with classical Gram-Schmidt plus one
 * reorthogonalisation), so CPU and GPU time the same algorithm.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

void csc_matvec(int64_t n, const int64_t* colptr, const int32_t* rowval, const double* nzval, const double* v,
                double* out) {
  memset(out, 0, sizeof(double) * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    const double vj = v[j];
    for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) out[rowval[k]] += nzval[k] * vj;
  }
}

/* V: (m+1) x n, V[0] = w / beta on entry (filled here); H: (m+2)^2 row-major, zeroed here.
 * Returns the breakdown column + 1 (mb) or m when there was none; *avnorm = |A v_m|. */
int arnoldi_cgs2(int64_t n, const int64_t* colptr, const int32_t* rowval, const double* nzval, int m, const double* w,
                 double beta, double* V, double* H, double* p, double* avnorm) {
  const int ld = m + 2;
  memset(H, 0, sizeof(double) * (size_t)ld * ld);
  for (int64_t i = 0; i < n; ++i) V[i] = w[i] / beta;
  double h[64], h2[64];
  for (int j = 0; j < m; ++j) {
    csc_matvec(n, colptr, rowval, nzval, V + (int64_t)j * n, p);
    for (int pass = 0; pass < 2; ++pass) {
      double* hh = pass ? h2 : h;
      for (int i = 0; i <= j; ++i) {
        const double* vi = V + (int64_t)i * n;
        double s = 0.0;
        for (int64_t k = 0; k < n; ++k) s += vi[k] * p[k];
        hh[i] = s;
      }
      for (int i = 0; i <= j; ++i) {
        const double* vi = V + (int64_t)i * n;
        const double c = hh[i];
        for (int64_t k = 0; k < n; ++k) p[k] -= c * vi[k];
      }
    }
    double s = 0.0;
    for (int64_t k = 0; k < n; ++k) s += p[k] * p[k];
    s = sqrt(s);
    for (int i = 0; i <= j; ++i) H[i * ld + j] = h[i] + h2[i];
    H[(j + 1) * ld + j] = s;
    if (s < 1e-7) return j + 1;
    double* vn = V + (int64_t)(j + 1) * n;
    for (int64_t k = 0; k < n; ++k) vn[k] = p[k] / s;
  }
  csc_matvec(n, colptr, rowval, nzval, V + (int64_t)m * n, p);
  double s = 0.0;
  for (int64_t k = 0; k < n; ++k) s += p[k] * p[k];
  *avnorm = sqrt(s);
  return m;
}

/* w = sum_{i<mx} c[i] V_i ; returns |w|. */
double combine(int64_t n, int mx, const double* V, const double* c, double* w) {
  double s = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    double acc = 0.0;
    for (int i = 0; i < mx; ++i) acc += c[i] * V[(int64_t)i * n + k];
    w[k] = acc;
    s += acc * acc;
  }
  return sqrt(s);
}
