/*
 * CPU restatement of the reference's hot loop (TEST INFRASTRUCTURE: the timed CPU baseline of
 * bench.py and the fast exact solutions of the parity tests; never part of the product).
 *
 * NumCME.jl applies its generator with SparseArrays' CSC mul! (src/fspmatrix/sparse/
 * fspsparsematrix.jl:200-214): out is cleared, then each column j scatters nzval[k] * v[j]
 * into out[rowval[k]]. The single-threaded variant restates that loop. The multi-threaded
 * variant stores the same matrix by rows (CSR) and gives each OpenMP thread a range of output
 * rows, which is how a row-parallel CPU SpMV uses all cores.
 *
 * krylov_arnoldi and krylov_combine restate, in plain C and FP64, the Krylov expv work the GPU
 * integrator does per step (Arnoldi with classical Gram-Schmidt and one reorthogonalisation,
 * then the combination of the basis), so CPU and GPU time the same algorithm. uniformize is
 * the series of the uniformization method (oracle/solve.py) for large exact solutions.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int cpu_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* out = A v; csr == 0: CSC column scatter (single thread), csr != 0: CSR rows over threads */
static void spmv(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, int csr, int threads,
                 const double* v, double* out) {
  if (!csr) {
    memset(out, 0, sizeof(double) * (size_t)n);
    for (int64_t j = 0; j < n; ++j) {
      const double vj = v[j];
      for (int64_t k = ptr[j]; k < ptr[j + 1]; ++k) out[idx[k]] += val[k] * vj;
    }
    return;
  }
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * v[idx[k]];
    out[i] = s;
  }
}

void csc_matvec(int64_t n, const int64_t* colptr, const int32_t* rowval, const double* nzval, const double* v,
                double* out) {
  spmv(n, colptr, rowval, nzval, 0, 1, v, out);
}

void csr_matvec(int64_t n, const int64_t* rowptr, const int32_t* colval, const double* nzval, int threads,
                const double* v, double* out) {
  spmv(n, rowptr, colval, nzval, 1, threads, v, out);
}

static double dot(int64_t n, int threads, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(threads)
  for (int64_t k = 0; k < n; ++k) s += a[k] * b[k];
  return s;
}

/* h[i] = V_i . p for i < nv in one sweep over p (each thread sums its rows, then the
 * threads' partial sums are added). */
static void multidot(int64_t n, int threads, int nv, const double* V, const double* p, double* h) {
  for (int i = 0; i < nv; ++i) h[i] = 0.0;
#pragma omp parallel num_threads(threads)
  {
    double loc[64];
    for (int i = 0; i < nv; ++i) loc[i] = 0.0;
#pragma omp for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
      const double pk = p[k];
      for (int i = 0; i < nv; ++i) loc[i] += V[(int64_t)i * n + k] * pk;
    }
#pragma omp critical
    for (int i = 0; i < nv; ++i) h[i] += loc[i];
  }
}

/* V: (m+1) x n, V[0] = w / beta on entry (filled here); H: (m+2)^2 row-major, zeroed here.
 * Returns the breakdown column + 1 (mb) or m when there was none; *avnorm = |A v_m|. */
int krylov_arnoldi(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, int csr, int threads, int m,
                   const double* w, double beta, double* V, double* H, double* p, double* avnorm) {
  const int ld = m + 2;
  memset(H, 0, sizeof(double) * (size_t)ld * ld);
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t i = 0; i < n; ++i) V[i] = w[i] / beta;
  double h[64], h2[64];
  for (int j = 0; j < m; ++j) {
    spmv(n, ptr, idx, val, csr, threads, V + (int64_t)j * n, p);
    for (int pass = 0; pass < 2; ++pass) {
      double* hh = pass ? h2 : h;
      multidot(n, threads, j + 1, V, p, hh);
#pragma omp parallel for schedule(static) num_threads(threads)
      for (int64_t k = 0; k < n; ++k) {
        double acc = p[k];
        for (int i = 0; i <= j; ++i) acc -= hh[i] * V[(int64_t)i * n + k];
        p[k] = acc;
      }
    }
    const double s = sqrt(dot(n, threads, p, p));
    for (int i = 0; i <= j; ++i) H[i * ld + j] = h[i] + h2[i];
    H[(j + 1) * ld + j] = s;
    if (s < 1e-7) return j + 1;
    double* vn = V + (int64_t)(j + 1) * n;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t k = 0; k < n; ++k) vn[k] = p[k] / s;
  }
  spmv(n, ptr, idx, val, csr, threads, V + (int64_t)m * n, p);
  *avnorm = sqrt(dot(n, threads, p, p));
  return m;
}

/* w = sum_{i<mx} c[i] V_i ; returns |w|. */
double krylov_combine(int64_t n, int threads, int mx, const double* V, const double* c, double* w) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(threads)
  for (int64_t k = 0; k < n; ++k) {
    double acc = 0.0;
    for (int i = 0; i < mx; ++i) acc += c[i] * V[(int64_t)i * n + k];
    w[k] = acc;
    s += acc * acc;
  }
  return sqrt(s);
}

/* acc = sum_{k=0..kmax} wts[k] P^k u for the CSR matrix P (the uniformized generator);
 * term and tmp are n-vectors of scratch. */
void uniformize(int64_t n, const int64_t* rowptr, const int32_t* colval, const double* nzval, int threads,
                const double* u, int kmax, const double* wts, double* acc, double* term, double* tmp) {
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t i = 0; i < n; ++i) {
    term[i] = u[i];
    acc[i] = wts[0] * u[i];
  }
  for (int k = 1; k <= kmax; ++k) {
    spmv(n, rowptr, colval, nzval, 1, threads, term, tmp);
    const double wk = wts[k];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t i = 0; i < n; ++i) {
      term[i] = tmp[i];
      acc[i] += wk * tmp[i];
    }
  }
}
