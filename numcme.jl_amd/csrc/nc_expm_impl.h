// Small dense matrix exponential kernels (host), included twice by nc_util.hip: once
// compiled for AVX2 and FMA, once for the baseline instruction set. Scaling and squaring
// with the [13/13] Pade approximant and the theta_13 threshold of N. J. Higham, "The
// scaling and squaring method for the matrix exponential revisited", SIAM J. Matrix Anal.
// Appl. 26(4), 2005. Matrices are padded to a multiple of 8 rows (zero blocks, which
// exponentiate to the identity and stay decoupled) in arrays of row stride kExpmMax.
// (No include guard: each inclusion defines the kernels in its own namespace.)
typedef double v4d __attribute__((vector_size(32)));

// C = A B on NP x NP blocks (NP = 8 NB) of row stride kExpmMax; a row of C stays in vector
// registers over the k loop
template <int NB>
void mm_nb(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c) {
  constexpr int NP = 8 * NB;
  for (int i = 0; i < NP; ++i) {
    v4d acc[2 * NB];
    for (int j = 0; j < 2 * NB; ++j) acc[j] = (v4d){0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < NP; ++k) {
      const double aik = a[i * kExpmMax + k];
      const v4d ak = (v4d){aik, aik, aik, aik};
      const double* brow = b + k * kExpmMax;
      for (int j = 0; j < 2 * NB; ++j) {
        v4d bv;
        __builtin_memcpy(&bv, brow + 4 * j, sizeof(bv));
        acc[j] += ak * bv;
      }
    }
    for (int j = 0; j < 2 * NB; ++j) __builtin_memcpy(c + i * kExpmMax + 4 * j, &acc[j], sizeof(v4d));
  }
}

inline void mm(int np, const double* a, const double* b, double* c) {
  switch (np / 8) {
    case 1: mm_nb<1>(a, b, c); break;
    case 2: mm_nb<2>(a, b, c); break;
    case 3: mm_nb<3>(a, b, c); break;
    case 4: mm_nb<4>(a, b, c); break;
    case 5: mm_nb<5>(a, b, c); break;
    case 6: mm_nb<6>(a, b, c); break;
    case 7: mm_nb<7>(a, b, c); break;
    default: mm_nb<8>(a, b, c); break;
  }
}

// row r -= f * row q over columns [j0, np) (np a multiple of 4, j0 rounded down to 4)
inline void axpy_row(double* r, const double* q, double f, int j0, int np) {
  const v4d fv = (v4d){f, f, f, f};
  for (int j = j0 & ~3; j < np; j += 4) {
    v4d x, y;
    __builtin_memcpy(&x, r + j, sizeof(x));
    __builtin_memcpy(&y, q + j, sizeof(y));
    x -= fv * y;
    __builtin_memcpy(r + j, &x, sizeof(x));
  }
}
struct ExpmWork {
  alignas(64) double A[kExpmMax * kExpmMax], A2[kExpmMax * kExpmMax], A4[kExpmMax * kExpmMax],
      A6[kExpmMax * kExpmMax], T[kExpmMax * kExpmMax], W[kExpmMax * kExpmMax], U[kExpmMax * kExpmMax],
      V[kExpmMax * kExpmMax];
};

// R = r_13(A / 2^sq) into w.W (padded blocks of stride kExpmMax); returns sq
inline int pade13(int n, const double* a, ExpmWork& w) {
  static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                               129060195264000.0,   10559470521600.0,    670442572800.0,     33522128640.0,
                               1323241920.0,        40840800.0,          960960.0,           16380.0,
                               182.0,               1.0};
  const double theta13 = 5.371920351148152;
  double norm1 = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(a[i * n + j]);
    norm1 = std::max(norm1, s);
  }
  if (!std::isfinite(norm1)) throw Error(NC_ERR_NUMERIC, "non-finite matrix in expm");
  int sq = 0;
  if (norm1 > theta13) sq = std::max(0, (int)std::ceil(std::log2(norm1 / theta13)));
  const double scale = std::ldexp(1.0, -sq);
  const int np = (n + 7) / 8 * 8;
  constexpr int L = kExpmMax;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) w.A[i * L + j] = (i < n && j < n) ? a[i * n + j] * scale : 0.0;
  mm(np, w.A, w.A, w.A2);
  mm(np, w.A2, w.A2, w.A4);
  mm(np, w.A4, w.A2, w.A6);
  // U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I]
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      const int k = i * L + j;
      w.T[k] = b[13] * w.A6[k] + b[11] * w.A4[k] + b[9] * w.A2[k];
    }
  mm(np, w.A6, w.T, w.W);
  for (int i = 0; i < np; ++i) {
    for (int j = 0; j < np; ++j) {
      const int k = i * L + j;
      w.W[k] += b[7] * w.A6[k] + b[5] * w.A4[k] + b[3] * w.A2[k];
    }
    w.W[i * L + i] += b[1];
  }
  mm(np, w.A, w.W, w.U);
  // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      const int k = i * L + j;
      w.T[k] = b[12] * w.A6[k] + b[10] * w.A4[k] + b[8] * w.A2[k];
    }
  mm(np, w.A6, w.T, w.V);
  for (int i = 0; i < np; ++i) {
    for (int j = 0; j < np; ++j) {
      const int k = i * L + j;
      w.V[k] += b[6] * w.A6[k] + b[4] * w.A4[k] + b[2] * w.A2[k];
    }
    w.V[i * L + i] += b[0];
  }
  // (V - U) R = (V + U): Gaussian elimination with partial pivoting on Q = V - U (in T),
  // the right-hand sides P = V + U (in W) carried along row by row, then back substitution
  double* Q = w.T;
  double* P = w.W;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      const int k = i * L + j;
      Q[k] = w.V[k] - w.U[k];
      P[k] = w.V[k] + w.U[k];
    }
  for (int col = 0; col < np; ++col) {
    int piv = col;
    double best = std::fabs(Q[col * L + col]);
    for (int r = col + 1; r < np; ++r)
      if (std::fabs(Q[r * L + col]) > best) best = std::fabs(Q[r * L + col]), piv = r;
    if (best == 0.0) throw Error(NC_ERR_NUMERIC, "singular Pade denominator in expm");
    if (piv != col)
      for (int j = 0; j < np; ++j) std::swap(Q[col * L + j], Q[piv * L + j]), std::swap(P[col * L + j], P[piv * L + j]);
    const double inv = 1.0 / Q[col * L + col];
    for (int r = col + 1; r < np; ++r) {
      const double f = Q[r * L + col] * inv;
      if (f == 0.0) continue;
      axpy_row(Q + r * L, Q + col * L, f, col, np);
      axpy_row(P + r * L, P + col * L, f, 0, np);
    }
  }
  for (int col = np - 1; col >= 0; --col) {
    const double inv = 1.0 / Q[col * L + col];
    for (int j = 0; j < np; ++j) P[col * L + j] *= inv;
    for (int r = 0; r < col; ++r) {
      const double f = Q[r * L + col];
      if (f != 0.0) axpy_row(P + r * L, P + col * L, f, 0, np);
    }
  }
  return sq;
}


// exp(A) (full) or its first column, for n <= kExpmMax, from the work area w
inline void expm_small(int n, const double* a, double* out, bool first_column, ExpmWork& w) {
  const int sq = pade13(n, a, w);
  const int np = (n + 7) / 8 * 8;
  double* R = w.W;
  double* S = w.U;
  // the first column: R^(2^sq) e1 = R^(2^(sq-1)) (R^(2^(sq-1)) e1), the last squaring a product
  // with a column
  const int full_squarings = first_column ? std::max(sq - 1, 0) : sq;
  for (int k = 0; k < full_squarings; ++k) {
    mm(np, R, R, S);
    std::swap(R, S);
  }
  if (!first_column) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) out[i * n + j] = R[i * kExpmMax + j];
  } else if (sq == 0) {
    for (int i = 0; i < n; ++i) out[i] = R[i * kExpmMax];
  } else {
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += R[i * kExpmMax + k] * R[k * kExpmMax];
      out[i] = s;
    }
  }
}
