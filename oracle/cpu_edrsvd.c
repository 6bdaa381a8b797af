/*
 * CPU restatement of the reference's randomized SVD step (EDrSVD, basicRandSVD_LU,
 * ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:557-775) in C with OpenMP:
 * FP32 data, FP64 accumulation and solves.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the GPU, checked in the tests
 * against the FP64 numpy oracle (oracle/ttsvd.py rsvd_step).  The engine never loads it.
 *
 * Order of operations (rsvd_lu_half.cu:585-763): Y = A Omega; Q = qr(Y); q times (Z = A^T Q,
 * Qz = qr(Z), Y = A Qz, Q = qr(Y)); B = Q^T A; W = B B^T; eigen-decomposition of W (descending);
 * U = Q U_B; SVT = U_B^T B.  qr is CholeskyQR2 in FP64 (the Q factor with a positive-diagonal R is
 * the one Householder QR gives, up to column signs), the eigensolver a cyclic Jacobi in FP64.  The
 * columns of U get the oracle's sign gauge (largest |entry| positive), the rows of SVT follow.
 *
 * Layouts: A m x n column-major (element (i, j) at i + m j); Omega n x ell row-major (as ttx_omega
 * writes it); U m x ell column-major; sigma ell descending; SVT ell x n column-major.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

/* Y (m x ell, row-major FP64) = A X, X n x ell row-major (float or double, one of the two non-null) */
static void a_times_x(const float* A, int64_t m, int64_t n, const float* Xf, const double* Xd, int ell,
                      double* Y) {
  const int nt = omp_get_max_threads();
  double* part = (double*)calloc((size_t)nt * m * ell, sizeof(double));
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    double* Yt = part + (size_t)t * m * ell;
    double x[64];
#pragma omp for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
      for (int c = 0; c < ell; ++c) x[c] = Xf ? (double)Xf[j * ell + c] : Xd[j * ell + c];
      const float* a = A + j * m;
      for (int64_t i = 0; i < m; ++i) {
        const double ai = (double)a[i];
        double* y = Yt + i * ell;
        for (int c = 0; c < ell; ++c) y[c] += ai * x[c];
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m * ell; ++e) {
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += part[(size_t)t * m * ell + e];
    Y[e] = s;
  }
  free(part);
}

/* Z (n x ell, row-major FP64) = A^T Q, Q m x ell row-major FP64 */
static void at_times_q(const float* A, int64_t m, int64_t n, const double* Q, int ell, double* Z) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) {
    double z[64];
    for (int c = 0; c < ell; ++c) z[c] = 0.0;
    const float* a = A + j * m;
    for (int64_t i = 0; i < m; ++i) {
      const double ai = (double)a[i];
      const double* q = Q + i * ell;
      for (int c = 0; c < ell; ++c) z[c] += ai * q[c];
    }
    for (int c = 0; c < ell; ++c) Z[j * ell + c] = z[c];
  }
}

/* one CholeskyQR round on X (rows x k, row-major FP64), in place; returns 0, or -1 on a pivot that
   is not positive */
static int cholqr_round(double* X, int64_t rows, int k) {
  const int nt = omp_get_max_threads();
  double* part = (double*)calloc((size_t)nt * k * k, sizeof(double));
  double G[64 * 64], R[64 * 64];
#pragma omp parallel
  {
    double* Gt = part + (size_t)omp_get_thread_num() * k * k;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < rows; ++i) {
      const double* x = X + i * k;
      for (int a = 0; a < k; ++a)
        for (int b = a; b < k; ++b) Gt[a * k + b] += x[a] * x[b];
    }
  }
  for (int a = 0; a < k; ++a)
    for (int b = a; b < k; ++b) {
      double s = 0.0;
      for (int t = 0; t < nt; ++t) s += part[(size_t)t * k * k + a * k + b];
      G[a * k + b] = G[b * k + a] = s;
    }
  free(part);
  /* G = R^T R, R upper triangular */
  memset(R, 0, sizeof(R));
  for (int j = 0; j < k; ++j) {
    double d = G[j * k + j];
    for (int p = 0; p < j; ++p) d -= R[p * k + j] * R[p * k + j];
    if (!(d > 0.0)) return -1;
    const double rjj = sqrt(d);
    R[j * k + j] = rjj;
    for (int c = j + 1; c < k; ++c) {
      double s = G[j * k + c];
      for (int p = 0; p < j; ++p) s -= R[p * k + j] * R[p * k + c];
      R[j * k + c] = s / rjj;
    }
  }
  /* X <- X R^-1 (forward substitution per row) */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < rows; ++i) {
    double* x = X + i * k;
    for (int c = 0; c < k; ++c) {
      double s = x[c];
      for (int p = 0; p < c; ++p) s -= x[p] * R[p * k + c];
      x[c] = s / R[c * k + c];
    }
  }
  return 0;
}

static int cholqr2(double* X, int64_t rows, int k) {
  if (cholqr_round(X, rows, k) != 0) return -1;
  return cholqr_round(X, rows, k);
}

/* symmetric eigen-decomposition of W (k x k row-major FP64, destroyed): cyclic Jacobi, eigenvalues
   descending in lam, eigenvectors in the columns of V (row-major k x k) */
static void jacobi_eig(double* W, int k, double* lam, double* V) {
  double Vt[64 * 64];
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) Vt[i * k + j] = (i == j) ? 1.0 : 0.0;
  double tot = 0.0;
  for (int e = 0; e < k * k; ++e) tot += W[e] * W[e];
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) off += 2.0 * W[p * k + q] * W[p * k + q];
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = W[p * k + q];
        if (apq == 0.0) continue;
        const double tau = (W[q * k + q] - W[p * k + p]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int r = 0; r < k; ++r) {
          const double wp = W[r * k + p], wq = W[r * k + q];
          W[r * k + p] = c * wp - s * wq;
          W[r * k + q] = s * wp + c * wq;
        }
        for (int r = 0; r < k; ++r) {
          const double wp = W[p * k + r], wq = W[q * k + r];
          W[p * k + r] = c * wp - s * wq;
          W[q * k + r] = s * wp + c * wq;
        }
        for (int r = 0; r < k; ++r) {
          const double vp = Vt[r * k + p], vq = Vt[r * k + q];
          Vt[r * k + p] = c * vp - s * vq;
          Vt[r * k + q] = s * vp + c * vq;
        }
      }
  }
  /* descending order */
  int idx[64];
  for (int i = 0; i < k; ++i) idx[i] = i;
  for (int i = 1; i < k; ++i) {
    const int v = idx[i];
    int j = i - 1;
    while (j >= 0 && W[idx[j] * k + idx[j]] < W[v * k + v]) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
  for (int c = 0; c < k; ++c) {
    lam[c] = W[idx[c] * k + idx[c]];
    for (int r = 0; r < k; ++r) V[r * k + c] = Vt[r * k + idx[c]];
  }
}

/* the randomized step; returns 0, -1 on bad arguments, -2 when a CholeskyQR pivot is not positive */
int ttx_cpu_rsvd(const float* A, int64_t m, int64_t n, int ell, int q, const float* Omega, float* U, double* sigma,
                 float* SVT) {
  if (!A || !Omega || !U || !sigma || !SVT || m <= 0 || n <= 0 || ell <= 0 || ell > 64 || ell > m || ell > n)
    return -1;
  double* Y = (double*)malloc((size_t)m * ell * sizeof(double));
  double* Z = (double*)malloc((size_t)n * ell * sizeof(double));
  int rc = 0;
  a_times_x(A, m, n, Omega, NULL, ell, Y);
  if (cholqr2(Y, m, ell) != 0) rc = -2;
  for (int it = 0; it < q && rc == 0; ++it) {
    at_times_q(A, m, n, Y, ell, Z);
    if (cholqr2(Z, n, ell) != 0) { rc = -2; break; }
    a_times_x(A, m, n, NULL, Z, ell, Y);
    if (cholqr2(Y, m, ell) != 0) rc = -2;
  }
  if (rc == 0) {
    /* B^T = A^T Q (n x ell) = Z; W = B B^T = Z^T Z */
    at_times_q(A, m, n, Y, ell, Z);
    double W[64 * 64], V[64 * 64], lam[64];
    const int nt = omp_get_max_threads();
    double* part = (double*)calloc((size_t)nt * ell * ell, sizeof(double));
#pragma omp parallel
    {
      double* Wt = part + (size_t)omp_get_thread_num() * ell * ell;
#pragma omp for schedule(static)
      for (int64_t j = 0; j < n; ++j) {
        const double* z = Z + j * ell;
        for (int a = 0; a < ell; ++a)
          for (int b = a; b < ell; ++b) Wt[a * ell + b] += z[a] * z[b];
      }
    }
    for (int a = 0; a < ell; ++a)
      for (int b = a; b < ell; ++b) {
        double s = 0.0;
        for (int t = 0; t < nt; ++t) s += part[(size_t)t * ell * ell + a * ell + b];
        W[a * ell + b] = W[b * ell + a] = s;
      }
    free(part);
    jacobi_eig(W, ell, lam, V);
    /* U = Q U_B with the sign gauge: the largest |entry| of each column positive (first on ties) */
    double sgn[64];
    for (int c = 0; c < ell; ++c) {
      double best = -1.0;
      double sign = 1.0;
      for (int64_t i = 0; i < m; ++i) {
        double u = 0.0;
        for (int a = 0; a < ell; ++a) u += Y[i * ell + a] * V[a * ell + c];
        if (fabs(u) > best) {
          best = fabs(u);
          sign = (u < 0.0) ? -1.0 : 1.0;
        }
      }
      sgn[c] = sign;
      sigma[c] = sqrt(lam[c] > 0.0 ? lam[c] : 0.0);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
      for (int c = 0; c < ell; ++c) {
        double u = 0.0;
        for (int a = 0; a < ell; ++a) u += Y[i * ell + a] * V[a * ell + c];
        U[i + m * c] = (float)(u * sgn[c]);
      }
    /* SVT = U_B^T B: SVT[c, j] = sum_a V[a][c] Z[j][a] */
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
      for (int c = 0; c < ell; ++c) {
        double s = 0.0;
        for (int a = 0; a < ell; ++a) s += V[a * ell + c] * Z[j * ell + a];
        SVT[c + (int64_t)ell * j] = (float)(s * sgn[c]);
      }
  }
  free(Y);
  free(Z);
  return rc;
}

int ttx_cpu_threads(void) { return omp_get_max_threads(); }
