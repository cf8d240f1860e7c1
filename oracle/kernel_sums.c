/*
 * FP64 CPU oracle for the point-set kernel sums.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP product path is
 * compared against; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load the library built from it.  Nothing in
 * registration_amd/ links, imports or calls it.
 *
 * It restates, in plain C with OpenMP over target rows, the three Fortran
 * subroutines of py-lddmm/kernelFunctions.f90 that the hot path calls:
 *
 *   oracle_apply_k            <- applyk             kernelFunctions.f90:1-54
 *   oracle_apply_diff_kt      <- applykdifft        kernelFunctions.f90:56-114
 *   oracle_apply_ddiff_k11and12 <- applykdiff11and12 kernelFunctions.f90:430-496
 *   oracle_apply_diff_k       <- applykdiff1 / applykdiff2 / applykdiff1and2
 *                                                   kernelFunctions.f90:116-290
 *   oracle_apply_k_mat        <- applykmat          kernelFunctions.f90:499-550
 *   oracle_apply_diff_k_mat   <- applykdiffmat      kernelFunctions.f90:552-610
 * (the last two carry one weight per (target, source) pair: the varifold data
 * term's matrix weights, surfaces.py:1184-1276).
 *
 * Semantics kept from the Fortran:
 *   - each target row k sums over the sources l in increasing l order, in FP64
 *     (the OpenMP split is over rows only, so a row's result does not depend
 *     on the thread count);
 *   - ut = |x_k - y_l| / sigma; the Gaussian (ord > 4) squares ut and uses
 *     exp(-ut/2); the Laplacian (ord 0..4) uses P_ord(ut) exp(-ut);
 *   - the diagonal cut-off "ut < 1e-8" compares against the single-precision
 *     literal 1e-8 widened to double (9.99999993922529e-09);
 *   - the Laplacian coefficient tables are default-REAL (float32) literals
 *     widened to double, e.g. 1./3 -> 0.3333333432674408.
 *
 * Arrays are row-major C arrays: x(n, d) is x[k*d + c].  a1/a2 of
 * applyDiffKT are (na, n, m) row-major, i.e. a1[(a*n + k)*m + c].
 */
#include <math.h>
#include <stdint.h>

/* float32 literals widened to double, as gfortran/flang evaluate "1./3". */
#define F(v) ((double)(float)(v))

/* c_(ord+1, j): P_ord(u) = sum_j LAP_C[ord][j] u^j   (kernelFunctions.f90:25-26) */
static const double LAP_C[5][5] = {
    {1, 0, 0, 0, 0},
    {1, 1, 0, 0, 0},
    {1, 1, F(1.0f / 3.0f), 0, 0},
    {1, 1, F(0.4f), F(1.0f / 15.0f), 0},
    {1, 1, F(3.0f / 7.0f), F(2.0f / 21.0f), F(1.0f / 105.0f)}};
static double lap_c(int ord, int j) { return LAP_C[ord][j]; }

/* c1_(ord, j): Q_ord(u) for the first derivative, ord = 1..4
 * (kernelFunctions.f90:85-86). */
static double lap_c1(int ord, int j) {
  double t[4][4] = {
      {1, 0, 0, 0},
      {F(1.0f / 3.0f), F(1.0f / 3.0f), 0, 0},
      {F(1.0f / 5.0f), F(1.0f / 5.0f), F(1.0f / 15.0f), 0},
      {F(1.0f / 7.0f), F(1.0f / 7.0f), F(2.0f / 35.0f), F(1.0f / 105.0f)}};
  return t[ord - 1][j];
}

/* c2_(ord-1, j): R_ord(u) for the second derivative, ord = 2..4
 * (kernelFunctions.f90:461). */
static double lap_c2(int ord, int j) {
  double t[3][3] = {
      {F(1.0f / 3.0f), 0, 0},
      {F(1.0f / 15.0f), F(1.0f / 15.0f), 0},
      {F(1.0f / 35.0f), F(1.0f / 35.0f), F(1.0f / 105.0f)}};
  return t[ord - 2][j];
}

static const double CUTOFF = (double)1e-8f;

static double dist(const double *a, const double *b, int d) {
  double s = 0.0;
  for (int c = 0; c < d; ++c) {
    double t = a[c] - b[c];
    s += t * t;
  }
  return sqrt(s);
}

/* K(r) of applyk, kernelFunctions.f90:33-48 */
static double kval(double r, double sig, int ord) {
  double ut = r / sig;
  if (ord > 4) {
    ut = ut * ut;
    return ut < CUTOFF ? 1.0 : exp(-0.5 * ut);
  }
  if (ut < CUTOFF) return 1.0;
  double lpt = lap_c(ord, 0) + lap_c(ord, 1) * ut + lap_c(ord, 2) * ut * ut +
               lap_c(ord, 3) * ut * ut * ut + lap_c(ord, 4) * ut * ut * ut * ut;
  return lpt * exp(-1.0 * ut);
}

/* K'(r^2) of applykdifft, kernelFunctions.f90:93-108 */
static double kdiff(double r, double sig, int ord) {
  double ut = r / sig;
  if (ord > 4) {
    ut = ut * ut;
    if (ut < CUTOFF) return -1.0 / (2 * sig * sig);
    return -1.0 * exp(-0.5 * ut) / (2 * sig * sig);
  }
  if (ut < CUTOFF) return -lap_c1(ord, 0) / (2 * sig * sig);
  double lpt = lap_c1(ord, 0) + lap_c1(ord, 1) * ut + lap_c1(ord, 2) * ut * ut +
               lap_c1(ord, 3) * ut * ut * ut;
  return -lpt * exp(-1.0 * ut) / (2 * sig * sig);
}

/* K''(r^2) of applykdiff11and12, kernelFunctions.f90:471-490 */
static double kdiff2(double r, double sig, int ord) {
  double ut = r / sig;
  if (ord > 4) {
    ut = ut * ut;
    if (ut < CUTOFF) return 1.0 / (4 * sig * sig * sig * sig);
    return exp(-0.5 * ut) / (4 * sig * sig * sig * sig);
  }
  if (ut < CUTOFF) return lap_c2(ord, 0) / (4 * sig * sig * sig * sig);
  double lpt = lap_c2(ord, 0) + lap_c2(ord, 1) * ut + lap_c2(ord, 2) * ut * ut;
  return lpt * exp(-1.0 * ut) / (4 * sig * sig * sig * sig);
}

/* f(k,:) = sum_l K(|y_k - x_l|) w(l,:)      (applyk, f90:1-54)
 * y: targets (ny, d); x: sources (nx, d); w: (nx, m); f: (ny, m). */
int oracle_apply_k(const double *y, const double *x, const double *w,
                   int64_t ny, int64_t nx, int d, int m, int ord,
                   double sig, double *f) {
  if (d < 1 || d > 8 || m < 1 || m > 16 || ord < 0) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < ny; ++k) {
    double df[16];
    for (int c = 0; c < m; ++c) df[c] = 0.0;
    for (int64_t l = 0; l < nx; ++l) {
      double kh = kval(dist(y + k * d, x + l * d, d), sig, ord);
      for (int c = 0; c < m; ++c) df[c] += kh * w[l * m + c];
    }
    for (int c = 0; c < m; ++c) f[k * m + c] = df[c];
  }
  return 0;
}

/* f(k,:) = sum_l 2 K'(r^2) (y_k - x_l) sum_{a,c} a1[a,k,c] a2[a,l,c]
 * (applykdifft, f90:56-114).  a1: (na, ny, m); a2: (na, nx, m). */
int oracle_apply_diff_kt(const double *y, const double *x, const double *a1,
                         const double *a2, int64_t ny, int64_t nx, int d,
                         int m, int na, int ord, double sig, double *f) {
  if (d < 1 || d > 8 || m < 1 || na < 1 || ord < 1) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < ny; ++k) {
    double df[8];
    for (int c = 0; c < d; ++c) df[c] = 0.0;
    for (int64_t l = 0; l < nx; ++l) {
      double kd = kdiff(dist(y + k * d, x + l * d, d), sig, ord);
      double s = 0.0;
      /* Fortran sum(a1(:,k,:)*a2(:,l,:)) runs column-major: a fastest. */
      for (int c = 0; c < m; ++c)
        for (int a = 0; a < na; ++a)
          s += a1[(a * ny + k) * m + c] * a2[(a * nx + l) * m + c];
      for (int c = 0; c < d; ++c)
        df[c] += kd * 2 * (y[k * d + c] - x[l * d + c]) * s;
    }
    for (int c = 0; c < d; ++c) f[k * d + c] = df[c];
  }
  return 0;
}

/* f(k,:) = sum_l 2 <a1_k, a2_l> (2 K'' <dx,dp> dx + K' dp),
 * dx = x_k - x_l, dp = p_k - p_l   (applykdiff11and12, f90:430-496). */
int oracle_apply_ddiff_k11and12(const double *x, const double *a1,
                                const double *a2, const double *p, int64_t n,
                                int d, int ord, double sig, double *f) {
  if (d < 1 || d > 8 || ord < 2) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    double df[8];
    for (int c = 0; c < d; ++c) df[c] = 0.0;
    for (int64_t l = 0; l < n; ++l) {
      double dx[8], dp[8];
      for (int c = 0; c < d; ++c) {
        dx[c] = x[k * d + c] - x[l * d + c];
        dp[c] = p[k * d + c] - p[l * d + c];
      }
      double r = 0.0;
      for (int c = 0; c < d; ++c) r += dx[c] * dx[c];
      r = sqrt(r);
      double kd = kdiff(r, sig, ord), kd2 = kdiff2(r, sig, ord);
      double aa = 0.0, xp = 0.0;
      for (int c = 0; c < d; ++c) {
        aa += a1[k * d + c] * a2[l * d + c];
        xp += dx[c] * dp[c];
      }
      for (int c = 0; c < d; ++c)
        df[c] += 2 * aa * (2 * kd2 * xp * dx[c] + kd * dp[c]);
    }
    for (int c = 0; c < d; ++c) f[k * d + c] = df[c];
  }
  return 0;
}

/* f(k,:) = sum_l K(|y_k - x_l|) beta(k,l,:)      (applykmat, f90:499-550)
 * beta: (ny, nx, m) row-major, i.e. beta[(k*nx + l)*m + c]. */
int oracle_apply_k_mat(const double *y, const double *x, const double *beta,
                       int64_t ny, int64_t nx, int d, int m, int ord, double sig,
                       double *f) {
  if (d < 1 || d > 8 || m < 1 || m > 16 || ord < 0) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < ny; ++k) {
    double df[16];
    for (int c = 0; c < m; ++c) df[c] = 0.0;
    for (int64_t l = 0; l < nx; ++l) {
      double kh = kval(dist(y + k * d, x + l * d, d), sig, ord);
      for (int c = 0; c < m; ++c) df[c] += kh * beta[(k * nx + l) * m + c];
    }
    for (int c = 0; c < m; ++c) f[k * m + c] = df[c];
  }
  return 0;
}

/* f(k,:) = sum_l 2 K'(r^2) (y_k - x_l) beta(k,l)   (applykdiffmat, f90:552-610)
 * beta: (ny, nx) row-major.  Order 0 is rejected: the Fortran indexes its
 * derivative table at row 0 there. */
int oracle_apply_diff_k_mat(const double *y, const double *x, const double *beta,
                            int64_t ny, int64_t nx, int d, int ord, double sig,
                            double *f) {
  if (d < 1 || d > 8 || ord < 1) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < ny; ++k) {
    double df[8];
    for (int c = 0; c < d; ++c) df[c] = 0.0;
    for (int64_t l = 0; l < nx; ++l) {
      double kd = kdiff(dist(y + k * d, x + l * d, d), sig, ord);
      for (int c = 0; c < d; ++c)
        df[c] += kd * 2 * (y[k * d + c] - x[l * d + c]) * beta[k * nx + l];
    }
    for (int c = 0; c < d; ++c) f[k * d + c] = df[c];
  }
  return 0;
}

/* f(k,:) = sum_l 2 K'(r^2) <x_k - x_l, b_kl> a2(l,:) with b_kl = a1_k (form 1,
 * applykdiff1, f90:116-173), -a1_l (form 2, applykdiff2, f90:175-231) or
 * a1_k - a1_l (form 3, applykdiff1and2, f90:234-290).  The first two add into
 * their f2py output array, which f2py zero-fills, so f is the sum.
 * x, a1, a2: (n, d). */
int oracle_apply_diff_k(const double *x, const double *a1, const double *a2,
                        int64_t n, int d, int form, int ord, double sig, double *f) {
  if (d < 1 || d > 8 || ord < 1 || form < 1 || form > 3) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    double df[8];
    for (int c = 0; c < d; ++c) df[c] = 0.0;
    for (int64_t l = 0; l < n; ++l) {
      double kd = kdiff(dist(x + k * d, x + l * d, d), sig, ord);
      double s = 0.0;
      for (int c = 0; c < d; ++c) {
        double dx = x[k * d + c] - x[l * d + c];
        double b = form == 1 ? a1[k * d + c] : (form == 2 ? a1[l * d + c] : a1[k * d + c] - a1[l * d + c]);
        s += dx * b;
      }
      for (int c = 0; c < d; ++c) {
        if (form == 2)
          df[c] -= kd * 2 * s * a2[l * d + c];
        else
          df[c] += kd * 2 * s * a2[l * d + c];
      }
    }
    for (int c = 0; c < d; ++c) f[k * d + c] = df[c];
  }
  return 0;
}
