// TEST INFRASTRUCTURE ONLY -- small FP64 linear algebra of the CPU oracle.
#include <cmath>

#include "oracle_common.hpp"

namespace oracle {

static void rotate(double A[3][3], double V[3][3], int p, int q) {
  const double apq = A[p][q];
  if (apq == 0.0) return;
  // theta = cot(2 phi), t = tan(phi) of the annihilating rotation
  const double theta = 0.5 * (A[q][q] - A[p][p]) / apq;
  const double abs_theta = std::fabs(theta);
  double t;
  if (abs_theta > 1.0e100) {
    t = 0.5 / theta;
  } else {
    t = 1.0 / (abs_theta + std::sqrt(1.0 + theta * theta));
    if (theta < 0.0) t = -t;
  }
  const double c = 1.0 / std::sqrt(1.0 + t * t);
  const double s = t * c;
  const double tau = s / (1.0 + c);
  const double h = t * apq;
  A[p][p] = A[p][p] - h;
  A[q][q] = A[q][q] + h;
  A[p][q] = 0.0;
  A[q][p] = 0.0;
  const int r = 3 - p - q;  // the remaining index
  {
    const double arp = A[r][p];
    const double arq = A[r][q];
    const double new_rp = arp - s * (arq + arp * tau);
    const double new_rq = arq + s * (arp - arq * tau);
    A[r][p] = new_rp;
    A[p][r] = new_rp;
    A[r][q] = new_rq;
    A[q][r] = new_rq;
  }
  for (int row = 0; row < 3; ++row) {
    const double vp = V[row][p];
    const double vq = V[row][q];
    V[row][p] = vp - s * (vq + vp * tau);
    V[row][q] = vq + s * (vp - vq * tau);
  }
}

void eig3(const double a[6], double w[3], double V[3][3]) {
  double A[3][3] = {{a[0], a[1], a[2]}, {a[1], a[3], a[4]}, {a[2], a[4], a[5]}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  const double mass = std::fabs(a[0]) + std::fabs(a[3]) + std::fabs(a[5]) +
                      2.0 * (std::fabs(a[1]) + std::fabs(a[2]) + std::fabs(a[4]));
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = std::fabs(A[0][1]) + std::fabs(A[0][2]) + std::fabs(A[1][2]);
    if (!(off > 1.0e-17 * mass)) break;
    rotate(A, V, 0, 1);
    rotate(A, V, 0, 2);
    rotate(A, V, 1, 2);
  }
  w[0] = A[0][0];
  w[1] = A[1][1];
  w[2] = A[2][2];
}

void eigvec(const double V[3][3], int k, double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = V[i][k];
  int big = 0;
  for (int i = 1; i < 3; ++i)
    if (std::fabs(out[i]) > std::fabs(out[big])) big = i;
  if (out[big] < 0.0)
    for (int i = 0; i < 3; ++i) out[i] = -out[i];
}

int argmax3(const double w[3]) {
  int k = 0;
  for (int i = 1; i < 3; ++i)
    if (w[i] > w[k]) k = i;
  return k;
}

int argmin3(const double w[3]) {
  int k = 0;
  for (int i = 1; i < 3; ++i)
    if (w[i] < w[k]) k = i;
  return k;
}

void basis(const double a[3], double e1[3], double e2[3]) {
  int k = 0;
  for (int i = 1; i < 3; ++i)
    if (std::fabs(a[i]) < std::fabs(a[k])) k = i;
  // r = unit vector along axis k; e1 = normalise(r - (r.a) a)
  double x = (k == 0 ? 1.0 : 0.0) - a[k] * a[0];
  double y = (k == 1 ? 1.0 : 0.0) - a[k] * a[1];
  double z = (k == 2 ? 1.0 : 0.0) - a[k] * a[2];
  const double len = std::sqrt(x * x + y * y + z * z);
  x = x / len;
  y = y / len;
  z = z / len;
  e1[0] = x;
  e1[1] = y;
  e1[2] = z;
  // e2 = a x e1
  e2[0] = a[1] * z - a[2] * y;
  e2[1] = a[2] * x - a[0] * z;
  e2[2] = a[0] * y - a[1] * x;
}

bool solve3(double M[3][3], double b[3], double x[3]) {
  for (int col = 0; col < 3; ++col) {
    int piv = col;
    for (int r = col + 1; r < 3; ++r)
      if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
    if (!(std::fabs(M[piv][col]) > 0.0)) return false;
    if (piv != col) {
      for (int k = 0; k < 3; ++k) {
        const double t = M[col][k];
        M[col][k] = M[piv][k];
        M[piv][k] = t;
      }
      const double t = b[col];
      b[col] = b[piv];
      b[piv] = t;
    }
    for (int r = col + 1; r < 3; ++r) {
      const double f = M[r][col] / M[col][col];
      for (int k = col; k < 3; ++k) M[r][k] = M[r][k] - f * M[col][k];
      b[r] = b[r] - f * b[col];
    }
  }
  for (int r = 2; r >= 0; --r) {
    double t = b[r];
    for (int k = r + 1; k < 3; ++k) t = t - M[r][k] * x[k];
    x[r] = t / M[r][r];
  }
  return true;
}

bool solve_n(double* M, double* b, double* x, int n) {
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(M[r * n + col]) > std::fabs(M[piv * n + col])) piv = r;
    if (!(std::fabs(M[piv * n + col]) > 0.0)) return false;
    if (piv != col) {
      for (int k = 0; k < n; ++k) {
        const double t = M[col * n + k];
        M[col * n + k] = M[piv * n + k];
        M[piv * n + k] = t;
      }
      const double t = b[col];
      b[col] = b[piv];
      b[piv] = t;
    }
    for (int r = col + 1; r < n; ++r) {
      const double f = M[r * n + col] / M[col * n + col];
      for (int k = col; k < n; ++k) M[r * n + k] = M[r * n + k] - f * M[col * n + k];
      b[r] = b[r] - f * b[col];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double t = b[r];
    for (int k = r + 1; k < n; ++k) t = t - M[r * n + k] * x[k];
    x[r] = t / M[r * n + r];
  }
  return true;
}

template <int N>
static bool cholesky_solve(const double A[N][N], const double b[N], double x[N]) {
  double L[N][N] = {};
  for (int j = 0; j < N; ++j) {
    double s = A[j][j];
    for (int k = 0; k < j; ++k) s = s - L[j][k] * L[j][k];
    if (!(s > 0.0)) return false;
    const double d = std::sqrt(s);
    L[j][j] = d;
    for (int i = j + 1; i < N; ++i) {
      double t = A[i][j];
      for (int k = 0; k < j; ++k) t = t - L[i][k] * L[j][k];
      L[i][j] = t / d;
    }
  }
  double y[N];
  for (int i = 0; i < N; ++i) {
    double t = b[i];
    for (int k = 0; k < i; ++k) t = t - L[i][k] * y[k];
    y[i] = t / L[i][i];
  }
  for (int i = N - 1; i >= 0; --i) {
    double t = y[i];
    for (int k = i + 1; k < N; ++k) t = t - L[k][i] * x[k];
    x[i] = t / L[i][i];
  }
  return true;
}

bool cholesky_solve5(const double A[5][5], const double b[5], double x[5]) {
  return cholesky_solve<5>(A, b, x);
}

bool cholesky_solve6(const double A[6][6], const double b[6], double x[6]) {
  return cholesky_solve<6>(A, b, x);
}

}  // namespace oracle
