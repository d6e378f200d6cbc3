// Small linear-algebra helpers of the drop-in types.
#include "definitions.h"

#include <cmath>
#include <algorithm>
#include <utility>

Matrix4 Matrix4::inverse() const {
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[r][c];
      a[r][4 + c] = r == c ? 1.0 : 0.0;
    }
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
    if (piv != col)
      for (int c = 0; c < 8; ++c) std::swap(a[col][c], a[piv][c]);
    const double d = a[col][col];
    for (int c = 0; c < 8; ++c) a[col][c] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == col) continue;
      const double f = a[r][col];
      for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
    }
  }
  Matrix4 inv;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) inv.m[r][c] = a[r][4 + c];
  return inv;
}

void Quat::toRotationMatrix(double R[3][3]) const {
  R[0][0] = 1 - 2 * (y * y + z * z);
  R[0][1] = 2 * (x * y - z * w);
  R[0][2] = 2 * (x * z + y * w);
  R[1][0] = 2 * (x * y + z * w);
  R[1][1] = 1 - 2 * (x * x + z * z);
  R[1][2] = 2 * (y * z - x * w);
  R[2][0] = 2 * (x * z - y * w);
  R[2][1] = 2 * (y * z + x * w);
  R[2][2] = 1 - 2 * (x * x + y * y);
}

Vector3 Quat::operator*(const Vector3& v) const {
  double R[3][3];
  toRotationMatrix(R);
  return {R[0][0] * v[0] + R[0][1] * v[1] + R[0][2] * v[2],
          R[1][0] * v[0] + R[1][1] * v[1] + R[1][2] * v[2],
          R[2][0] * v[0] + R[2][1] * v[1] + R[2][2] * v[2]};
}

Matrix4 SE3::matrix() const {
  double R[3][3];
  q_.toRotationMatrix(R);
  Matrix4 M;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) M.m[r][c] = R[r][c];
    M.m[r][3] = t_[r];
  }
  M.m[3][0] = M.m[3][1] = M.m[3][2] = 0.0;
  M.m[3][3] = 1.0;
  return M;
}

Quat Quat::FromTwoVectors(const Vector3& a, const Vector3& b) {
  const Vector3 v0 = a.normalized(), v1 = b.normalized();
  const double c = v1.dot(v0);
  if (c < -1.0 + 1e-12) {
    // (anti)parallel: a half turn about an axis normal to v0; the axis is
    // v0 x e with e the coordinate axis least aligned with v0
    const double ax = std::fabs(v0[0]), ay = std::fabs(v0[1]), az = std::fabs(v0[2]);
    const Vector3 e = (ax <= ay && ax <= az) ? Vector3(1, 0, 0)
                      : (ay <= az)           ? Vector3(0, 1, 0)
                                             : Vector3(0, 0, 1);
    const Vector3 axis = v0.cross(e).normalized();
    const double cc = std::max(c, -1.0);
    const double w2 = (1.0 + cc) * 0.5;
    const Vector3 v = axis * std::sqrt(1.0 - w2);
    return {std::sqrt(w2), v[0], v[1], v[2]};
  }
  const Vector3 axis = v0.cross(v1);
  const double s = std::sqrt((1.0 + c) * 2.0);
  const double invs = 1.0 / s;
  return {s * 0.5, axis[0] * invs, axis[1] * invs, axis[2] * invs};
}

Vector3 eulerAngles012(const double R[3][3]) {
  // first angle from the last column; it is folded into (-pi, 0] before the
  // final sign flip, the second angle's cosine takes the matching sign
  double a = std::atan2(R[1][2], R[2][2]);
  const double c2 = std::sqrt(R[0][0] * R[0][0] + R[0][1] * R[0][1]);
  double b;
  if (a > 0) {
    a -= M_PI;
    b = std::atan2(-R[0][2], -c2);
  } else {
    b = std::atan2(-R[0][2], c2);
  }
  const double s1 = std::sin(a), c1 = std::cos(a);
  const double c = std::atan2(s1 * R[2][0] - c1 * R[1][0], c1 * R[1][1] - s1 * R[2][1]);
  return {-a, -b, -c};
}

void angleAxisToQuaternion(const double aa[3], double q[4]) {
  const double t2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  if (t2 > 0.0) {
    const double t = std::sqrt(t2);
    const double h = t * 0.5;
    const double k = std::sin(h) / t;
    q[0] = std::cos(h);
    q[1] = aa[0] * k;
    q[2] = aa[1] * k;
    q[3] = aa[2] * k;
  } else {
    // first-order expansion at zero
    q[0] = 1.0;
    q[1] = aa[0] * 0.5;
    q[2] = aa[1] * 0.5;
    q[3] = aa[2] * 0.5;
  }
}

void quaternionToAngleAxis(const double q[4], double aa[3]) {
  const double s2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double k = 2.0;
  if (s2 > 0.0) {
    const double s = std::sqrt(s2), c = q[0];
    // the angle of the shorter rotation, in [-pi, pi]
    const double two_theta = 2.0 * (c < 0.0 ? std::atan2(-s, -c) : std::atan2(s, c));
    k = two_theta / s;
  }
  aa[0] = q[1] * k;
  aa[1] = q[2] * k;
  aa[2] = q[3] * k;
}
