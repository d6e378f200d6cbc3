// Core types of the C++ drop-in: the names and fields of the reference's
// sloam/include/helpers/definitions.h (PointT, CloudT, VectorType, SE3,
// TreeVertex, FeatureModelParams, Scalar, PIDEF) over small self-contained
// types instead of PCL / Eigen / Sophus.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#define PIDEF 3.14159265

// The vector, quaternion and rigid-transform helpers below are also compiled
// into the GPU keyframe estimator (sloam_amd/csrc/keyframe.hip), so that the
// device evaluates the same expressions in the same order as the host; their
// sines, cosines and arc tangents are detmath.h's, which give the same bits
// on both (the host's and the GPU's libm differ in the last place).
#ifndef SLOAM_HD
#if defined(__HIPCC__)
#define SLOAM_HD __host__ __device__
#else
#define SLOAM_HD
#endif
#endif

#include "detmath.h"

using Scalar = double;

// ---------------------------------------------------------------------------
// fixed-size double vectors / matrices (the subset of the Eigen API the
// reference's object code uses)
// ---------------------------------------------------------------------------
struct Vector3 {
  double v[3] = {0.0, 0.0, 0.0};
  SLOAM_HD Vector3() {}
  SLOAM_HD Vector3(double a, double b, double c) : v{a, b, c} {}
  SLOAM_HD double& operator()(int i) { return v[i]; }
  SLOAM_HD double operator()(int i) const { return v[i]; }
  SLOAM_HD double& operator[](int i) { return v[i]; }
  SLOAM_HD double operator[](int i) const { return v[i]; }
  SLOAM_HD Vector3 operator+(const Vector3& o) const {
    return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]};
  }
  SLOAM_HD Vector3 operator-(const Vector3& o) const {
    return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]};
  }
  SLOAM_HD Vector3 operator-() const { return {-v[0], -v[1], -v[2]}; }
  SLOAM_HD Vector3 operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
  SLOAM_HD Vector3 operator/(double s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
  SLOAM_HD double dot(const Vector3& o) const {
    return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2];
  }
  SLOAM_HD Vector3 cross(const Vector3& o) const {
    return {v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2],
            v[0] * o.v[1] - v[1] * o.v[0]};
  }
  SLOAM_HD double squaredNorm() const { return dot(*this); }
  SLOAM_HD double norm() const { return sqrt(squaredNorm()); }
  SLOAM_HD Vector3 normalized() const { return *this / norm(); }
};
SLOAM_HD inline Vector3 operator*(double s, const Vector3& a) { return a * s; }

struct Vector4 {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  SLOAM_HD Vector4() {}
  SLOAM_HD Vector4(double a, double b, double c, double d) : v{a, b, c, d} {}
  SLOAM_HD double& operator()(int i) { return v[i]; }
  SLOAM_HD double operator()(int i) const { return v[i]; }
  SLOAM_HD double& operator[](int i) { return v[i]; }
  SLOAM_HD double operator[](int i) const { return v[i]; }
  // the reference writes plane.segment(0, 3): the normal part
  SLOAM_HD Vector3 head3() const { return {v[0], v[1], v[2]}; }
};

struct Matrix4 {
  double m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  SLOAM_HD double& operator()(int r, int c) { return m[r][c]; }
  SLOAM_HD double operator()(int r, int c) const { return m[r][c]; }
  SLOAM_HD Vector4 operator*(const Vector4& x) const {
    Vector4 y;
    for (int r = 0; r < 4; ++r)
      y.v[r] = m[r][0] * x.v[0] + m[r][1] * x.v[1] + m[r][2] * x.v[2] + m[r][3] * x.v[3];
    return y;
  }
  SLOAM_HD Matrix4 operator*(const Matrix4& o) const {
    Matrix4 y;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += m[r][k] * o.m[k][c];
        y.m[r][c] = s;
      }
    return y;
  }
  SLOAM_HD Matrix4 transpose() const {
    Matrix4 t;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) t.m[r][c] = m[c][r];
    return t;
  }
  // general 4x4 (Gauss-Jordan, partial pivoting)
  SLOAM_HD Matrix4 inverse() const {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        a[r][c] = m[r][c];
        a[r][4 + c] = r == c ? 1.0 : 0.0;
      }
    for (int col = 0; col < 4; ++col) {
      int piv = col;
      for (int r = col + 1; r < 4; ++r)
        if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
      if (piv != col)
        for (int c = 0; c < 8; ++c) {
          const double t = a[col][c];
          a[col][c] = a[piv][c];
          a[piv][c] = t;
        }
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
};

struct Quat {
  double w = 1.0, x = 0.0, y = 0.0, z = 0.0;
  SLOAM_HD Quat() {}
  SLOAM_HD Quat(double w_, double x_, double y_, double z_) : w(w_), x(x_), y(y_), z(z_) {}
  SLOAM_HD Quat normalized() const {
    const double n = sqrt(w * w + x * x + y * y + z * z);
    return {w / n, x / n, y / n, z / n};
  }
  SLOAM_HD Quat conjugate() const { return {w, -x, -y, -z}; }
  SLOAM_HD Quat operator*(const Quat& o) const {
    return {w * o.w - x * o.x - y * o.y - z * o.z, w * o.x + x * o.w + y * o.z - z * o.y,
            w * o.y - x * o.z + y * o.w + z * o.x, w * o.z + x * o.y - y * o.x + z * o.w};
  }
  SLOAM_HD void toRotationMatrix(double R[3][3]) const {
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
  // R(q) v
  SLOAM_HD Vector3 operator*(const Vector3& v) const {
    double R[3][3];
    toRotationMatrix(R);
    return {R[0][0] * v[0] + R[0][1] * v[1] + R[0][2] * v[2],
            R[1][0] * v[0] + R[1][1] * v[1] + R[1][2] * v[2],
            R[2][0] * v[0] + R[2][1] * v[1] + R[2][2] * v[2]};
  }
  // Eigen's Quaternion::FromTwoVectors: the rotation taking the direction of
  // a to the direction of b about their common normal (half-angle form
  // w = s / 2, v = (a x b) / s, s = sqrt(2 (1 + a.b)) on unit vectors)
  SLOAM_HD static Quat FromTwoVectors(const Vector3& a, const Vector3& b) {
    const Vector3 v0 = a.normalized(), v1 = b.normalized();
    const double c = v1.dot(v0);
    if (c < -1.0 + 1e-12) {
      // (anti)parallel: a half turn about an axis normal to v0; the axis is
      // v0 x e with e the coordinate axis least aligned with v0
      const double ax = fabs(v0[0]), ay = fabs(v0[1]), az = fabs(v0[2]);
      const Vector3 e = (ax <= ay && ax <= az) ? Vector3(1, 0, 0)
                        : (ay <= az)           ? Vector3(0, 1, 0)
                                               : Vector3(0, 0, 1);
      const Vector3 axis = v0.cross(e).normalized();
      const double cc = c < -1.0 ? -1.0 : c;
      const double w2 = (1.0 + cc) * 0.5;
      const Vector3 v = axis * sqrt(1.0 - w2);
      return {sqrt(w2), v[0], v[1], v[2]};
    }
    const Vector3 axis = v0.cross(v1);
    const double s = sqrt((1.0 + c) * 2.0);
    const double invs = 1.0 / s;
    return {s * 0.5, axis[0] * invs, axis[1] * invs, axis[2] * invs};
  }
};

// Eigen 3.3 MatrixBase::eulerAngles(0, 1, 2) of a rotation matrix: angles
// (a, b, c) with R = Rx(a) Ry(b) Rz(c) and a in [0, pi]
SLOAM_HD inline Vector3 eulerAngles012(const double R[3][3]) {
  // first angle from the last column; it is folded into (-pi, 0] before the
  // final sign flip, the second angle's cosine takes the matching sign
  double a = sloam_dm::atan2(R[1][2], R[2][2]);
  const double c2 = sqrt(R[0][0] * R[0][0] + R[0][1] * R[0][1]);
  double b;
  if (a > 0) {
    a -= M_PI;
    b = sloam_dm::atan2(-R[0][2], -c2);
  } else {
    b = sloam_dm::atan2(-R[0][2], c2);
  }
  double s1, c1;
  sloam_dm::sincos(a, &s1, &c1);
  const double c = sloam_dm::atan2(s1 * R[2][0] - c1 * R[1][0], c1 * R[1][1] - s1 * R[2][1]);
  return {-a, -b, -c};
}

// ceres::AngleAxisToQuaternion / QuaternionToAngleAxis (rotation.h of Ceres
// Solver): quaternions as (w, x, y, z)
SLOAM_HD inline void angleAxisToQuaternion(const double aa[3], double q[4]) {
  const double t2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  if (t2 > 0.0) {
    const double t = sqrt(t2);
    const double h = t * 0.5;
    double sh, ch;
    sloam_dm::sincos(h, &sh, &ch);
    const double k = sh / t;
    q[0] = ch;
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

SLOAM_HD inline void quaternionToAngleAxis(const double q[4], double aa[3]) {
  const double s2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double k = 2.0;
  if (s2 > 0.0) {
    const double s = sqrt(s2), c = q[0];
    // the angle of the shorter rotation, in [-pi, pi]
    const double two_theta =
        2.0 * (c < 0.0 ? sloam_dm::atan2(-s, -c) : sloam_dm::atan2(s, c));
    k = two_theta / s;
  }
  aa[0] = q[1] * k;
  aa[1] = q[2] * k;
  aa[2] = q[3] * k;
}

// rigid transform, the Sophus::SE3d operations the reference uses
class SE3 {
 public:
  SLOAM_HD SE3() {}
  SLOAM_HD SE3(const Quat& q, const Vector3& t) : q_(q.normalized()), t_(t) {}
  // a transform whose quaternion is already of unit length, kept bit for bit
  // (the constructor above normalises again); no counterpart in Sophus
  SLOAM_HD static SE3 fromUnitQuaternion(const Quat& q, const Vector3& t) {
    SE3 r;
    r.q_ = q;
    r.t_ = t;
    return r;
  }
  SLOAM_HD Vector3& translation() { return t_; }
  SLOAM_HD const Vector3& translation() const { return t_; }
  SLOAM_HD const Quat& unit_quaternion() const { return q_; }
  SLOAM_HD void setQuaternion(const Quat& q) { q_ = q.normalized(); }
  SLOAM_HD SE3 inverse() const {
    const Quat qi = q_.conjugate();
    return SE3(qi, -(qi * t_));
  }
  SLOAM_HD SE3 operator*(const SE3& o) const { return SE3(q_ * o.q_, q_ * o.t_ + t_); }
  SLOAM_HD Vector3 operator*(const Vector3& p) const { return q_ * p + t_; }
  SLOAM_HD Matrix4 matrix() const {
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

 private:
  Quat q_;
  Vector3 t_;
};

// ---------------------------------------------------------------------------
// points and clouds (pcl::PointXYZI / pcl::PointCloud subset)
// ---------------------------------------------------------------------------
struct PointT {
  float x = 0.f, y = 0.f, z = 0.f, intensity = 0.f;
};

using VectorType = std::vector<PointT>;
using Slash = VectorType;

struct CloudT {
  using Ptr = std::shared_ptr<CloudT>;
  std::vector<PointT> points;
  uint32_t width = 0;
  uint32_t height = 1;
  bool is_dense = true;
  size_t size() const { return points.size(); }
  void push_back(const PointT& p) {
    points.push_back(p);
    width = (uint32_t)points.size();
    height = 1;
  }
  const PointT& at(int col, int row) const { return points[(size_t)row * width + col]; }
};

// ---------------------------------------------------------------------------
// trellis vertex and model parameters (field names of the reference)
// ---------------------------------------------------------------------------
struct TreeVertex {
  int treeId = 0;
  int beam = 0;
  int prevVertexSize = 0;
  Scalar radius = 0.0;
  bool isValid = false;
  PointT coords;
  Slash points;
};

// Defaults: sloam/params/sloam.yaml, plus the values the node hard-codes
// (maxTreeRadius 0.3, maxAxisTheta 10, maxFocusOutlierDistance 0.5,
// sloamNode.cpp:104-106).  The reference leaves the struct uninitialised.
struct FeatureModelParams {
  int scansPerSweep = 1;
  Scalar minTreeModels = 5;
  Scalar minGroundModels = 36;
  Scalar maxLidarDist = 20.0;
  Scalar maxGroundLidarDist = 25.0;
  Scalar minGroundLidarDist = 5.0;
  bool twoStepOptim = true;
  int groundRadiiBins = 2;
  int groundThetaBins = 18;
  Scalar groundRetainThresh = 0.05;
  Scalar groundMatchThresh = 1.0;
  Scalar roughTreeMatchThresh = 2.5;
  Scalar treeMatchThresh = 0.5;
  Scalar maxTreeRadius = 0.3;
  Scalar maxAxisTheta = 10.0;
  Scalar maxFocusOutlierDistance = 0.5;
  Scalar AddNewTreeThreshDist = 1.5;
  int featuresPerTree = 20;
  int numGroundFeatures = 5;
  Scalar defaultTreeRadius = 0.2;
};
