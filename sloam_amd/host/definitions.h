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

using Scalar = double;

// ---------------------------------------------------------------------------
// fixed-size double vectors / matrices (the subset of the Eigen API the
// reference's object code uses)
// ---------------------------------------------------------------------------
struct Vector3 {
  double v[3] = {0.0, 0.0, 0.0};
  Vector3() = default;
  Vector3(double a, double b, double c) : v{a, b, c} {}
  double& operator()(int i) { return v[i]; }
  double operator()(int i) const { return v[i]; }
  double& operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
  Vector3 operator+(const Vector3& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
  Vector3 operator-(const Vector3& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
  Vector3 operator-() const { return {-v[0], -v[1], -v[2]}; }
  Vector3 operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
  Vector3 operator/(double s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
  double dot(const Vector3& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
  Vector3 cross(const Vector3& o) const {
    return {v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2],
            v[0] * o.v[1] - v[1] * o.v[0]};
  }
  double squaredNorm() const { return dot(*this); }
  double norm() const { return std::sqrt(squaredNorm()); }
  Vector3 normalized() const { return *this / norm(); }
};
inline Vector3 operator*(double s, const Vector3& a) { return a * s; }

struct Vector4 {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  Vector4() = default;
  Vector4(double a, double b, double c, double d) : v{a, b, c, d} {}
  double& operator()(int i) { return v[i]; }
  double operator()(int i) const { return v[i]; }
  double& operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
  // the reference writes plane.segment(0, 3): the normal part
  Vector3 head3() const { return {v[0], v[1], v[2]}; }
};

struct Matrix4 {
  double m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  double& operator()(int r, int c) { return m[r][c]; }
  double operator()(int r, int c) const { return m[r][c]; }
  Vector4 operator*(const Vector4& x) const {
    Vector4 y;
    for (int r = 0; r < 4; ++r) y.v[r] = m[r][0] * x.v[0] + m[r][1] * x.v[1] + m[r][2] * x.v[2] + m[r][3] * x.v[3];
    return y;
  }
  Matrix4 operator*(const Matrix4& o) const {
    Matrix4 y;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += m[r][k] * o.m[k][c];
        y.m[r][c] = s;
      }
    return y;
  }
  Matrix4 transpose() const {
    Matrix4 t;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) t.m[r][c] = m[c][r];
    return t;
  }
  Matrix4 inverse() const;  // general 4x4 (Gauss-Jordan, partial pivoting)
};

struct Quat {
  double w = 1.0, x = 0.0, y = 0.0, z = 0.0;
  Quat() = default;
  Quat(double w_, double x_, double y_, double z_) : w(w_), x(x_), y(y_), z(z_) {}
  Quat normalized() const {
    const double n = std::sqrt(w * w + x * x + y * y + z * z);
    return {w / n, x / n, y / n, z / n};
  }
  Quat conjugate() const { return {w, -x, -y, -z}; }
  Quat operator*(const Quat& o) const {
    return {w * o.w - x * o.x - y * o.y - z * o.z, w * o.x + x * o.w + y * o.z - z * o.y,
            w * o.y - x * o.z + y * o.w + z * o.x, w * o.z + x * o.y - y * o.x + z * o.w};
  }
  // R(q) v
  Vector3 operator*(const Vector3& v) const;
  void toRotationMatrix(double R[3][3]) const;
  // Eigen's Quaternion::FromTwoVectors: the rotation taking the direction of
  // a to the direction of b about their common normal (half-angle form
  // w = s / 2, v = (a x b) / s, s = sqrt(2 (1 + a.b)) on unit vectors)
  static Quat FromTwoVectors(const Vector3& a, const Vector3& b);
};

// Eigen 3.3 MatrixBase::eulerAngles(0, 1, 2) of a rotation matrix: angles
// (a, b, c) with R = Rx(a) Ry(b) Rz(c) and a in [0, pi]
Vector3 eulerAngles012(const double R[3][3]);

// ceres::AngleAxisToQuaternion / QuaternionToAngleAxis (rotation.h of Ceres
// Solver): quaternions as (w, x, y, z)
void angleAxisToQuaternion(const double aa[3], double q[4]);
void quaternionToAngleAxis(const double q[4], double aa[3]);

// rigid transform, the Sophus::SE3d operations the reference uses
class SE3 {
 public:
  SE3() = default;
  SE3(const Quat& q, const Vector3& t) : q_(q.normalized()), t_(t) {}
  Vector3& translation() { return t_; }
  const Vector3& translation() const { return t_; }
  const Quat& unit_quaternion() const { return q_; }
  void setQuaternion(const Quat& q) { q_ = q.normalized(); }
  SE3 inverse() const {
    const Quat qi = q_.conjugate();
    return SE3(qi, -(qi * t_));
  }
  SE3 operator*(const SE3& o) const { return SE3(q_ * o.q_, q_ * o.t_ + t_); }
  Vector3 operator*(const Vector3& p) const { return q_ * p + t_; }
  Matrix4 matrix() const;

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
