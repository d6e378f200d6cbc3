// Tests of the C++ drop-in classes (Plane, Cylinder, sloam::sloam), the
// cases of the reference's gtests (sloam/src/tests/plane_test.cpp,
// cylinder_test.cpp, core_test.cpp) over the same data, plus host-only checks
// of the rotation helpers.  gtest is not part of this toolchain: a small
// harness with the same EXPECT_* names.
//
//   dropin_test <fixture dir> [host|gpu|dump <out file>]
//
// The fixture directory holds still_landmarks_t{0,1}.lm and
// still_ground_t{0,1}.xyz, written from tests/golden/ by the pytest wrapper
// (tests/test_dropin.py):
//   .lm : int32 n_trees; per tree int32 n_vertices; per vertex int32 treeId,
//         beam, prevVertexSize, isValid, float64 radius, float32 coords[3],
//         int32 n_points, float32 points[3 n_points]
//   .xyz: int32 n; float32 xyz[3 n]
//   .org: int32 height, width; float32 xyz[3 height width] (NaN = empty)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "cylinder.h"
#include "mapManager.h"
#include "pcdIO.h"
#include "definitions.h"
#include "detmath.h"
#include "plane.h"
#include "rosmsg.h"
#include "segmentation.h"
#include "sloam.h"
#include "trellis.h"

namespace {

int g_failures = 0;
const char* g_test = "";

#define EXPECT_TRUE(c)                                                                    \
  do {                                                                                    \
    if (!(c)) {                                                                           \
      std::fprintf(stderr, "[%s] %s:%d: EXPECT_TRUE(%s) failed\n", g_test, __FILE__,      \
                   __LINE__, #c);                                                         \
      ++g_failures;                                                                       \
    }                                                                                     \
  } while (0)
#define EXPECT_EQ(a, b)                                                                   \
  do {                                                                                    \
    const double va_ = (a), vb_ = (b);                                                    \
    if (!(va_ == vb_)) {                                                                  \
      std::fprintf(stderr, "[%s] %s:%d: EXPECT_EQ(%s, %s): %.17g vs %.17g\n", g_test,     \
                   __FILE__, __LINE__, #a, #b, va_, vb_);                                 \
      ++g_failures;                                                                       \
    }                                                                                     \
  } while (0)
#define EXPECT_NEAR(a, b, tol)                                                            \
  do {                                                                                    \
    const double va_ = (a), vb_ = (b);                                                    \
    if (!(std::fabs(va_ - vb_) <= (tol))) {                                               \
      std::fprintf(stderr, "[%s] %s:%d: EXPECT_NEAR(%s, %s, %s): %.17g vs %.17g\n",       \
                   g_test, __FILE__, __LINE__, #a, #b, #tol, va_, vb_);                   \
      ++g_failures;                                                                       \
    }                                                                                     \
  } while (0)

std::string g_dir;

using Landmarks = std::vector<std::vector<TreeVertex>>;

template <typename T>
void rd(FILE* f, T* p, size_t n, const std::string& path) {
  if (n && std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short read: " + path);
}

Landmarks read_landmarks(const std::string& name) {
  const std::string path = g_dir + "/" + name + ".lm";
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  Landmarks out;
  int32_t nt;
  rd(f, &nt, 1, path);
  out.resize(nt);
  for (auto& tree : out) {
    int32_t nv;
    rd(f, &nv, 1, path);
    tree.resize(nv);
    for (auto& v : tree) {
      int32_t vi[4];
      rd(f, vi, 4, path);
      v.treeId = vi[0];
      v.beam = vi[1];
      v.prevVertexSize = vi[2];
      v.isValid = vi[3] != 0;
      rd(f, &v.radius, 1, path);
      float c[3];
      rd(f, c, 3, path);
      v.coords.x = c[0];
      v.coords.y = c[1];
      v.coords.z = c[2];
      int32_t np;
      rd(f, &np, 1, path);
      std::vector<float> p(3 * (size_t)np);
      rd(f, p.data(), p.size(), path);
      for (int i = 0; i < np; ++i) v.points.push_back(PointT{p[3 * i], p[3 * i + 1], p[3 * i + 2], 0.f});
    }
  }
  std::fclose(f);
  return out;
}

CloudT::Ptr read_cloud(const std::string& name) {
  const std::string path = g_dir + "/" + name + ".xyz";
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  int32_t n;
  rd(f, &n, 1, path);
  std::vector<float> p(3 * (size_t)n);
  rd(f, p.data(), p.size(), path);
  std::fclose(f);
  CloudT::Ptr c(new CloudT);
  for (int i = 0; i < n; ++i) c->push_back(PointT{p[3 * i], p[3 * i + 1], p[3 * i + 2], 0.f});
  return c;
}

CloudT::Ptr read_organized(const std::string& name) {
  const std::string path = g_dir + "/" + name + ".org";
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  int32_t hw[2];
  rd(f, hw, 2, path);
  std::vector<float> p(3 * (size_t)hw[0] * hw[1]);
  rd(f, p.data(), p.size(), path);
  std::fclose(f);
  CloudT::Ptr c(new CloudT);
  for (size_t i = 0; i < p.size() / 3; ++i)
    c->points.push_back(PointT{p[3 * i], p[3 * i + 1], p[3 * i + 2], 0.f});
  c->height = (uint32_t)hw[0];
  c->width = (uint32_t)hw[1];
  c->is_dense = false;
  return c;
}

// ---------------------------------------------------------------------------
// host-only: rotation helpers against their defining identities
// ---------------------------------------------------------------------------
void rot_x(double a, double R[3][3]) {
  const double c = std::cos(a), s = std::sin(a);
  const double M[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
  std::memcpy(R, M, sizeof(M));
}
void rot_y(double a, double R[3][3]) {
  const double c = std::cos(a), s = std::sin(a);
  const double M[3][3] = {{c, 0, s}, {0, 1, 0}, {-s, 0, c}};
  std::memcpy(R, M, sizeof(M));
}
void rot_z(double a, double R[3][3]) {
  const double c = std::cos(a), s = std::sin(a);
  const double M[3][3] = {{c, -s, 0}, {s, c, 0}, {0, 0, 1}};
  std::memcpy(R, M, sizeof(M));
}
void mul(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

void test_euler_angles() {
  std::mt19937 g(7);
  std::uniform_real_distribution<double> ua(0.01, M_PI - 0.01), ub(-1.5, 1.5), uc(-3.1, 3.1);
  for (int it = 0; it < 200; ++it) {
    const double a = ua(g), b = ub(g), c = uc(g);
    double X[3][3], Y[3][3], Z[3][3], XY[3][3], R[3][3];
    rot_x(a, X);
    rot_y(b, Y);
    rot_z(c, Z);
    mul(X, Y, XY);
    mul(XY, Z, R);
    const Vector3 e = eulerAngles012(R);
    // the decomposition is unique for a in (0, pi), |b| < pi / 2
    EXPECT_NEAR(e[0], a, 1e-9);
    EXPECT_NEAR(e[1], b, 1e-9);
    EXPECT_NEAR(e[2], c, 1e-9);
  }
  // identity: all zero (the first angle is folded to +0, not pi)
  const double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const Vector3 e = eulerAngles012(I);
  EXPECT_EQ(std::fabs(e[0]) + std::fabs(e[1]) + std::fabs(e[2]), 0.0);
}

void test_from_two_vectors() {
  std::mt19937 g(8);
  std::normal_distribution<double> n(0.0, 1.0);
  for (int it = 0; it < 200; ++it) {
    const Vector3 a(n(g), n(g), n(g)), b(n(g), n(g), n(g));
    const Quat q = Quat::FromTwoVectors(a, b);
    EXPECT_NEAR(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z, 1.0, 1e-12);
    const Vector3 r = q * a.normalized();
    const Vector3 bn = b.normalized();
    for (int i = 0; i < 3; ++i) EXPECT_NEAR(r[i], bn[i], 1e-12);
    // about the common normal: the axis is orthogonal to both
    EXPECT_NEAR(q.x * a[0] + q.y * a[1] + q.z * a[2], 0.0, 1e-12 * a.norm());
  }
  // parallel and antiparallel
  const Quat q0 = Quat::FromTwoVectors(Vector3(0, 0, 1), Vector3(0, 0, 2));
  EXPECT_EQ(q0.w, 1.0);
  const Quat q1 = Quat::FromTwoVectors(Vector3(0, 0, 1), Vector3(0, 0, -1));
  const Vector3 r1 = q1 * Vector3(0, 0, 1);
  EXPECT_NEAR(r1[2], -1.0, 1e-12);
}

void test_angle_axis() {
  std::mt19937 g(9);
  std::normal_distribution<double> n(0.0, 1.0);
  std::uniform_real_distribution<double> angle(0.0, 3.1);
  for (int it = 0; it < 200; ++it) {
    // angles in the canonical range [0, pi) that the inverse returns
    double w[3] = {n(g), n(g), n(g)};
    const double th = angle(g);
    const double s = th / std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (double& x : w) x *= s;
    double q[4], back[3];
    angleAxisToQuaternion(w, q);
    quaternionToAngleAxis(q, back);
    for (int i = 0; i < 3; ++i) EXPECT_NEAR(back[i], w[i], 1e-12);
    // the quaternion of the rotation by th about w / th
    EXPECT_NEAR(q[0], std::cos(th / 2), 1e-15);
  }
  const double z[3] = {0, 0, 0};
  double q[4];
  angleAxisToQuaternion(z, q);
  EXPECT_EQ(q[0], 1.0);
}

// error of a double result in units of the last place of the exact value,
// the exact value taken from the long double function (64-bit significand)
double ulp_error(double got, long double exact) {
  const double e = (double)exact;
  const double ulp = std::nextafter(std::fabs(e), INFINITY) - std::fabs(e);
  return (double)(std::fabs((long double)got - exact) / (long double)ulp);
}

// sloam_dm's sine, cosine, arc tangent and arc cosine against the long double
// libm over the ranges the pose and the model gates use and beyond; the
// quadrants, signed zeros and NaN of atan2 as in C
void test_detmath() {
  std::mt19937_64 g(11);
  double worst[4] = {0, 0, 0, 0};
  std::uniform_real_distribution<double> small(-4.0, 4.0), wide(-1000.0, 1000.0), unit(-1.0, 1.0);
  std::normal_distribution<double> n(0.0, 1.0);
  std::uniform_real_distribution<double> mag(-30.0, 30.0);
  for (int it = 0; it < 400000; ++it) {
    const double x = (it & 1) ? small(g) : wide(g);
    double s, c;
    sloam_dm::sincos(x, &s, &c);
    worst[0] = std::max(worst[0], ulp_error(s, sinl((long double)x)));
    worst[1] = std::max(worst[1], ulp_error(c, cosl((long double)x)));
    // atan2 over every quadrant and magnitude ratios from 1e-13 to 1e13
    const double y2 = n(g) * std::pow(10.0, mag(g) / 5), x2 = n(g) * std::pow(10.0, mag(g) / 5);
    worst[2] = std::max(worst[2],
                        ulp_error(sloam_dm::atan2(y2, x2), atan2l((long double)y2, (long double)x2)));
    const double a = unit(g);
    worst[3] = std::max(worst[3], ulp_error(sloam_dm::acos(a), acosl((long double)a)));
  }
  std::printf("detmath worst ulp error: sin %.3f cos %.3f atan2 %.3f acos %.3f\n", worst[0],
              worst[1], worst[2], worst[3]);
  EXPECT_TRUE(worst[0] <= 0.8);
  EXPECT_TRUE(worst[1] <= 0.8);
  EXPECT_TRUE(worst[2] <= 1.3);
  EXPECT_TRUE(worst[3] <= 1.8);
  // exact points and the sign conventions of C's atan2
  EXPECT_EQ(sloam_dm::atan2(0.0, 1.0), 0.0);
  EXPECT_EQ(sloam_dm::atan2(0.0, -1.0), M_PI);
  EXPECT_EQ(sloam_dm::atan2(-0.0, -1.0), -M_PI);
  EXPECT_EQ(sloam_dm::atan2(1.0, 0.0), M_PI / 2);
  EXPECT_EQ(sloam_dm::atan2(-1.0, 0.0), -M_PI / 2);
  EXPECT_EQ(sloam_dm::atan2(1.0, 1.0), M_PI / 4);
  EXPECT_NEAR(sloam_dm::atan2(-1.0, -1.0), -3 * M_PI / 4, 4.5e-16);
  EXPECT_TRUE(std::signbit(sloam_dm::atan2(-0.0, 1.0)));
  EXPECT_TRUE(std::isnan(sloam_dm::atan2(NAN, 1.0)));
  EXPECT_TRUE(std::isnan(sloam_dm::acos(1.0 + 1e-15)));
  EXPECT_EQ(sloam_dm::acos(1.0), 0.0);
  EXPECT_EQ(sloam_dm::acos(-1.0), M_PI);
  EXPECT_EQ(sloam_dm::sin(0.0), 0.0);
  EXPECT_EQ(sloam_dm::cos(0.0), 1.0);
  double s, c;
  sloam_dm::sincos(2e6, &s, &c);
  EXPECT_TRUE(std::isnan(s) && std::isnan(c));
}

void test_se3() {
  const Quat q = Quat(0.9, 0.1, -0.3, 0.2).normalized();
  const SE3 T(q, Vector3(1, 2, 3));
  const SE3 I = T * T.inverse();
  EXPECT_NEAR(I.translation().norm(), 0.0, 1e-15);
  EXPECT_NEAR(std::fabs(I.unit_quaternion().w), 1.0, 1e-15);
  const Matrix4 M = T.matrix(), Mi = M.inverse(), P = M * Mi;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) EXPECT_NEAR(P(r, c), r == c ? 1.0 : 0.0, 1e-14);
}

// ---------------------------------------------------------------------------
// plane_test.cpp
// ---------------------------------------------------------------------------
FeatureModelParams plane_params() {
  FeatureModelParams p;
  p.scansPerSweep = 1;
  p.maxLidarDist = 15.0;
  p.maxGroundLidarDist = 30.0;
  p.minGroundLidarDist = 0.0;
  p.groundRadiiBins = 1;
  p.groundThetaBins = 18;
  p.groundMatchThresh = 2.0;
  p.groundRetainThresh = 0.1;
  p.numGroundFeatures = 1;
  p.maxTreeRadius = 0.3;
  p.maxAxisTheta = 10;
  p.maxFocusOutlierDistance = 0.5;
  p.roughTreeMatchThresh = 3.0;
  p.treeMatchThresh = 1.0;
  p.AddNewTreeThreshDist = 2.0;
  p.featuresPerTree = 2;
  p.defaultTreeRadius = 0.1;
  return p;
}

VectorType three_points() {
  PointT a, b, c;
  b.y = 1;
  c.x = 1;
  return {a, b, c};
}

void plane_initializes() {
  Plane plane(three_points(), plane_params());
  EXPECT_TRUE(plane.isValid);
}

void plane_distance_to_feature() {
  const VectorType g = three_points();
  Plane plane(g, plane_params());
  const float d = plane.distance(g[0]);
  EXPECT_NEAR(0.0, d, 0.1);
}

void plane_translate_model() {
  Plane plane(three_points(), plane_params());
  SE3 tf;
  tf.translation()[2] = 1;
  const Vector3 centroid = plane.model.centroid;
  plane.project(tf);
  EXPECT_EQ(centroid[2] + 1, plane.model.centroid[2]);
}

// ---------------------------------------------------------------------------
// cylinder_test.cpp
// ---------------------------------------------------------------------------
FeatureModelParams cylinder_params() {
  FeatureModelParams p = plane_params();
  p.maxLidarDist = 30.0;
  p.groundThetaBins = 1;
  p.numGroundFeatures = 3;
  return p;
}

std::vector<Cylinder> valid_cylinders() {
  const FeatureModelParams params = cylinder_params();
  VectorType g = three_points();
  PointT d;
  d.x = 1;
  d.y = 1;
  g.push_back(d);
  const Plane plane(g, params);
  EXPECT_TRUE(plane.isValid);
  std::vector<Cylinder> out;
  for (const auto& l : read_landmarks("still_landmarks_t0")) {
    Cylinder c(l, plane, params);
    if (c.isValid) out.push_back(c);
  }
  return out;
}

void cylinder_distance_to_model() {
  auto cyl = valid_cylinders();
  EXPECT_TRUE(!cyl.empty());
  if (cyl.empty()) return;
  const float d = (float)cyl[0].distance(cyl[0].model);
  EXPECT_EQ(0.0, d);
}

void cylinder_distance_to_feature() {
  auto cyl = valid_cylinders();
  EXPECT_TRUE(!cyl.empty());
  if (cyl.empty()) return;
  const float d = (float)cyl[0].distance(cyl[0].features[0]);
  EXPECT_NEAR(0.0, d, 0.1);
}

void cylinder_translate_model() {
  auto cyl = valid_cylinders();
  EXPECT_TRUE(!cyl.empty());
  if (cyl.empty()) return;
  SE3 tf;
  tf.translation()[0] = 1;
  const Vector3 root = cyl[0].model.root;
  cyl[0].project(tf);
  EXPECT_EQ(root[0] + 1, cyl[0].model.root[0]);
}

// Cylinder from a cylinder fit (no device calls): axis pointing up, root the
// bottom vertex's foot on the fitted axis (within 2 m of the ground, so
// valid; the reference's ray-plane intersection only moves a root forward
// along the ray, so an upward axis keeps it), the radius floor and the
// validity rules of the line-fit model.
void cylinder_from_fit() {
  FeatureModelParams fm;  // sloam.yaml defaults
  PlaneParameters pp;
  pp.plane = Vector4(0, 0, 1, 0);  // z = 0
  pp.centroid = Vector3(0, 0, 0);
  const Plane ground(pp, VectorType(), true);
  auto tree = [](double x, double y, double z0, double z1) {
    std::vector<TreeVertex> v;
    for (int k = 0; k < 10; ++k) {
      TreeVertex t;
      t.treeId = 7;
      t.radius = 0.25;
      t.isValid = true;
      t.coords.x = (float)x;
      t.coords.y = (float)y;
      t.coords.z = (float)(z0 + (z1 - z0) * k / 9.0);
      for (int i = 0; i < 5; ++i) t.points.push_back(t.coords);
      v.push_back(t);
    }
    return v;
  };
  auto fit = [](double cx, double cy, double cz, double ax, double ay, double az, double r) {
    CylinderFit f;
    f.ok = true;
    f.root[0] = cx;
    f.root[1] = cy;
    f.root[2] = cz;
    const double n = std::sqrt(ax * ax + ay * ay + az * az);
    f.ray[0] = ax / n;
    f.ray[1] = ay / n;
    f.ray[2] = az / n;
    f.radius = r;
    f.inliers = 50;
    return f;
  };
  const auto t5 = tree(5.0, 0.0, 0.3, 3.0);  // vertex [1] at z = 0.6
  {
    // vertical axis given pointing down
    Cylinder c(t5, ground, fm, fit(5.1, 0.05, 1.7, 0, 0, -1, 0.25));
    EXPECT_TRUE(c.isValid);
    EXPECT_NEAR(1.0, c.model.ray[2], 1e-12);
    EXPECT_NEAR(5.1, c.model.root[0], 1e-12);
    EXPECT_NEAR(0.05, c.model.root[1], 1e-12);
    EXPECT_NEAR((double)t5[1].coords.z, c.model.root[2], 1e-12);
    EXPECT_EQ(0.25, c.model.radius);
    EXPECT_EQ((size_t)7, c.id);
    EXPECT_EQ((size_t)fm.featuresPerTree, c.features.size());
  }
  {
    // tilted axis through (5, 0, 1): the foot of vertex [1] = (5, 0, z1) is
    // (5, 0, 1) + (z1 - 1) cos(a) (sin a, 0, cos a)
    const double a = 4.0 * PIDEF / 180.0;
    Cylinder c(t5, ground, fm, fit(5.0, 0.0, 1.0, std::sin(a), 0, std::cos(a), 0.22));
    EXPECT_TRUE(c.isValid);
    const double s = ((double)t5[1].coords.z - 1.0) * std::cos(a);
    EXPECT_NEAR(5.0 + s * std::sin(a), c.model.root[0], 1e-12);
    EXPECT_NEAR(1.0 + s * std::cos(a), c.model.root[2], 1e-12);
  }
  {
    // radius below the default: raised to it
    Cylinder c(t5, ground, fm, fit(5.0, 0.0, 1.0, 0, 0, 1, 0.08));
    EXPECT_TRUE(c.isValid);
    EXPECT_EQ(fm.defaultTreeRadius, c.model.radius);
  }
  // rejected: too wide, axis too tilted, failed fit, tree too short, out of range
  EXPECT_TRUE(!Cylinder(t5, ground, fm, fit(5.0, 0.0, 1.0, 0, 0, 1, 0.35)).isValid);
  {
    const double a = 15.0 * PIDEF / 180.0;
    EXPECT_TRUE(!Cylinder(t5, ground, fm, fit(5.0, 0.0, 1.0, std::sin(a), 0, std::cos(a), 0.25))
                     .isValid);
  }
  {
    CylinderFit bad = fit(5.0, 0.0, 1.0, 0, 0, 1, 0.25);
    bad.ok = false;
    EXPECT_TRUE(!Cylinder(t5, ground, fm, bad).isValid);
  }
  EXPECT_TRUE(!Cylinder(tree(5.0, 0.0, 0.3, 1.0), ground, fm, fit(5.0, 0.0, 1.0, 0, 0, 1, 0.25))
                   .isValid);
  EXPECT_TRUE(!Cylinder(tree(25.0, 0.0, 0.3, 3.0), ground, fm, fit(25.0, 0.0, 1.0, 0, 0, 1, 0.25))
                   .isValid);
}

// MapManager (mapManager.cpp): new landmarks, hits and replacement through
// the last submap's indices, the hit threshold of getMap, and the submap's
// nearest-first order over the 200 newest landmarks (at most 100).
void map_manager() {
  FeatureModelParams fm;
  PlaneParameters pp;
  pp.plane = Vector4(0, 0, 1, 0);
  pp.centroid = Vector3(0, 0, 0);
  const Plane ground(pp, VectorType(), true);
  // a valid tree model rooted at (x, y, 0)
  auto tree = [&](double x, double y) {
    std::vector<TreeVertex> v;
    for (int k = 0; k < 10; ++k) {
      TreeVertex t;
      t.treeId = 3;
      t.radius = 0.25;
      t.coords.x = (float)x;
      t.coords.y = (float)y;
      t.coords.z = (float)(0.3 * k);
      for (int i = 0; i < 5; ++i) t.points.push_back(t.coords);
      v.push_back(t);
    }
    CylinderFit f;
    f.ok = true;
    f.root[0] = x;
    f.root[1] = y;
    f.root[2] = 1.0;
    f.ray[2] = 1.0;
    f.radius = 0.25;
    f.inliers = 40;
    // the vertices lie within 20 m of the plane centroid only near the origin
    PlaneParameters q = pp;
    q.centroid = Vector3(x, y, 0);
    return Cylinder(v, Plane(q, VectorType(), true), fm, f);
  };
  MapManager m;
  std::vector<Cylinder> sub;
  m.getSubmap(SE3(), sub);
  EXPECT_TRUE(sub.empty());
  // first scan: three new landmarks at x = 1, 5, 9
  std::vector<Cylinder> obs = {tree(1, 0), tree(5, 0), tree(9, 0)};
  for (const auto& c : obs) EXPECT_TRUE(c.isValid);
  m.updateMap(obs, {-1, -1, -1});
  EXPECT_EQ((size_t)3, m.size());
  EXPECT_TRUE(m.getMap().empty());  // one hit each
  // the submap from x = 6: nearest first (5, 9, 1)
  SE3 at6;
  at6.translation()[0] = 6;
  m.getSubmap(at6, sub);
  EXPECT_EQ((size_t)3, sub.size());
  EXPECT_NEAR(5.0, sub[0].model.root[0], 1e-12);
  EXPECT_NEAR(9.0, sub[1].model.root[0], 1e-12);
  EXPECT_NEAR(1.0, sub[2].model.root[0], 1e-12);
  // observation 0 matches submap entry 0 (map entry 1) and moves it; the
  // other is new
  std::vector<Cylinder> obs2 = {tree(5.2, 0), tree(20, 0)};
  m.updateMap(obs2, {0, -1});
  EXPECT_EQ((size_t)4, m.size());
  // two more hits on the same landmark: it enters the map
  for (int r = 0; r < 2; ++r) {
    std::vector<Cylinder> s2;
    m.getSubmap(at6, s2);
    EXPECT_NEAR(5.2, s2[0].model.root[0], 1e-6);
    std::vector<Cylinder> o = {tree(5.2, 0)};
    m.updateMap(o, {0});
  }
  const auto full = m.getMap();
  EXPECT_EQ((size_t)1, full.size());
  if (!full.empty()) EXPECT_NEAR(5.2, full[0].model.root[0], 1e-6);
  // only the 200 newest landmarks and at most 100 of them take part
  MapManager big;
  std::vector<Cylinder> many;
  std::vector<int> none;
  for (int i = 0; i < 250; ++i) {
    many.push_back(tree(0.01 * i, 0));
    none.push_back(-1);
  }
  big.updateMap(many, none);
  // the nearest 100 of the 250 (from (0, 0, 1)) are landmarks 0..99; of
  // those, only the newest 200 of the map (250 - index < 200: 51..99) count
  std::vector<Cylinder> sb;
  big.getSubmap(SE3(), sb);
  EXPECT_EQ((size_t)49, sb.size());
  bool ordered = true;
  for (size_t i = 0; i < sb.size(); ++i)
    ordered = ordered && std::fabs(sb[i].model.root[0] - 0.01 * (51 + (int)i)) < 1e-6;
  EXPECT_TRUE(ordered);
  // the submap by reference: the same trees in the same order, no copies;
  // an update through it moves the referenced landmark
  MapView view;
  big.getSubmap(SE3(), view);
  EXPECT_EQ(sb.size(), view.size());
  bool same = view.size() == sb.size();
  for (size_t i = 0; same && i < view.size(); ++i)
    same = view[i]->model.root[0] == sb[i].model.root[0];
  EXPECT_TRUE(same);
  std::vector<Cylinder> moved = {tree(0.6, 0.3)};
  big.updateMap(moved, {0});
  EXPECT_EQ((size_t)250, big.size());
  MapView view2;
  big.getSubmap(SE3(), view2);
  bool found = false;
  for (const Cylinder* c : view2) found = found || std::fabs(c->model.root[1] - 0.3) < 1e-9;
  EXPECT_TRUE(found);
}

// PCD files (pcdIO.h): organised ascii and binary round trips with NaN
// points, fields found by name in another order with extra fields of other
// types and counts, and the failures PCL reports as -1.
// ROS 1 wire format: byte layouts worked out by hand from the
// serialisation rules (little-endian scalars, uint32 counts before strings and
// T[], fixed arrays inline), round trips of every message type, the
// PointXYZI layout of pcl::toROSMsg, fields by name at any offset and type,
// and malformed input refused.
void ros_messages() {
  using namespace sloam_wire;
  {
    Header h;
    h.seq = 7;
    h.stamp = {1, 2};
    h.frame_id = "map";
    std::vector<uint8_t> b;
    serialize(h, b);
    const uint8_t want[] = {7, 0, 0, 0, 1, 0, 0, 0, 2, 0, 0, 0, 3, 0, 0, 0, 'm', 'a', 'p'};
    EXPECT_EQ(b.size(), sizeof(want));
    EXPECT_TRUE(b.size() == sizeof(want) && std::memcmp(b.data(), want, sizeof(want)) == 0);
    EXPECT_EQ(serializedLength(h), 19);
    Header back;
    EXPECT_TRUE(deserialize(b.data(), b.size(), back));
    EXPECT_TRUE(back.seq == 7 && back.stamp.nsec == 2 && back.frame_id == "map");
    EXPECT_TRUE(!deserialize(b.data(), b.size() - 1, back));  // truncated string
    b.push_back(0);
    EXPECT_TRUE(!deserialize(b.data(), b.size(), back));  // trailing byte
  }
  {
    ROSCylinder c;
    c.root[0] = 1.5f;
    c.ray[2] = 1.0f;
    c.radii = {0.25};
    c.radius = 0.5f;
    c.id = -3;
    std::vector<uint8_t> b;
    serialize(c, b);
    // 3 + 3 floats, count 1, one double, one float, int64
    EXPECT_EQ(b.size(), 12 + 12 + 4 + 8 + 4 + 8);
    float root0;
    std::memcpy(&root0, b.data(), 4);
    EXPECT_EQ(root0, 1.5);
    uint32_t cnt;
    std::memcpy(&cnt, b.data() + 24, 4);
    EXPECT_EQ(cnt, 1);
    int64_t id;
    std::memcpy(&id, b.data() + 40, 8);
    EXPECT_EQ((double)id, -3);
    // a count larger than the bytes left is refused before any allocation
    std::vector<uint8_t> bad = b;
    const uint32_t huge = 0x7fffffffu;
    std::memcpy(bad.data() + 24, &huge, 4);
    ROSCylinder back;
    EXPECT_TRUE(!deserialize(bad.data(), bad.size(), back));
  }
  // pcl::toROSMsg layout of PointXYZI and the way back
  CloudT cloud;
  cloud.width = 3;
  cloud.height = 2;
  for (int i = 0; i < 6; ++i) {
    PointT p;
    p.x = 0.5f * i;
    p.y = -1.0f * i;
    p.z = 2.0f + i;
    p.intensity = 10.0f * i;
    cloud.points.push_back(p);
  }
  cloud.points[4].x = std::numeric_limits<float>::quiet_NaN();
  cloud.is_dense = false;
  PointCloud2 msg;
  toROSMsg(cloud, msg);
  EXPECT_TRUE(msg.height == 2 && msg.width == 3 && msg.point_step == 32 && msg.row_step == 96);
  EXPECT_EQ(msg.fields.size(), 4);
  EXPECT_TRUE(msg.fields[3].name == "intensity" && msg.fields[3].offset == 16 &&
              msg.fields[3].datatype == FLOAT32 && msg.fields[3].count == 1);
  EXPECT_EQ(msg.data.size(), 6 * 32);
  float v;
  std::memcpy(&v, msg.data.data() + 5 * 32 + 16, 4);
  EXPECT_EQ(v, 50.0);
  std::vector<uint8_t> b;
  serialize(msg, b);
  // header (seq, stamp, empty frame id: 16), height, width, fields (count,
  // then per field 4 + name + offset 4 + datatype 1 + count 4), bigendian,
  // steps, data (count + 192), is_dense
  const size_t fields = 4 + (4 + 1 + 9) * 3 + (4 + 9 + 9);
  EXPECT_EQ(b.size(), 16 + 8 + fields + 1 + 8 + 4 + 192 + 1);
  PointCloud2 msg2;
  EXPECT_TRUE(deserialize(b.data(), b.size(), msg2));
  CloudT back;
  EXPECT_TRUE(fromROSMsg(msg2, back));
  EXPECT_TRUE(back.width == 3 && back.height == 2 && !back.is_dense);
  bool same = back.points.size() == 6;
  for (size_t i = 0; same && i < 6; ++i) {
    const PointT &a = cloud.points[i], &c = back.points[i];
    same = std::memcmp(&a, &c, sizeof(PointT)) == 0;
  }
  EXPECT_TRUE(same);
  // another producer's layout: intensity as uint16 first, then x y z as
  // float64, a padded row; fields found by name
  PointCloud2 other;
  other.height = 2;
  other.width = 1;
  other.fields = {PointField{"intensity", 0, UINT16, 1}, PointField{"x", 8, FLOAT64, 1},
                  PointField{"y", 16, FLOAT64, 1}, PointField{"z", 24, FLOAT64, 1}};
  other.point_step = 32;
  other.row_step = 40;
  other.data.assign(80, 0);
  for (int r = 0; r < 2; ++r) {
    const uint16_t inten = (uint16_t)(100 + r);
    std::memcpy(other.data.data() + 40 * r, &inten, 2);
    const double xyz[3] = {1.25 + r, -2.5, 3.0};
    std::memcpy(other.data.data() + 40 * r + 8, xyz, 24);
  }
  other.is_dense = true;
  CloudT oc;
  EXPECT_TRUE(fromROSMsg(other, oc));
  EXPECT_TRUE(oc.points.size() == 2 && oc.points[1].x == 2.25f && oc.points[1].y == -2.5f &&
              oc.points[1].z == 3.0f && oc.points[1].intensity == 101.0f && oc.is_dense);
  // malformed: short data, field past point_step, big-endian
  PointCloud2 bad = other;
  bad.data.resize(60);
  EXPECT_TRUE(!fromROSMsg(bad, oc));
  bad = other;
  bad.fields[1].offset = 28;
  EXPECT_TRUE(!fromROSMsg(bad, oc));
  bad = other;
  bad.is_bigendian = true;
  EXPECT_TRUE(!fromROSMsg(bad, oc));
  // an observation as sloamNode would publish it, through bytes and back
  ROSObservation obs;
  obs.header.frame_id = "map";
  obs.header.stamp = {1700000000u, 5u};
  SE3 pose(Quat(0.9, 0.1, -0.2, 0.3), Vector3(1.0, 2.0, 3.0));
  obs.pose.pose = toPose(pose);
  obs.initialGuess.pose = toPose(SE3());
  obs.pc = msg;
  ROSCylinder cyl;
  cyl.radii = {0.1, 0.2};
  cyl.id = 9;
  obs.treeModels = {cyl, cyl};
  obs.ground.coefs[2] = 1.0f;
  obs.ground.features = msg;
  obs.matches = {3, -1, 4};
  obs.success = true;
  b.clear();
  serialize(obs, b);
  ROSObservation ob;
  EXPECT_TRUE(deserialize(b.data(), b.size(), ob));
  std::vector<uint8_t> b2;
  serialize(ob, b2);
  EXPECT_TRUE(b == b2);
  EXPECT_TRUE(ob.matches.size() == 3 && ob.matches[2] == 4 && ob.success &&
              ob.treeModels.size() == 2 && ob.treeModels[1].radii[1] == 0.2);
  const SE3 pb = fromPose(ob.pose.pose);
  EXPECT_NEAR(pb.unit_quaternion().x, pose.unit_quaternion().x, 1e-15);
  EXPECT_NEAR(pb.translation()[2], 3.0, 0);
  for (size_t cut = 0; cut < b.size(); cut += 97)
    EXPECT_TRUE(!deserialize(b.data(), cut, ob));  // every prefix is refused
  // the remaining message types round-trip through their bytes
  ROSSubMap sm;
  sm.treeModels = {cyl};
  sm.t_map_anchor.covariance[35] = 0.5;
  ROSSweep sw;
  ROSScan sc;
  sc.treeModels = {cyl};
  sw.scans = {sc, sc};
  LoopClosure lc;
  lc.source_idx = 4;
  ObservationPair op;
  op.source = obs;
  op.target_idx = 2;
  auto round = [](const auto& m) {
    std::vector<uint8_t> x, y;
    serialize(m, x);
    typename std::decay<decltype(m)>::type back;
    if (!deserialize(x.data(), x.size(), back)) return false;
    serialize(back, y);
    return x == y && serializedLength(m) == x.size();
  };
  EXPECT_TRUE(round(sm));
  EXPECT_TRUE(round(sw));
  EXPECT_TRUE(round(lc));
  EXPECT_TRUE(round(op));
}

void pcd_io() {
  const std::string dir = g_dir + "/";
  CloudT c;
  c.width = 4;
  c.height = 3;
  c.points.resize(12);
  for (int i = 0; i < 12; ++i) {
    c.points[i].x = 0.1f * i - 0.3f;
    c.points[i].y = 1e-7f * i;
    c.points[i].z = -12345.678f + i;
    c.points[i].intensity = (float)(i % 5);
  }
  c.points[5].x = c.points[5].y = c.points[5].z = std::numeric_limits<float>::quiet_NaN();
  c.is_dense = false;
  auto same = [](const CloudT& a, const CloudT& b) {
    if (a.width != b.width || a.height != b.height || a.points.size() != b.points.size() ||
        a.is_dense != b.is_dense)
      return false;
    for (size_t i = 0; i < a.points.size(); ++i) {
      const float va[4] = {a.points[i].x, a.points[i].y, a.points[i].z, a.points[i].intensity};
      const float vb[4] = {b.points[i].x, b.points[i].y, b.points[i].z, b.points[i].intensity};
      for (int k = 0; k < 4; ++k)
        if (!(va[k] == vb[k] || (std::isnan(va[k]) && std::isnan(vb[k])))) return false;
    }
    return true;
  };
  CloudT r;
  EXPECT_EQ(0, sloam_io::savePCDFileASCII(dir + "rt_ascii.pcd", c));
  EXPECT_EQ(0, sloam_io::loadPCDFile(dir + "rt_ascii.pcd", r));
  EXPECT_TRUE(same(c, r));
  EXPECT_EQ(0, sloam_io::savePCDFileBinary(dir + "rt_bin.pcd", c));
  EXPECT_EQ(0, sloam_io::loadPCDFile(dir + "rt_bin.pcd", r));
  EXPECT_TRUE(same(c, r));
  // fields in another order, an extra uint16 field and a two-element field
  {
    FILE* f = std::fopen((dir + "mixed.pcd").c_str(), "wb");
    std::fprintf(f,
                 "VERSION 0.7\nFIELDS intensity ring x y z t\nSIZE 4 2 4 8 4 4\n"
                 "TYPE F U F F F F\nCOUNT 1 1 1 1 1 2\nWIDTH 2\nHEIGHT 1\nPOINTS 2\n"
                 "DATA binary\n");
    for (int i = 0; i < 2; ++i) {
      const float in = 7.0f + i, x = 1.5f * i, z = -2.0f, t[2] = {9.f, 9.f};
      const uint16_t ring = (uint16_t)(3 + i);
      const double y = 0.25 + i;
      std::fwrite(&in, 4, 1, f);
      std::fwrite(&ring, 2, 1, f);
      std::fwrite(&x, 4, 1, f);
      std::fwrite(&y, 8, 1, f);
      std::fwrite(&z, 4, 1, f);
      std::fwrite(t, 4, 2, f);
    }
    std::fclose(f);
    CloudT m;
    EXPECT_EQ(0, sloam_io::loadPCDFile(dir + "mixed.pcd", m));
    EXPECT_EQ((size_t)2, m.points.size());
    if (m.points.size() == 2) {
      EXPECT_EQ(1.5f, m.points[1].x);
      EXPECT_EQ(1.25f, m.points[1].y);
      EXPECT_EQ(-2.0f, m.points[1].z);
      EXPECT_EQ(8.0f, m.points[1].intensity);
      EXPECT_TRUE(m.is_dense);
    }
  }
  // the reference's ascii header style with nan tokens and no intensity
  {
    FILE* f = std::fopen((dir + "ref.pcd").c_str(), "wb");
    std::fprintf(f,
                 "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z\n"
                 "SIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 1\nWIDTH 3\nHEIGHT 1\n"
                 "VIEWPOINT 0 0 0 1 0 0 0\nPOINTS 3\nDATA ascii\n1 2 3\nnan nan nan\n"
                 "-2.8259385 2.4685013 1.4757745\n");
    std::fclose(f);
    CloudT a;
    EXPECT_EQ(0, sloam_io::loadPCDFile(dir + "ref.pcd", a));
    EXPECT_EQ((size_t)3, a.points.size());
    if (a.points.size() == 3) {
      EXPECT_TRUE(std::isnan(a.points[1].x));
      EXPECT_EQ(-2.8259385f, a.points[2].x);
      EXPECT_EQ(0.0f, a.points[2].intensity);
      EXPECT_TRUE(!a.is_dense);
    }
  }
  // failures: no file, compressed data, POINTS != WIDTH * HEIGHT
  CloudT e;
  EXPECT_EQ(-1, sloam_io::loadPCDFile(dir + "missing.pcd", e));
  {
    FILE* f = std::fopen((dir + "bad.pcd").c_str(), "wb");
    std::fprintf(f, "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nWIDTH 1\nPOINTS 1\n"
                    "DATA binary_compressed\n");
    std::fclose(f);
    EXPECT_EQ(-1, sloam_io::loadPCDFile(dir + "bad.pcd", e));
    f = std::fopen((dir + "bad2.pcd").c_str(), "wb");
    std::fprintf(f, "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nWIDTH 2\nHEIGHT 2\nPOINTS 3\n"
                    "DATA ascii\n1 2 3\n");
    std::fclose(f);
    EXPECT_EQ(-1, sloam_io::loadPCDFile(dir + "bad2.pcd", e));
  }
  // hostile headers fail with -1 before any allocation of their claimed size:
  // a width * height past 2^31 - 1, a width that overflows the product, a
  // binary section shorter than POINTS records, an absurd COUNT
  {
    const char* headers[] = {
        "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nWIDTH 2000000000\nHEIGHT 2000000000\n"
        "DATA binary\n",
        "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nWIDTH 9223372036854775807\nHEIGHT 2\n"
        "DATA ascii\n1 2 3\n",
        "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nWIDTH 100000000\nHEIGHT 1\nPOINTS 100000000\n"
        "DATA binary\n0123456789ab",
        "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nWIDTH 1000000\nHEIGHT 1\nDATA ascii\n1 2 3\n",
        "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 2000000000\nWIDTH 1\nHEIGHT 1\n"
        "DATA binary\n0123456789ab",
    };
    int k = 0;
    for (const char* h : headers) {
      const std::string p = dir + "hostile" + std::to_string(k++) + ".pcd";
      FILE* f = std::fopen(p.c_str(), "wb");
      std::fputs(h, f);
      std::fclose(f);
      CloudT c;
      EXPECT_EQ(-1, sloam_io::loadPCDFile(p, c));
      EXPECT_EQ((size_t)0, c.points.size());
    }
  }
}

// ---------------------------------------------------------------------------
// trellis: the reference's tree clouds give its landmark archives
// (sloam/src/tests/aux/<seq>_tree_<t>.pcd -> <seq>_landmarks_<t>); the
// archives hold half the code's vertex radius
// ---------------------------------------------------------------------------
void instance_reference_clouds() {
  for (const char* stamp : {"t0", "t1"}) {
    Instance inst;
    inst.set_params(Instance::Params());
    inst.reset_tree_id();
    std::vector<std::vector<TreeVertex>> lm;
    CloudT::Ptr tree = read_organized(std::string("still_tree_") + stamp);
    inst.computeGraph(tree, tree, lm);
    const Landmarks ref = read_landmarks(std::string("still_landmarks_") + stamp);
    EXPECT_EQ(lm.size(), ref.size());
    if (lm.size() != ref.size()) continue;
    for (size_t t = 0; t < lm.size(); ++t) {
      EXPECT_EQ(lm[t].size(), ref[t].size());
      if (lm[t].size() != ref[t].size()) continue;
      for (size_t v = 0; v < lm[t].size(); ++v) {
        const TreeVertex &a = lm[t][v], &b = ref[t][v];
        EXPECT_EQ(a.treeId, b.treeId);
        EXPECT_EQ(a.points.size(), b.points.size());
        EXPECT_NEAR(a.coords.x, b.coords.x, 1e-6 * (1 + std::fabs(b.coords.x)));
        EXPECT_NEAR(a.coords.y, b.coords.y, 1e-6 * (1 + std::fabs(b.coords.y)));
        EXPECT_NEAR(a.coords.z, b.coords.z, 1e-6 * (1 + std::fabs(b.coords.z)));
        EXPECT_NEAR(a.radius, 2.0 * b.radius, 2e-6);
      }
    }
  }
}

void instance_find_clusters() {
  Instance inst;
  CloudT::Ptr tree = read_organized("still_tree_t0");
  std::vector<int> labels;
  std::vector<std::vector<int>> idx;
  inst.findClusters(tree, labels, idx);
  EXPECT_EQ(labels.size(), tree->points.size());
  size_t n = 0;
  for (const auto& l : idx) n += l.size();
  size_t finite = 0;
  for (const auto& p : tree->points) finite += std::isfinite(p.x) ? 1 : 0;
  EXPECT_EQ(n, finite);
  EXPECT_TRUE(!idx.empty());
}

// ---------------------------------------------------------------------------
// core_test.cpp
// ---------------------------------------------------------------------------
FeatureModelParams core_params() {
  FeatureModelParams p;
  p.scansPerSweep = 1;
  p.maxLidarDist = 15.0;
  p.maxGroundLidarDist = 30.0;
  p.minGroundLidarDist = 0.0;
  p.groundRadiiBins = 1;
  p.groundThetaBins = 18;
  p.groundMatchThresh = 2.0;
  p.groundRetainThresh = 0.05;
  p.maxTreeRadius = 0.3;
  p.maxAxisTheta = 10;
  p.maxFocusOutlierDistance = 0.5;
  p.roughTreeMatchThresh = 3.0;
  p.treeMatchThresh = 1.0;
  p.AddNewTreeThreshDist = 2.0;
  p.featuresPerTree = 2;
  p.numGroundFeatures = 60;
  p.defaultTreeRadius = 0.1;
  return p;
}

SloamInput core_input(const char* stamp) {
  SloamInput in;
  in.landmarks = read_landmarks(std::string("still_landmarks_") + stamp);
  // core_test.cpp:78-81: every vertex holds exactly 5 points (zero-padded)
  for (auto& tree : in.landmarks)
    for (auto& v : tree) v.points.resize(5);
  in.groundCloud = read_cloud(std::string("still_ground_") + stamp);
  in.poseEstimate = SE3();
  return in;
}

void core_first_scan() {
  sloam::sloam s;
  SloamOutput out;
  s.setFmParams(core_params());
  SloamInput in = core_input("t0");
  s.RunSloam(in, out);
  EXPECT_TRUE(out.tm.size() > 0);
  EXPECT_TRUE(s.getPrevGroundModel().size() > 0);
}

void core_second_scan() {
  sloam::sloam s;
  SloamOutput out;
  s.setFmParams(core_params());
  SloamInput in = core_input("t1");
  s.RunSloam(in, out);
  EXPECT_TRUE(out.tm.size() > 0);
  EXPECT_TRUE(s.getPrevGroundModel().size() > 0);
}

void run_two_scans(sloam::sloam& s, SloamOutput& out, bool* success, bool twoStep) {
  FeatureModelParams p = core_params();
  p.twoStepOptim = twoStep;
  s.setFmParams(p);
  SloamInput t0 = core_input("t0"), t1 = core_input("t1");
  s.RunSloam(t0, out);
  t1.mapModels = out.tm;
  *success = s.RunSloam(t1, out);
}

void core_sloam_success() {
  for (bool twoStep : {true, false}) {
    sloam::sloam s;
    SloamOutput out;
    bool success = false;
    run_two_scans(s, out, &success, twoStep);
    EXPECT_TRUE(success);
  }
}

void core_pose_optimization() {
  // the two-step solve of sloam.yaml (twoStepOptim: true); the joint solve
  // is checked against the oracle on the same matches by the dump
  sloam::sloam s;
  SloamOutput out;
  bool success = false;
  run_two_scans(s, out, &success, true);
  const float translation = (float)out.T_Delta.translation().norm();
  EXPECT_NEAR(0.0, translation, 0.1);
  // the still sequence: the estimated pose stays near the origin
  EXPECT_NEAR(0.0, out.T_Map_Curr.translation().norm(), 0.1);
}

void core_object_association() {
  sloam::sloam s;
  SloamOutput out;
  bool success = false;
  run_two_scans(s, out, &success, true);
  bool matched = false;
  for (int m : out.matches) matched = matched || m != -1;
  EXPECT_TRUE(matched);
}

// ---------------------------------------------------------------------------
// dump: the models and pose of the two-scan run, for comparison against the
// CPU restatement by the pytest wrapper
// ---------------------------------------------------------------------------
int dump(const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return 2;
  for (bool twoStep : {true, false}) {
    FeatureModelParams p = core_params();
    p.twoStepOptim = twoStep;
    sloam::sloam s;
    s.setFmParams(p);
    SloamInput t0 = core_input("t0"), t1 = core_input("t1");
    std::vector<Cylinder> lm;
    std::vector<Plane> pl;
    s.computeModels(t0, lm, pl);
    std::fprintf(f, "scan t0 twostep %d planes %zu trees %zu\n", (int)twoStep, pl.size(), lm.size());
    for (const auto& g : pl)
      std::fprintf(f, "plane %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", g.model.plane[0],
                   g.model.plane[1], g.model.plane[2], g.model.plane[3], g.model.centroid[0],
                   g.model.centroid[1], g.model.centroid[2]);
    for (const auto& c : lm)
      std::fprintf(f, "tree %zu %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", c.id, c.model.root[0],
                   c.model.root[1], c.model.root[2], c.model.ray[0], c.model.ray[1], c.model.ray[2],
                   c.model.radius);
    SloamOutput out;
    sloam::sloam s2;
    s2.setFmParams(p);
    s2.RunSloam(t0, out);
    t1.mapModels = out.tm;
    const bool ok = s2.RunSloam(t1, out);
    const Vector3& t = out.T_Map_Curr.translation();
    const Quat& q = out.T_Map_Curr.unit_quaternion();
    std::fprintf(f, "pose %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (int)twoStep, (int)ok,
                 t[0], t[1], t[2], q.x, q.y, q.z, q.w);
    std::fprintf(f, "matches");
    for (int m : out.matches) std::fprintf(f, " %d", m);
    std::fprintf(f, "\n");
  }
  // the pose problems of scan t1 against the map of scan t0, and the three
  // solves over them
  {
    const FeatureModelParams p = core_params();
    sloam::sloam s;
    s.setFmParams(p);
    SloamInput t0 = core_input("t0"), t1 = core_input("t1");
    SloamOutput out0;
    s.RunSloam(t0, out0);
    std::vector<Cylinder> lm;
    std::vector<Plane> pl;
    s.computeModels(t1, lm, pl);
    const auto tm = s.matchFeatures(t1.poseEstimate, lm, out0.tm, p.treeMatchThresh);
    const auto gm = s.matchFeatures(t1.poseEstimate, pl, s.getPrevGroundModel(), 1.0);
    for (const auto& m : tm) {
      const auto& o = m.object.model;
      std::fprintf(f, "tmatch %.9g %.9g %.9g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                   m.feature[0], m.feature[1], m.feature[2], o.root[0], o.root[1], o.root[2],
                   o.ray[0], o.ray[1], o.ray[2], o.radius);
    }
    for (const auto& m : gm) {
      const auto& o = m.object.model;
      std::fprintf(f, "gmatch %.9g %.9g %.9g %.17g %.17g %.17g %.17g\n", m.feature[0],
                   m.feature[1], m.feature[2], o.plane[0], o.plane[1], o.plane[2], o.plane[3]);
    }
    SE3 tf;
    const bool ok = s.OptimizePose(t1.poseEstimate, tm, gm, tf);
    const Vector3& t = tf.translation();
    const Quat& q = tf.unit_quaternion();
    std::fprintf(f, "joint %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (int)ok, t[0], t[1],
                 t[2], q.x, q.y, q.z, q.w);
    double xy[3], zrp[3];
    s.OptimizeXYYaw(t1.poseEstimate, true, tm, xy);
    s.OptimizeZRollPitch(t1.poseEstimate, true, gm, zrp);
    std::fprintf(f, "xyyaw %.17g %.17g %.17g\n", xy[0], xy[1], xy[2]);
    std::fprintf(f, "zrollpitch %.17g %.17g %.17g\n", zrp[0], zrp[1], zrp[2]);
  }
  std::fclose(f);
  return 0;
}

// the pose solves' arithmetic setting: FP64 by default, 0 or 1 accepted,
// anything else refused without a change
void pose_precision_setting() {
  sloam::sloam s;
  EXPECT_EQ(s.posePrecision(), 1.0);
  s.setPosePrecision(0);
  EXPECT_EQ(s.posePrecision(), 0.0);
  bool threw = false;
  try {
    s.setPosePrecision(2);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  EXPECT_EQ(s.posePrecision(), 0.0);
  s.setPosePrecision(1);
  EXPECT_EQ(s.posePrecision(), 1.0);
}

struct Case {
  const char* name;
  bool gpu;
  std::function<void()> fn;
};

}  // namespace

// seg::Segmentation on a labelled scan: <in> holds int32 H, W, n, destagger,
// float32 xyzi [n][4], uint8 labels [n]; <out> receives the mask (H*W
// bytes), then for each of ground (val 1), the dense tree cloud (val 255,
// dense), the non-ground non-tree points (val 0) and the two first again
// from a mask edited in a pixel of value 0 (so selected from the kept
// projection, not taken from the GPU partition): int32 count, float32
// points [count][4]
int segmentation(const char* in_path, const char* out_path) {
  FILE* f = std::fopen(in_path, "rb");
  if (!f) return 2;
  int32_t hdr[4];
  rd(f, hdr, 4, in_path);
  const int H = hdr[0], W = hdr[1], n = hdr[2];
  std::vector<float> xyzi(4 * (size_t)n);
  std::vector<uint8_t> labels((size_t)n);
  rd(f, xyzi.data(), xyzi.size(), in_path);
  rd(f, labels.data(), labels.size(), in_path);
  std::fclose(f);
  auto cloud = std::make_shared<CloudT>();
  cloud->points.resize((size_t)n);
  for (int i = 0; i < n; ++i)
    cloud->points[i] = PointT{xyzi[4 * (size_t)i], xyzi[4 * (size_t)i + 1],
                              xyzi[4 * (size_t)i + 2], xyzi[4 * (size_t)i + 3]};
  cloud->width = (uint32_t)W;
  cloud->height = (uint32_t)H;
  // as SLOAMNode makes it (sloamNode.cpp:81): fov, -fov, width, height, depth 1
  seg::Segmentation segm("unused.onnx", 22.5f, -22.5f, W, H, 1, hdr[3] != 0);
  segm.setPointLabels(labels);
  seg::MaskImage mask;
  segm.run(cloud, mask);
  CloudT::Ptr ground, tree, other, ground2, tree2;
  segm.maskCloud(cloud, mask, ground, 1);
  segm.maskCloud(cloud, mask, tree, 255, true);
  segm.maskCloud(cloud, mask, other, 0);
  seg::MaskImage edited = mask;
  for (auto& b : edited.bytes)
    if (b == 0) {
      b = 7;
      break;
    }
  segm.maskCloud(cloud, edited, ground2, 1);
  segm.maskCloud(cloud, edited, tree2, 255, true);
  EXPECT_EQ(tree->width, (double)W);
  EXPECT_EQ(tree->height, (double)H);
  EXPECT_TRUE(tree->is_dense && !ground->is_dense);
  EXPECT_EQ(ground->width, (double)ground->points.size());
  FILE* o = std::fopen(out_path, "wb");
  if (!o) return 2;
  std::fwrite(mask.data(), 1, mask.bytes.size(), o);
  for (const CloudT::Ptr& c : {ground, tree, other, ground2, tree2}) {
    const int32_t cnt = (int32_t)c->points.size();
    std::fwrite(&cnt, 4, 1, o);
    std::fwrite(c->points.data(), sizeof(PointT), c->points.size(), o);
  }
  std::fclose(o);
  return g_failures ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <fixture dir> [host|gpu|dump <out>]\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  const std::string mode = argc > 2 ? argv[2] : "all";
  if (mode == "seg") {
    if (argc < 4) return 2;
    try {
      return segmentation((g_dir + "/scan.bin").c_str(), argv[3]);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "seg: %s\n", e.what());
      return 1;
    }
  }
  if (mode == "dump") {
    if (argc < 4) return 2;
    try {
      return dump(argv[3]);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "dump: %s\n", e.what());
      return 1;
    }
  }
  const std::vector<Case> cases = {
      {"EulerAngles012", false, test_euler_angles},
      {"FromTwoVectors", false, test_from_two_vectors},
      {"AngleAxisQuaternion", false, test_angle_axis},
      {"SE3", false, test_se3},
      {"DetMath.Accuracy", false, test_detmath},
      {"CylinderTest.FromCylinderFit", false, cylinder_from_fit},
      {"MapManager", false, map_manager},
      {"PCD.ReadWrite", false, pcd_io},
      {"ROSMsg.WireFormat", false, ros_messages},
      {"Sloam.PosePrecision", false, pose_precision_setting},
      {"PlaneTest.Initalizes", true, plane_initializes},
      {"PlaneTest.DistanceToFeature", true, plane_distance_to_feature},
      {"PlaneTest.TranslateModel", true, plane_translate_model},
      {"CylinderTest.DistanceToModel", true, cylinder_distance_to_model},
      {"CylinderTest.DistanceToFeature", true, cylinder_distance_to_feature},
      {"CylinderTest.TranslateModel", true, cylinder_translate_model},
      {"InstanceTest.ReferenceClouds", true, instance_reference_clouds},
      {"InstanceTest.FindClusters", true, instance_find_clusters},
      {"SLOAMTest.FirstScan", true, core_first_scan},
      {"SLOAMTest.SecondScan", true, core_second_scan},
      {"SLOAMTest.SLOAMSucess", true, core_sloam_success},
      {"SLOAMTest.PoseOptimization", true, core_pose_optimization},
      {"SLOAMTest.ObjectAssociation", true, core_object_association},
  };
  int ran = 0;
  for (const auto& c : cases) {
    if ((mode == "host" && c.gpu) || (mode == "gpu" && !c.gpu)) continue;
    g_test = c.name;
    const int before = g_failures;
    std::printf("[ RUN      ] %s\n", c.name);
    std::fflush(stdout);
    try {
      c.fn();
    } catch (const std::exception& e) {
      std::fprintf(stderr, "[%s] exception: %s\n", c.name, e.what());
      ++g_failures;
    }
    std::printf("[ %s ] %s\n", g_failures == before ? "      OK" : " FAILED ", c.name);
    std::fflush(stdout);
    ++ran;
  }
  std::printf("%d tests, %d failures\n", ran, g_failures);
  return g_failures ? 1 : 0;
}
