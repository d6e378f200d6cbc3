// Model arithmetic of the reference's tree and ground objects and of
// computeModels' plane gate, as free functions over the drop-in's vector
// types.  The host classes (Cylinder, Plane, sloam::sloam) and the GPU
// keyframe estimator (sloam_amd/csrc/keyframe.hip) both call these, so both
// evaluate the same expressions in the same order.
#pragma once

#include <cmath>

#include "definitions.h"
#include "utils.h"

namespace sloam_geom {

// The parts of CylinderParameters (sloam/include/objects/cylinder.h:14-22)
// that association and the pose solve read.
struct CylModel {
  Vector3 root;
  Vector3 ray;
  double radius = 0.0;
};

// A cylinder fit record of one tree (sloam_cyl_plan model record).
struct CylFitRecord {
  bool ok = false;
  double root[3] = {0, 0, 0};
  double ray[3] = {0, 0, 0};
  double radius = -1.0;
};

// ---- plane (sloam/src/objects/plane.cpp) ------------------------------------
// Plane::distance(const PointT&) (plane.cpp:136-151)
SLOAM_HD inline double plane_point_distance(const Vector4& plane, const PointT& p) {
  return fabs(plane[0] * p.x + plane[1] * p.y + plane[2] * p.z + plane[3]) /
         plane.head3().norm();
}

// Plane::distance(const PlaneParameters&) (plane.cpp:131-134)
SLOAM_HD inline double plane_model_distance(const Vector3& centroid, const Vector3& tgt_centroid) {
  return (centroid - tgt_centroid).norm();
}

// Plane::project (plane.cpp:153-175): coefficients by the inverse transpose
// of the transform, centroid by the transform
SLOAM_HD inline void plane_project(const SE3& tf, const Vector4& plane, const Vector3& centroid,
                                   Vector4& plane_out, Vector3& centroid_out) {
  plane_out = tf.matrix().inverse().transpose() * plane;
  centroid_out = tf * centroid;
}

// computeModels' gate of a ground cell (sloam/src/core/sloam.cpp:401-408):
// the normal taken to the map frame is within 0.1 rad of vertical in every
// Euler angle (or of its opposite), and the centroid lies below the sensor.
// tfm = poseEstimate.inverse().matrix().transpose().
SLOAM_HD inline bool plane_accepted(const SE3& pose, const Matrix4& tfm, const Vector4& plane,
                                    const Vector3& centroid) {
  const Vector4 normal = tfm * plane;
  const Quat q_rot = Quat::FromTwoVectors(Vector3(0, 0, 1), normal.head3());
  double R[3][3];
  q_rot.toRotationMatrix(R);
  const Vector3 angles = eulerAngles012(R);
  const bool angleCheck = (angles[0] < 0.1 && angles[1] < 0.1 && angles[2] < 0.1) ||
                          (M_PI - fabs(angles[0]) < 0.1 && M_PI - fabs(angles[1]) < 0.1 &&
                           M_PI - fabs(angles[2]) < 0.1);
  const bool heightCheck = (pose * centroid)[2] < pose.translation()[2];
  return angleCheck && heightCheck;
}

// ---- cylinder (sloam/src/objects/cylinder.cpp) -----------------------------
SLOAM_HD inline float squared_distance_f(const PointT& a, const PointT& b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

// cylinder.cpp:5-13: vertex [1] within maxLidarDist (2-D) of the plane centroid
SLOAM_HD inline bool cyl_within_range(const PointT& vertex1, const Vector3& plane_centroid,
                                      double maxLidarDist) {
  PointT c;
  c.x = (float)plane_centroid(0);
  c.y = (float)plane_centroid(1);
  c.z = (float)plane_centroid(2);
  return euclideanDist2D(vertex1, c) < maxLidarDist;
}

// cylinder.cpp:100-113: a tree with vertices [1] and [n-2] at least
// sqrt(1.5) m apart has a model
SLOAM_HD inline bool cyl_needs_line_fit(int n_vertices, const PointT& v1, const PointT& v_n2) {
  if (n_vertices < 3) return false;
  return !(squared_distance_f(v1, v_n2) < 1.5f);
}

// The model from a cylinder fit of the tree's points: the fitted axis
// pointing up, the bottom vertex's foot on it as the root, the fitted radius
// with the default radius as its floor; radius -1 when the fit failed.
SLOAM_HD inline void cyl_model_from_fit(const PointT& bottom, const CylFitRecord& fit,
                                        double defaultTreeRadius, CylModel& m) {
  m.root = Vector3(bottom.x, bottom.y, bottom.z);
  if (!fit.ok || !(fit.radius > 0.0)) {
    m.radius = -1;
    return;
  }
  const double sgn = fit.ray[2] < 0.0 ? -1.0 : 1.0;
  m.ray = Vector3(sgn * fit.ray[0], sgn * fit.ray[1], sgn * fit.ray[2]);
  const Vector3 c(fit.root[0], fit.root[1], fit.root[2]);
  m.root = c + (m.root - c).dot(m.ray) * m.ray;
  m.radius = fit.radius < defaultTreeRadius ? defaultTreeRadius : fit.radius;
}

// Cylinder::groundBasedRoot (cylinder.cpp:41-55): the root moves to the
// axis' intersection with the ground plane when within 2 m of it
SLOAM_HD inline bool cyl_ground_based_root(const Vector4& g, const Vector3& g_centroid,
                                           CylModel& m) {
  const float dist = (float)(fabs(g[0] * m.root(0) + g[1] * m.root(1) + g[2] * m.root(2) + g[3]) /
                             g.head3().norm());
  if (dist < 2.0) {
    m.root = rayPlaneIntersection(g_centroid, g.head3(), m.root, m.ray);
    return true;
  }
  return false;
}

// The rules of Cylinder::filter a model fails (bits; 0 when it passes).
enum CylFilterFail : int { kCylNoModel = 1, kCylRadiusCap = 2, kCylAxisTilt = 4 };

// Cylinder::filter (cylinder.cpp:57-73): radius below the cap and the axis
// within maxAxisTheta degrees of the ground normal (either direction); the
// failing rules
SLOAM_HD inline int cyl_filter_fails(const CylModel& m, double maxTreeRadius,
                                     double maxAxisTheta, const Vector4& g) {
  if (m.radius == -1) return kCylNoModel;
  const Vector3 up = g.head3();
  const double theta = (180 / PIDEF) * sloam_dm::acos(m.ray.dot(up) / (m.ray.norm() * up.norm()));
  int fails = 0;
  if (!(m.radius < maxTreeRadius)) fails |= kCylRadiusCap;
  if (!((theta <= maxAxisTheta) || (theta >= 180 - maxAxisTheta))) fails |= kCylAxisTilt;
  return fails;
}

SLOAM_HD inline bool cyl_filter(const CylModel& m, double maxTreeRadius, double maxAxisTheta,
                                const Vector4& g) {
  return cyl_filter_fails(m, maxTreeRadius, maxAxisTheta, g) == 0;
}

// The validity the Cylinder constructor leaves (cylinder.cpp:15-27) once the
// model is computed: ground-based root, root away from the origin, positive
// radius, filter.
SLOAM_HD inline bool cyl_valid(CylModel& m, const Vector4& g, const Vector3& g_centroid,
                               double maxTreeRadius, double maxAxisTheta) {
  const bool validZ = cyl_ground_based_root(g, g_centroid, m);
  const bool validNorm = m.root.norm() > 0.01;
  const bool validRadius = m.radius > 0.0;
  const bool validTree = cyl_filter(m, maxTreeRadius, maxAxisTheta, g);
  return validZ && validRadius && validTree && validNorm;
}

// Cylinder::distance(const CylinderParameters&) (cylinder.cpp:175-194): the
// mean gap between the two axes at heights 0, 3 and 6
SLOAM_HD inline double cyl_model_distance(const Vector3& root, const Vector3& ray,
                                          const Vector3& tgt_root, const Vector3& tgt_ray) {
  const double heights[3] = {0.0, 3.0, 6.0};
  double d = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double h = heights[i];
    const double st = (h - root[2]) / ray[2];
    const Vector3 mp = root + st * ray;
    const double tt = (h - tgt_root[2]) / tgt_ray[2];
    const Vector3 tp = tgt_root + tt * tgt_ray;
    d += (mp - tp).norm();
  }
  return d / 3.0;
}

// Cylinder::project (cylinder.cpp:205-210): root and root + ray transformed
SLOAM_HD inline void cyl_project(const SE3& tf, const Vector3& root, const Vector3& ray,
                                 Vector3& root_out, Vector3& ray_out) {
  Vector3 other = root + ray;
  root_out = tf * root;
  other = tf * other;
  ray_out = other - root_out;
}

}  // namespace sloam_geom
