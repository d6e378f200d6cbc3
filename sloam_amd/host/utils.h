// Geometric helpers with the reference's arithmetic
// (sloam/include/helpers/utils.h:9-52).
#pragma once

#include <cmath>

#include "definitions.h"

// squares in double, result rounded to float (utils.h:9)
SLOAM_HD inline float pow_2(const Scalar& x) { return (float)(x * x); }

SLOAM_HD inline float euclideanDist2D(const PointT& a, const PointT& b) {
  return sqrtf(pow_2(a.x - b.x) + pow_2(a.y - b.y));
}

// float running sums, divided by the count in double (utils.h:14-28)
inline PointT computeCentroid(const VectorType& features) {
  PointT c;
  for (const auto& p : features) {
    c.x += p.x;
    c.y += p.y;
    c.z += p.z;
  }
  c.x = (float)((double)c.x / (Scalar)features.size());
  c.y = (float)((double)c.y / (Scalar)features.size());
  c.z = (float)((double)c.z / (Scalar)features.size());
  return c;
}

// intersection of the ray root + t ray with the plane (utils.h:41-52): the
// denominator and t are floats, t must be >= 0.001
SLOAM_HD inline Vector3 rayPlaneIntersection(const Vector3& planeCentroid,
                                             const Vector3& planeNormal, const Vector3& rayOrigin,
                                             const Vector3& rayDirection) {
  const float denom = (float)planeNormal.dot(rayDirection);
  if (fabsf(denom) > 0.001f) {
    const float t = (float)((planeCentroid - rayOrigin).dot(planeNormal) / (double)denom);
    if (t >= 0.001f) return rayOrigin + (double)t * rayDirection;
  }
  return rayOrigin;
}
