// TEST INFRASTRUCTURE ONLY -- CPU restatement of SLOAM's per-keyframe
// estimator, the checker of the C++ drop-in (sloam_amd/host/sloam.cpp) and
// the CPU baseline of the keyframe stream.
//
// Restates, in FP64 on the host:
//   * sloam::computeModels (sloam/src/core/sloam.cpp:388-436): the ground
//     cells' acceptance (Euler angles of the rotation taking +z to the cell
//     normal in the map frame, sloam.cpp:395-406; height check :408) and each
//     tree's nearest accepted plane (:418-432);
//   * the Cylinder constructor (sloam/src/objects/cylinder.cpp:3-28) with
//     computeModel (:75-173, the sample-consensus line through the vertex
//     medians from oracle_line_ransac, the "upper median" radius, the default
//     radius floor), groundBasedRoot (:41-55, rayPlaneIntersection of
//     sloam/include/helpers/utils.h:41-52) and filter (:57-73); and the build's
//     cylinder-fit variant (a tree's axis, root and radius from a point-level
//     cylinder fit of its points; no counterpart in the reference);
//   * Cylinder::distance / project (cylinder.cpp:175-210), Plane::distance /
//     project (sloam/src/objects/plane.cpp:131-176);
//   * matchFeatures, matchModels (sloam.cpp:258-328);
//   * TwoStepOptimizePose / OptimizeXYYaw / OptimizeZRollPitch / OptimizePose
//     (sloam.cpp:33-255) through oracle_pose_lm_batch (oracle_pose.cpp);
//   * RunSloam (sloam.cpp:453-532) with its state (first scan, previous
//     ground planes).
// Independent of the drop-in's code: Eigen's Quaternion::FromTwoVectors is
// restated as the minimal rotation between two vectors (Rodrigues' formula),
// and eulerAngles(0, 1, 2) as the x-y-z decomposition mapped onto Eigen's
// documented range (first angle in [0, pi]).
//
// Where the reference is undefined the drop-in's definitions are taken:
// matchFeatures against an empty map list returns no match (the reference
// reads mapObjects[0]); minPlanes = 0.1 x radii bins x theta bins (the
// reference computes it from unset parameters, sloam.cpp:11); a tree with
// fewer than 3 vertices gets no model (the reference indexes [1] and [2]).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

extern "C" int oracle_line_ransac(const float* P, int n, double thr, int max_iter, double prob,
                                  uint32_t seed, int optimize, float coeff[6], uint8_t* inl,
                                  int32_t stats[4]);
extern "C" int oracle_pose_lm_batch(const float* pts, const int32_t* off, int n_problems,
                                    const double* cyl, const double* pl, const double* pose_in,
                                    const int32_t* free_masks, const int32_t cfg_i[3],
                                    const double cfg_d[6], double* pose_out, double* params_out,
                                    int32_t* status, double* cost, int n_threads);

namespace oracle {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kPiDef = 3.14159265;  // PIDEF of sloam/include/helpers/definitions.h

// params[] layout of oracle_sloam_create
struct Params {
  double maxLidarDist, maxTreeRadius, maxAxisTheta, defaultTreeRadius;
  double featuresPerTree, numGroundFeatures;
  int radiiBins, thetaBins;
  double treeMatchThresh, addNewTreeThreshDist, minTreeModels, minGroundModels;
  bool twoStep;
};

struct Pose {
  double t[3];
  double q[4];  // x, y, z, w
};

void rot(const double q[4], double R[3][3]) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
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

void apply(const Pose& T, const double p[3], double out[3]) {
  double R[3][3];
  rot(T.q, R);
  for (int i = 0; i < 3; ++i) out[i] = R[i][0] * p[0] + R[i][1] * p[1] + R[i][2] * p[2] + T.t[i];
}

double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm3(const double a[3]) { return std::sqrt(dot3(a, a)); }

// Rotation matrix of the smallest rotation taking unit direction of a onto
// that of b: R = I + [k]x + [k]x^2 / (1 + c), k = a x b, c = a . b
// (Rodrigues).  Antiparallel input (c = -1) has no unique answer; ground
// normals never reach it (they point up and the pose tilts little).
void rotation_between(const double a_in[3], const double b_in[3], double R[3][3]) {
  const double na = norm3(a_in), nb = norm3(b_in);
  const double a[3] = {a_in[0] / na, a_in[1] / na, a_in[2] / na};
  const double b[3] = {b_in[0] / nb, b_in[1] / nb, b_in[2] / nb};
  const double k[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                       a[0] * b[1] - a[1] * b[0]};
  const double c = dot3(a, b);
  const double K[3][3] = {{0, -k[2], k[1]}, {k[2], 0, -k[0]}, {-k[1], k[0], 0}};
  const double f = 1.0 / (1.0 + c);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double k2 = 0;
      for (int m = 0; m < 3; ++m) k2 += K[i][m] * K[m][j];
      R[i][j] = (i == j ? 1.0 : 0.0) + K[i][j] + f * k2;
    }
}

// Angles (a, b, c) with R = Rx(a) Ry(b) Rz(c), in the range Eigen's
// eulerAngles(0, 1, 2) documents: a in [0, pi], b and c in [-pi, pi].
// Plain decomposition: b = asin(R02) in [-pi/2, pi/2], a = atan2(-R12, R22),
// c = atan2(-R01, R00); the same rotation is also (a + pi, pi - b, c + pi),
// which is taken when a < 0.
void euler_xyz(const double R[3][3], double e[3]) {
  const double s = std::sqrt(R[0][0] * R[0][0] + R[0][1] * R[0][1]);
  double a = std::atan2(-R[1][2], R[2][2]);
  double b = std::atan2(R[0][2], s);
  double c = std::atan2(-R[0][1], R[0][0]);
  if (a < 0) {
    a += kPi;
    b = (b >= 0 ? kPi : -kPi) - b;
    c = c > 0 ? c - kPi : c + kPi;
  }
  e[0] = a;
  e[1] = b;
  e[2] = c;
}

struct PlaneM {
  double n[4];  // plane coefficients (normal, offset)
  double c[3];  // centroid
  std::vector<float> feat;  // 4 floats per feature point
};

struct CylM {
  double root[3], ray[3], radius;
  int id;
  bool valid;
  std::vector<float> feat;  // 4 floats per feature point (sensor frame)
};

struct Vtx {
  int id;
  double radius;
  float xyz[3];
  const float* pts;
  int npts;
};

// Plane::distance(const PointT&) (plane.cpp:136-151)
double plane_point_distance(const PlaneM& p, const float q[3]) {
  return std::fabs(p.n[0] * q[0] + p.n[1] * q[1] + p.n[2] * q[2] + p.n[3]) /
         std::sqrt(p.n[0] * p.n[0] + p.n[1] * p.n[1] + p.n[2] * p.n[2]);
}

// euclideanDist2D (utils.h:9-12): float differences, squares in double
// stored as float, float sum and square root
float dist2d(const float a[3], const float b[3]) {
  const double dx = (double)(a[0] - b[0]), dy = (double)(a[1] - b[1]);
  const float sx = (float)(dx * dx), sy = (float)(dy * dy);
  return std::sqrt(sx + sy);
}

// Cylinder(vertices, plane, params) and its variant from a cylinder fit
// (cfit != nullptr: ok flag, then root[3], ray[3], radius).  lfit: the line
// fit of the medians (ok, inliers, coeff[6]); computed here when null.
CylM make_cylinder(const std::vector<Vtx>& v, const PlaneM& g, const Params& prm,
                   const int32_t* lfit_int, const float* lfit_coeff, const int32_t* cfit_ok,
                   const double* cfit) {
  CylM m;
  m.valid = false;
  m.radius = 0;
  m.id = 0;
  for (int i = 0; i < 3; ++i) m.root[i] = m.ray[i] = 0;
  const float centroid[3] = {(float)g.c[0], (float)g.c[1], (float)g.c[2]};
  if (!(dist2d(v[1].xyz, centroid) < prm.maxLidarDist)) return m;
  // computeModel
  const Vtx& first = v[2];
  std::vector<float> valid_radii;
  for (const Vtx& x : v) {
    for (int p = 0; p < x.npts; ++p) {
      m.feat.push_back(x.pts[4 * p]);
      m.feat.push_back(x.pts[4 * p + 1]);
      m.feat.push_back(x.pts[4 * p + 2]);
      m.feat.push_back((float)first.id);
    }
    if (x.npts > 3) valid_radii.push_back((float)x.radius);
  }
  m.id = first.id;
  const float* bottom = v[1].xyz;
  const float* top = v[v.size() - 2].xyz;
  for (int i = 0; i < 3; ++i) m.root[i] = bottom[i];
  const float ddx = bottom[0] - top[0], ddy = bottom[1] - top[1], ddz = bottom[2] - top[2];
  const float sq = ddx * ddx + ddy * ddy + ddz * ddz;
  bool have_model = true;
  if (sq < 1.5f) {
    m.radius = -1;
    have_model = false;
  } else if (cfit) {
    if (!cfit_ok[0] || !(cfit[6] > 0.0)) {
      m.radius = -1;
      have_model = false;
    } else {
      const double sgn = cfit[5] < 0.0 ? -1.0 : 1.0;
      for (int i = 0; i < 3; ++i) m.ray[i] = sgn * cfit[3 + i];
      const double rel[3] = {m.root[0] - cfit[0], m.root[1] - cfit[1], m.root[2] - cfit[2]};
      const double h = dot3(rel, m.ray);
      for (int i = 0; i < 3; ++i) m.root[i] = cfit[i] + h * m.ray[i];
      m.radius = cfit[6] < prm.defaultTreeRadius ? prm.defaultTreeRadius : cfit[6];
    }
  } else {
    int32_t stats[4] = {0, 0, 0, 0};
    float coeff[6] = {0, 0, 0, 0, 0, 0};
    if (lfit_int) {
      stats[0] = lfit_int[0];
      stats[1] = lfit_int[1];
      std::memcpy(coeff, lfit_coeff, sizeof(coeff));
    } else {
      std::vector<float> med;
      for (const Vtx& x : v) med.insert(med.end(), x.xyz, x.xyz + 3);
      oracle_line_ransac(med.data(), (int)v.size(), 0.25, 50, 0.99, 12345u, 1, coeff, nullptr,
                         stats);
    }
    if (!stats[0] || stats[1] == 0) {
      m.radius = -1;
      have_model = false;
    } else {
      for (int i = 0; i < 3; ++i) m.ray[i] = coeff[3 + i];
      if (!valid_radii.empty()) {
        std::sort(valid_radii.begin(), valid_radii.end());
        int d = 0;
        while (d < (int)valid_radii.size() && !(valid_radii[d] > 0)) ++d;
        const int middle = std::min((int)valid_radii.size() - 1,
                                    (d + 1) + (int)((valid_radii.size() - d) / 2));
        m.radius = valid_radii[middle];
        if (m.radius == 0)
          m.radius = -1;
        else if (m.radius < prm.defaultTreeRadius)
          m.radius = prm.defaultTreeRadius;
      } else {
        m.radius = -1;
      }
    }
  }
  if (have_model) m.feat.resize(4 * (size_t)prm.featuresPerTree, 0.f);  // zero-padded
  // groundBasedRoot: float distance, ray-plane intersection with float
  // denominator and parameter
  bool validZ = false;
  {
    const double nn = std::sqrt(g.n[0] * g.n[0] + g.n[1] * g.n[1] + g.n[2] * g.n[2]);
    const float dist =
        (float)(std::fabs(g.n[0] * m.root[0] + g.n[1] * m.root[1] + g.n[2] * m.root[2] + g.n[3]) /
                nn);
    if (dist < 2.0) {
      const float denom = (float)(g.n[0] * m.ray[0] + g.n[1] * m.ray[1] + g.n[2] * m.ray[2]);
      if (std::fabs(denom) > 0.001f) {
        const double rel[3] = {g.c[0] - m.root[0], g.c[1] - m.root[1], g.c[2] - m.root[2]};
        const float t = (float)(dot3(rel, g.n) / denom);
        if (t >= 0.001f)
          for (int i = 0; i < 3; ++i) m.root[i] = m.root[i] + (double)t * m.ray[i];
      }
      validZ = true;
    }
  }
  const bool validNorm = norm3(m.root) > 0.01;
  const bool validRadius = m.radius > 0.0;
  bool validTree = false;
  if (m.radius != -1) {
    const double theta =
        (180 / kPiDef) * std::acos(dot3(m.ray, g.n) / (norm3(m.ray) * norm3(g.n)));
    validTree = (m.radius < prm.maxTreeRadius) &&
                ((theta <= prm.maxAxisTheta) || (theta >= 180 - prm.maxAxisTheta));
  }
  m.valid = validZ && validRadius && validTree && validNorm;
  return m;
}

// Cylinder::distance(const CylinderParameters&) (cylinder.cpp:175-194)
double cyl_model_distance(const CylM& a, const double root[3], const double ray[3]) {
  double d = 0;
  for (const double h : {0.0, 3.0, 6.0}) {
    const double sa = (h - a.root[2]) / a.ray[2];
    const double tb = (h - root[2]) / ray[2];
    double e[3];
    for (int i = 0; i < 3; ++i) e[i] = (a.root[i] + sa * a.ray[i]) - (root[i] + tb * ray[i]);
    d += norm3(e);
  }
  return d / 3.0;
}

void project_cyl(const Pose& T, const CylM& c, double root[3], double ray[3]) {
  const double other[3] = {c.root[0] + c.ray[0], c.root[1] + c.ray[1], c.root[2] + c.ray[2]};
  double o2[3];
  apply(T, c.root, root);
  apply(T, other, o2);
  for (int i = 0; i < 3; ++i) ray[i] = o2[i] - root[i];
}

// Plane::project model part: n' = [R n; d - t . (R n)], c' = R c + t
void project_plane_model(const Pose& T, const PlaneM& p, double n[4], double c[3]) {
  double R[3][3];
  rot(T.q, R);
  for (int i = 0; i < 3; ++i) n[i] = R[i][0] * p.n[0] + R[i][1] * p.n[1] + R[i][2] * p.n[2];
  n[3] = p.n[3] - (T.t[0] * n[0] + T.t[1] * n[1] + T.t[2] * n[2]);
  apply(T, p.c, c);
}

void project_plane(const Pose& T, PlaneM& p) {
  double n[4], c[3];
  project_plane_model(T, p, n, c);
  std::memcpy(p.n, n, sizeof(n));
  std::memcpy(p.c, c, sizeof(c));
  double R[3][3];
  rot(T.q, R);
  for (size_t k = 0; k + 4 <= p.feat.size(); k += 4) {
    float* f = &p.feat[k];
    if (!std::isfinite(f[0]) || !std::isfinite(f[1]) || !std::isfinite(f[2])) continue;
    const double x = f[0], y = f[1], z = f[2];
    double o[3];
    for (int i = 0; i < 3; ++i) o[i] = R[i][0] * x + R[i][1] * y + R[i][2] * z + T.t[i];
    f[0] = (float)o[0];
    f[1] = (float)o[1];
    f[2] = (float)o[2];
  }
}

void project_cyl_inplace(const Pose& T, CylM& c) {
  double r[3], a[3];
  project_cyl(T, c, r, a);
  std::memcpy(c.root, r, sizeof(r));
  std::memcpy(c.ray, a, sizeof(a));
}

// ceres::QuaternionToAngleAxis (q stored x, y, z, w)
void quat_to_aa(const double q[4], double aa[3]) {
  const double s2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  if (s2 > 0.0) {
    const double st = std::sqrt(s2);
    const double c = q[3];
    const double th2 = 2.0 * ((c < 0.0) ? std::atan2(-st, -c) : std::atan2(st, c));
    for (int i = 0; i < 3; ++i) aa[i] = q[i] * (th2 / st);
  } else {
    for (int i = 0; i < 3; ++i) aa[i] = 2.0 * q[i];
  }
}

void aa_to_quat(const double aa[3], double q[4]) {
  const double t2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  if (t2 > 0.0) {
    const double th = std::sqrt(t2);
    const double k = std::sin(0.5 * th) / th;
    for (int i = 0; i < 3; ++i) q[i] = aa[i] * k;
    q[3] = std::cos(0.5 * th);
  } else {
    for (int i = 0; i < 3; ++i) q[i] = 0.5 * aa[i];
    q[3] = 1.0;
  }
}

struct Match {
  float p[3];  // sensor-frame feature
  int obj;     // map object index
};

struct Problem {
  std::vector<float> pts;     // 4 floats per residual (x, y, z, id bits)
  std::vector<double> cyl;    // 8 per map cylinder
  std::vector<double> plane;  // 4 per plane
};

void push(Problem& pr, const float p[3], int32_t id) {
  pr.pts.push_back(p[0]);
  pr.pts.push_back(p[1]);
  pr.pts.push_back(p[2]);
  float bits;
  std::memcpy(&bits, &id, 4);
  pr.pts.push_back(bits);
}

// one LM solve; mode 0: joint quaternion (all free), 1: angle-axis subset
bool solve(const Problem& pr, const Pose& x0, int param, int32_t free_mask, double pose_out[7],
           double params_out[6], int32_t* iterations) {
  const int32_t off[2] = {0, (int32_t)(pr.pts.size() / 4)};
  const int32_t cfg_i[3] = {50, 1, param};
  const double cfg_d[6] = {0.1, 1e-6, 1e-10, 1e-8, 1e4, 1e-3};
  const double pin[7] = {x0.t[0], x0.t[1], x0.t[2], x0.q[0], x0.q[1], x0.q[2], x0.q[3]};
  int32_t st[2];
  double cost[2];
  const double dummy[8] = {0, 0, 0, 0, 0, 1, 0, 0};
  oracle_pose_lm_batch(pr.pts.data(), off, 1, pr.cyl.empty() ? dummy : pr.cyl.data(),
                       pr.plane.empty() ? dummy : pr.plane.data(), pin, &free_mask, cfg_i, cfg_d,
                       pose_out, params_out, st, cost, 1);
  *iterations = st[1];
  return st[0] == 0;
}

struct State {
  Params p;
  bool first = true;
  std::vector<PlaneM> prev;
};

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

// params: maxLidarDist, maxTreeRadius, maxAxisTheta, defaultTreeRadius,
// featuresPerTree, numGroundFeatures, groundRadiiBins, groundThetaBins,
// treeMatchThresh, AddNewTreeThreshDist, minTreeModels, minGroundModels,
// twoStepOptim (0 / 1)
void* oracle_sloam_create(const double* params) {
  State* s = new State();
  Params& p = s->p;
  p.maxLidarDist = params[0];
  p.maxTreeRadius = params[1];
  p.maxAxisTheta = params[2];
  p.defaultTreeRadius = params[3];
  p.featuresPerTree = params[4];
  p.numGroundFeatures = params[5];
  p.radiiBins = (int)params[6];
  p.thetaBins = (int)params[7];
  p.treeMatchThresh = params[8];
  p.addNewTreeThreshDist = params[9];
  p.minTreeModels = params[10];
  p.minGroundModels = params[11];
  p.twoStep = params[12] != 0.0;
  return s;
}

void oracle_sloam_destroy(void* h) { delete static_cast<State*>(h); }

// Replaces the state: first-scan flag and the previous ground planes (8
// doubles each: plane[4], centroid[3], unused; their features do not take
// part in association).
int oracle_sloam_set_state(void* h, int first, int n_prev, const double* prev) {
  State* s = static_cast<State*>(h);
  s->first = first != 0;
  s->prev.assign((size_t)n_prev, PlaneM());
  for (int k = 0; k < n_prev; ++k) {
    std::memcpy(s->prev[k].n, prev + 8 * k, 4 * sizeof(double));
    std::memcpy(s->prev[k].c, prev + 8 * k + 4, 3 * sizeof(double));
  }
  return 0;
}

// One RunSloam call.
//   pose: estimate (t, q x y z w)
//   cells: n_cells ground cells, 8 doubles each (plane[4], centroid[3], -),
//          cell_valid, nf feature points of 4 floats each
//   trees: tree_off (n_trees + 1) into vertices; per vertex vtx_int (id,
//          valid), vtx_f64 (radius, x, y, z), vtx_pt_off (n_vtx + 1) into pts
//          (4 floats each)
//   tree_mode: 0 line fits computed here (oracle_line_ransac), 1 line fits
//          given (line_int: ok, inliers; line_coeff: 6 floats), 2 cylinder
//          fits given (cyl_ok; cyl_fit: root[3], ray[3], radius)
//   map: n_map cylinders, 7 doubles each (root, ray, radius)
// Outputs:
//   info[12]: returned value, early return (no state change), treeCheck,
//          groundCheck, tree feature matches, plane feature matches, trees
//          kept, planes kept, first scan, pose solve converged, LM iterations
//          of the tree (or joint) solve and of the ground solve (-1: not run)
//   pose_out[7] = T_Map_Curr, delta_out[7] = T_Delta
//   tm_id / tm (7 doubles, projected) / matches per kept tree (capacity
//          n_trees); tree_valid per input tree
//   planes (8 doubles each, projected; capacity n_cells)
int oracle_sloam_run(void* h, const double pose[7], int n_cells, const double* cell_model,
                     const int32_t* cell_valid, int nf, const float* cell_feat, int n_trees,
                     const int32_t* tree_off, const int32_t* vtx_int, const double* vtx_f64,
                     const int32_t* vtx_pt_off, const float* pts, int tree_mode,
                     const int32_t* line_int, const float* line_coeff, const int32_t* cyl_ok,
                     const double* cyl_fit, int n_map, const double* map, int32_t info[12],
                     double pose_out[7], double delta_out[7], int32_t* tm_id, double* tm,
                     int32_t* matches, int32_t* tree_valid, double* planes_out) {
  State* s = static_cast<State*>(h);
  const Params& prm = s->p;
  for (int i = 0; i < 10; ++i) info[i] = 0;
  info[10] = info[11] = -1;
  Pose est;
  std::memcpy(est.t, pose, 3 * sizeof(double));
  std::memcpy(est.q, pose + 3, 4 * sizeof(double));
  double Rest[3][3];
  rot(est.q, Rest);
  // --- computeModels: ground cells
  std::vector<PlaneM> planes;
  for (int c = 0; c < n_cells; ++c) {
    if (!cell_valid[c]) continue;
    PlaneM p;
    std::memcpy(p.n, cell_model + 8 * c, 4 * sizeof(double));
    std::memcpy(p.c, cell_model + 8 * c + 4, 3 * sizeof(double));
    p.feat.assign(cell_feat + 4 * (size_t)c * nf, cell_feat + 4 * ((size_t)c + 1) * nf);
    // the normal in the map frame: the top rows of T^-T pi = R n
    double nm[3];
    for (int i = 0; i < 3; ++i) nm[i] = Rest[i][0] * p.n[0] + Rest[i][1] * p.n[1] + Rest[i][2] * p.n[2];
    const double z[3] = {0, 0, 1};
    double R[3][3], e[3];
    rotation_between(z, nm, R);
    euler_xyz(R, e);
    const bool angleCheck = (e[0] < 0.1 && e[1] < 0.1 && e[2] < 0.1) ||
                            (kPi - std::fabs(e[0]) < 0.1 && kPi - std::fabs(e[1]) < 0.1 &&
                             kPi - std::fabs(e[2]) < 0.1);
    double cm[3];
    apply(est, p.c, cm);
    const bool heightCheck = cm[2] < est.t[2];
    if (angleCheck && heightCheck) planes.push_back(std::move(p));
  }
  // --- computeModels: trees (none without planes)
  std::vector<CylM> trees;
  for (int t = 0; t < n_trees; ++t) tree_valid[t] = 0;
  if (!planes.empty()) {
    for (int t = 0; t < n_trees; ++t) {
      std::vector<Vtx> v;
      for (int k = tree_off[t]; k < tree_off[t + 1]; ++k) {
        Vtx x;
        x.id = vtx_int[2 * k];
        x.radius = vtx_f64[4 * k];
        for (int i = 0; i < 3; ++i) x.xyz[i] = (float)vtx_f64[4 * k + 1 + i];
        x.pts = pts + 4 * (size_t)vtx_pt_off[k];
        x.npts = vtx_pt_off[k + 1] - vtx_pt_off[k];
        v.push_back(x);
      }
      if (v.size() < 3) continue;
      double best = 100000;
      size_t bi = 0;
      for (size_t i = 0; i < planes.size(); ++i) {
        const double d = plane_point_distance(planes[i], v[1].xyz);
        if (d < best) {
          best = d;
          bi = i;
        }
      }
      CylM c = make_cylinder(v, planes[bi], prm, tree_mode == 1 ? line_int + 2 * t : nullptr,
                             tree_mode == 1 ? line_coeff + 6 * t : nullptr,
                             tree_mode == 2 ? cyl_ok + t : nullptr,
                             tree_mode == 2 ? cyl_fit + 7 * t : nullptr);
      if (c.valid) {
        tree_valid[t] = 1;
        trees.push_back(std::move(c));
      }
    }
  }
  std::vector<int> match_idx(trees.size(), -1);
  auto write_out = [&](const Pose& cur, const Pose& delta) {
    std::memcpy(pose_out, cur.t, 3 * sizeof(double));
    std::memcpy(pose_out + 3, cur.q, 4 * sizeof(double));
    std::memcpy(delta_out, delta.t, 3 * sizeof(double));
    std::memcpy(delta_out + 3, delta.q, 4 * sizeof(double));
    for (size_t k = 0; k < trees.size(); ++k) {
      tm_id[k] = trees[k].id;
      std::memcpy(tm + 7 * k, trees[k].root, 3 * sizeof(double));
      std::memcpy(tm + 7 * k + 3, trees[k].ray, 3 * sizeof(double));
      tm[7 * k + 6] = trees[k].radius;
      matches[k] = match_idx[k];
    }
    for (size_t k = 0; k < planes.size(); ++k) {
      std::memcpy(planes_out + 8 * k, planes[k].n, 4 * sizeof(double));
      std::memcpy(planes_out + 8 * k + 4, planes[k].c, 3 * sizeof(double));
      planes_out[8 * k + 7] = 0;
    }
    info[6] = (int32_t)trees.size();
    info[7] = (int32_t)planes.size();
  };
  Pose identity = {{0, 0, 0}, {0, 0, 0, 1}};
  if (s->first) {
    for (auto& p : planes) project_plane(est, p);
    for (auto& c : trees) project_cyl_inplace(est, c);
    info[0] = 1;
    info[8] = 1;
    write_out(est, identity);
    s->prev = planes;
    s->first = false;
    return 0;
  }
  if (n_map == 0 || planes.empty() || trees.empty()) {
    info[0] = 0;
    info[1] = 1;
    info[6] = (int32_t)trees.size();
    info[7] = (int32_t)planes.size();
    return 0;
  }
  // --- matchFeatures: trees against the map, planes against the previous
  // planes; the nearest model of the object projected by the estimate
  // (strict <, first wins, initialised at threshold + 100)
  std::vector<Match> tmatch, gmatch;
  for (const CylM& c : trees) {
    double r[3], a[3];
    project_cyl(est, c, r, a);
    double best = prm.treeMatchThresh + 100;
    int bi = 0;
    for (int k = 0; k < n_map; ++k) {
      CylM mm;
      std::memcpy(mm.root, map + 7 * k, 3 * sizeof(double));
      std::memcpy(mm.ray, map + 7 * k + 3, 3 * sizeof(double));
      const double d = cyl_model_distance(mm, r, a);
      if (d < best) {
        best = d;
        bi = k;
      }
    }
    if (best < prm.treeMatchThresh)
      for (size_t f = 0; f < c.feat.size(); f += 4) tmatch.push_back({{c.feat[f], c.feat[f + 1], c.feat[f + 2]}, bi});
  }
  if (!s->prev.empty())
    for (const PlaneM& p : planes) {
      double n[4], cc[3];
      project_plane_model(est, p, n, cc);
      double best = 1.0 + 100;
      int bi = 0;
      for (size_t k = 0; k < s->prev.size(); ++k) {
        const double e[3] = {s->prev[k].c[0] - cc[0], s->prev[k].c[1] - cc[1], s->prev[k].c[2] - cc[2]};
        const double d = norm3(e);
        if (d < best) {
          best = d;
          bi = (int)k;
        }
      }
      if (best < 1.0)
        for (size_t f = 0; f < p.feat.size(); f += 4) gmatch.push_back({{p.feat[f], p.feat[f + 1], p.feat[f + 2]}, bi});
    }
  const double minPlanes = (prm.radiiBins * prm.thetaBins) * 0.1;
  const bool treeCheck = trees.size() > prm.minTreeModels && tmatch.size() > 5 * prm.featuresPerTree;
  const bool groundCheck = gmatch.size() > prm.minGroundModels && planes.size() > minPlanes;
  info[2] = treeCheck;
  info[3] = groundCheck;
  info[4] = (int32_t)tmatch.size();
  info[5] = (int32_t)gmatch.size();
  // residual problems: consecutive matches of one map object share its entry
  auto tree_problem = [&](Problem& pr) {
    int last = -1;
    for (const Match& m : tmatch) {
      if (m.obj != last) {
        const double* o = map + 7 * m.obj;
        pr.cyl.insert(pr.cyl.end(), {o[0], o[1], o[2], o[3], o[4], o[5], o[6], 0.0});
        last = m.obj;
      }
      push(pr, m.p, (int32_t)(pr.cyl.size() / 8 - 1));
    }
  };
  auto ground_problem = [&](Problem& pr) {
    int last = -1;
    for (const Match& m : gmatch) {
      if (m.obj != last) {
        const PlaneM& o = s->prev[m.obj];
        pr.plane.insert(pr.plane.end(), {o.n[0], o.n[1], o.n[2], o.n[3]});
        last = m.obj;
      }
      push(pr, m.p, -(int32_t)(pr.plane.size() / 4));
    }
  };
  Pose cur = est, delta = identity;
  bool success = true;
  if (prm.twoStep) {
    double p0[6] = {est.t[0], est.t[1], est.t[2], 0, 0, 0};
    quat_to_aa(est.q, p0 + 3);
    double treeOut[3] = {p0[0], p0[1], p0[5]}, groundOut[3] = {p0[2], p0[3], p0[4]};
    bool conv = true;
    if (treeCheck && !tmatch.empty()) {
      Problem pr;
      tree_problem(pr);
      double po[7], pa[6];
      const bool ok = solve(pr, est, 1, (1 << 0) | (1 << 1) | (1 << 5), po, pa, &info[10]);
      conv = conv && ok;
      if (ok) {
        treeOut[0] = pa[0];
        treeOut[1] = pa[1];
        treeOut[2] = pa[5];
      }
    }
    if (groundCheck && !gmatch.empty()) {
      Problem pr;
      ground_problem(pr);
      double po[7], pa[6];
      const bool ok = solve(pr, est, 1, (1 << 2) | (1 << 3) | (1 << 4), po, pa, &info[11]);
      conv = conv && ok;
      if (ok) {
        groundOut[0] = pa[2];
        groundOut[1] = pa[3];
        groundOut[2] = pa[4];
      }
    }
    const double aa[3] = {groundOut[1], groundOut[2], treeOut[2]};
    aa_to_quat(aa, cur.q);
    cur.t[0] = treeOut[0];
    cur.t[1] = treeOut[1];
    cur.t[2] = groundOut[0];
    info[9] = conv;
    success = true;  // TwoStepOptimizePose always returns true
  } else if (treeCheck && groundCheck) {
    Problem pr;
    tree_problem(pr);
    ground_problem(pr);
    double po[7], pa[6];
    success = solve(pr, est, 0, 0x3f, po, pa, &info[10]);
    info[9] = success;
    if (success) {
      std::memcpy(delta.t, po, 3 * sizeof(double));
      std::memcpy(delta.q, po + 3, 4 * sizeof(double));
      cur = delta;
    }
  }
  for (auto& p : planes) project_plane(cur, p);
  for (auto& c : trees) project_cyl_inplace(cur, c);
  // --- matchModels (threshold AddNewTreeThreshDist)
  for (size_t k = 0; k < trees.size(); ++k) {
    double best = prm.treeMatchThresh + 100;
    int bi = 0;
    for (int m = 0; m < n_map; ++m) {
      CylM mm;
      std::memcpy(mm.root, map + 7 * m, 3 * sizeof(double));
      std::memcpy(mm.ray, map + 7 * m + 3, 3 * sizeof(double));
      const double d = cyl_model_distance(mm, trees[k].root, trees[k].ray);
      if (d < best) {
        best = d;
        bi = m;
      }
    }
    if (best < prm.addNewTreeThreshDist) match_idx[k] = bi;
  }
  info[0] = success;
  write_out(cur, delta);
  s->prev = planes;
  return 0;
}

// Euler angles (0, 1, 2) of the smallest rotation taking a onto b, as the
// plane acceptance reads them (exposed for its own tests).
int oracle_from_two_vectors_euler(const double a[3], const double b[3], double R_out[9],
                                  double e[3]) {
  double R[3][3];
  rotation_between(a, b, R);
  euler_xyz(R, e);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R_out[3 * i + j] = R[i][j];
  return 0;
}

}  // extern "C"
