// TEST INFRASTRUCTURE ONLY -- FP64 CPU oracle of SLOAM's pose estimation.
//
// Restates sloam::sloam::OptimizePose (sloam/src/core/sloam.cpp:174-255):
// parameter blocks t[3] and a unit quaternion q (sloam.cpp:187-193), one
// residual per matched feature,
//   cylinder  r = ||pp - lp|| - radius            (sloam/include/objects/cylinder.h:90-132)
//   plane     r = |pi . [lp; 1]| / ||pi_xyz||     (sloam/include/objects/plane.h:76-116)
// with lp = q * p + t, HuberLoss(0.1) (sloam.cpp:69) and Levenberg-Marquardt
// (sloam.cpp:221-227, max 50 iterations, success <=> CONVERGENCE).
//
// The solver is Ceres' documented LM trust-region loop, restated (Ceres is not
// available here; its version is pinned in docker/SLOAMDockerfile:30-32):
//   * Huber corrector: rho'' <= 0, so the corrected system is IRLS with weight
//     w = rho'(r^2) = min(1, delta / |r|); cost = 1/2 sum rho(r^2);
//   * step: (J'WJ + D / radius) dx = -J'Wr, D = diag(J'WJ) clamped to
//     [1e-6, 1e32]; initial radius 1e4;
//   * quality q = (F - F_new) / (model decrease), accepted if q > 1e-3; then
//     radius /= max(1/3, 1 - (2q - 1)^3), decrease factor = 2; otherwise
//     radius /= factor, factor *= 2;
//   * convergence: |dF| / F <= 1e-6 after an accepted step, ||g||_inf <= 1e-10,
//     or ||dx|| <= 1e-8 (||x|| + 1e-8); NO_CONVERGENCE at max iterations.
// The rotation update is a left perturbation exp(dtheta) * R (Ceres'
// EigenQuaternionParameterization, whose delta is half the rotation vector;
// the LM step is invariant to that diagonal scaling because D = diag(J'WJ)).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle_common.hpp"

namespace oracle {
namespace {

struct Pose {
  double t[3];
  double q[4];  // x, y, z, w
};

void quat_to_R(const double q[4], double R[3][3]) {
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

// q' = exp(dtheta) (x) q, normalised
void quat_left_update(const double q[4], const double dth[3], double out[4]) {
  const double ang = std::sqrt(dth[0] * dth[0] + dth[1] * dth[1] + dth[2] * dth[2]);
  double d[4];
  if (ang > 1e-300) {
    const double s = std::sin(0.5 * ang) / ang;
    d[0] = dth[0] * s;
    d[1] = dth[1] * s;
    d[2] = dth[2] * s;
    d[3] = std::cos(0.5 * ang);
  } else {
    d[0] = 0.5 * dth[0];
    d[1] = 0.5 * dth[1];
    d[2] = 0.5 * dth[2];
    d[3] = 1.0;
  }
  // Hamilton product d * q
  const double x = d[3] * q[0] + d[0] * q[3] + d[1] * q[2] - d[2] * q[1];
  const double y = d[3] * q[1] - d[0] * q[2] + d[1] * q[3] + d[2] * q[0];
  const double z = d[3] * q[2] + d[0] * q[1] - d[1] * q[0] + d[2] * q[3];
  const double w = d[3] * q[3] - d[0] * q[0] - d[1] * q[1] - d[2] * q[2];
  const double n = std::sqrt(x * x + y * y + z * z + w * w);
  out[0] = x / n;
  out[1] = y / n;
  out[2] = z / n;
  out[3] = w / n;
}

// residual + 1x6 Jacobian [d/dt, d/dtheta] of one feature at pose (R, t)
//   cyl = root[3], ray[3], radius; plane = n[3], d
double residual(const double R[3][3], const double t[3], const float* p, int32_t id,
                const double* cyl, const double* pl, double J[6]) {
  const double P[3] = {p[0], p[1], p[2]};
  double rp[3];
  for (int i = 0; i < 3; ++i) rp[i] = R[i][0] * P[0] + R[i][1] * P[1] + R[i][2] * P[2];
  const double lp[3] = {rp[0] + t[0], rp[1] + t[1], rp[2] + t[2]};
  double dir[3];  // d r / d lp
  double r;
  if (id >= 0) {
    const double* c = cyl + 8 * (size_t)id;
    const double rr = std::sqrt(c[3] * c[3] + c[4] * c[4] + c[5] * c[5]);
    const double a[3] = {c[3] / rr, c[4] / rr, c[5] / rr};
    const double v[3] = {lp[0] - c[0], lp[1] - c[1], lp[2] - c[2]};
    const double h = v[0] * a[0] + v[1] * a[1] + v[2] * a[2];
    const double w[3] = {v[0] - h * a[0], v[1] - h * a[1], v[2] - h * a[2]};
    const double rho = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    r = rho - c[6];
    for (int i = 0; i < 3; ++i) dir[i] = rho > 0 ? w[i] / rho : 0.0;
  } else {
    const double* q = pl + 4 * (size_t)(-id - 1);
    const double nn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    const double s = (q[0] * lp[0] + q[1] * lp[1] + q[2] * lp[2] + q[3]) / nn;
    r = std::fabs(s);
    const double sg = s < 0 ? -1.0 : 1.0;
    for (int i = 0; i < 3; ++i) dir[i] = sg * q[i] / nn;
  }
  J[0] = dir[0];
  J[1] = dir[1];
  J[2] = dir[2];
  // d lp / d dtheta = -[R p]x  ->  d r / d dtheta = (R p) x dir
  J[3] = rp[1] * dir[2] - rp[2] * dir[1];
  J[4] = rp[2] * dir[0] - rp[0] * dir[2];
  J[5] = rp[0] * dir[1] - rp[1] * dir[0];
  return r;
}

// ceres::QuaternionToAngleAxis / AngleAxisToQuaternion (q stored x, y, z, w)
void quat_to_aa(const double q[4], double aa[3]) {
  const double s2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  if (s2 > 0.0) {
    const double st = std::sqrt(s2);
    const double c = q[3];
    const double two_theta = 2.0 * ((c < 0.0) ? std::atan2(-st, -c) : std::atan2(st, c));
    const double k = two_theta / st;
    for (int i = 0; i < 3; ++i) aa[i] = q[i] * k;
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

// left Jacobian of SO(3): Jl(w) = I + (1 - cos t)/t^2 [w]x + (t - sin t)/t^3 [w]x^2
void left_jacobian(const double w[3], double J[3][3]) {
  const double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double a, b;
  if (t2 < 1e-12) {
    a = 0.5 - t2 / 24.0;
    b = 1.0 / 6.0 - t2 / 120.0;
  } else {
    const double t = std::sqrt(t2);
    a = (1.0 - std::cos(t)) / t2;
    b = (t - std::sin(t)) / (t2 * t);
  }
  const double K[3][3] = {{0.0, -w[2], w[1]}, {w[2], 0.0, -w[0]}, {-w[1], w[0], 0.0}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double k2 = 0.0;
      for (int m = 0; m < 3; ++m) k2 += K[i][m] * K[m][j];
      J[i][j] = (i == j ? 1.0 : 0.0) + a * K[i][j] + b * k2;
    }
}

struct Eval {
  double F;
  double A[6][6];
  double g[6];
};

// param 0: Jacobian w.r.t. [dt, left rotation]; param 1: w.r.t. [t, w]
// additive (the rotation columns mapped by Jl(w)), masked by free_mask
void evaluate(const Pose& x, const double* aa, int param, int free_mask, const float* pts, int n,
              const double* cyl, const double* pl, double delta, Eval* e) {
  double R[3][3];
  quat_to_R(x.q, R);
  double Jl[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (param == 1) left_jacobian(aa + 3, Jl);
  std::memset(e, 0, sizeof(Eval));
  const double d2 = delta * delta;
  for (int i = 0; i < n; ++i) {
    const float* p = pts + 4 * (size_t)i;
    int32_t id;
    std::memcpy(&id, p + 3, 4);
    double J[6];
    const double r = residual(R, x.t, p, id, cyl, pl, J);
    if (param == 1) {
      const double jt[3] = {J[3], J[4], J[5]};
      for (int c = 0; c < 3; ++c) J[3 + c] = jt[0] * Jl[0][c] + jt[1] * Jl[1][c] + jt[2] * Jl[2][c];
    }
    for (int a = 0; a < 6; ++a)
      if (!((free_mask >> a) & 1)) J[a] = 0.0;
    const double s = r * r;
    double w, rho;
    if (s <= d2) {
      w = 1.0;
      rho = s;
    } else {
      const double ar = std::sqrt(s);
      w = delta / ar;
      rho = 2.0 * delta * ar - d2;
    }
    e->F += 0.5 * rho;
    for (int a = 0; a < 6; ++a) {
      for (int b = a; b < 6; ++b) e->A[a][b] += w * J[a] * J[b];
      e->g[a] += w * J[a] * r;
    }
  }
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < a; ++b) e->A[a][b] = e->A[b][a];
}

Pose apply_step(const Pose& x, const double dx[6]) {
  Pose y;
  for (int i = 0; i < 3; ++i) y.t[i] = x.t[i] + dx[i];
  quat_left_update(x.q, dx + 3, y.q);
  return y;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

// Residuals and 1x6 Jacobians of every feature at pose (t, q = x,y,z,w).
// pts: float4 per feature (x, y, z, landmark id bits: >= 0 cylinder index,
// < 0 plane index -(k+1)); cyl: 8 doubles per cylinder (root, ray, radius,
// unused); pl: 4 doubles per plane.
int oracle_pose_residuals(const float* pts, int n, const double* cyl, const double* pl,
                          const double pose[7], double* r, double* J) {
  double R[3][3];
  quat_to_R(pose + 3, R);
  for (int i = 0; i < n; ++i) {
    int32_t id;
    std::memcpy(&id, pts + 4 * (size_t)i + 3, 4);
    r[i] = residual(R, pose, pts + 4 * (size_t)i, id, cyl, pl, J + 6 * (size_t)i);
  }
  return 0;
}

// Batched LM, one problem per residual range [off[k], off[k+1]).
// cfg_i = {max_iters, stop_on_convergence}
// cfg_d = {huber_delta, function_tol, gradient_tol, parameter_tol,
//          initial_radius, min_relative_decrease}
// pose_in/out: 7 doubles per problem (t, q xyzw); status: 2 ints per problem
// (termination: 0 CONVERGENCE, 1 NO_CONVERGENCE; iterations); cost: 2 doubles
// (initial, final).
int oracle_pose_lm_batch(const float* pts, const int32_t* off, int n_problems, const double* cyl,
                         const double* pl, const double* pose_in, const int32_t* free_masks,
                         const int32_t cfg_i[3], const double cfg_d[6], double* pose_out,
                         double* params_out, int32_t* status, double* cost, int n_threads) {
  const int max_iters = cfg_i[0];
  const bool stop_conv = cfg_i[1] != 0;
  const int param = cfg_i[2];
  const double delta = cfg_d[0], ftol = cfg_d[1], gtol = cfg_d[2], ptol = cfg_d[3];
  const double radius0 = cfg_d[4], min_rel = cfg_d[5];
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads > 0 ? n_threads : 1)
  for (int k = 0; k < n_problems; ++k) {
    const float* P = pts + 4 * (size_t)off[k];
    const int n = off[k + 1] - off[k];
    Pose x;
    std::memcpy(x.t, pose_in + 7 * k, 3 * sizeof(double));
    std::memcpy(x.q, pose_in + 7 * k + 3, 4 * sizeof(double));
    const int fm = free_masks ? free_masks[k] : 0x3f;
    double aa[6] = {x.t[0], x.t[1], x.t[2], 0, 0, 0};
    quat_to_aa(x.q, aa + 3);
    Eval cur;
    evaluate(x, aa, param, fm, P, n, cyl, pl, delta, &cur);
    const double F0 = cur.F;
    double radius = radius0, factor = 2.0;
    int term = 1, it = 0;
    bool done = false;
    {
      double gmax = 0;
      for (int a = 0; a < 6; ++a) gmax = std::fmax(gmax, std::fabs(cur.g[a]));
      if (gmax <= gtol) {
        term = 0;
        done = stop_conv;
      }
    }
    while (!done && it < max_iters) {
      ++it;
      double Aug[6][6];
      double rhs[6];
      for (int a = 0; a < 6; ++a) {
        for (int b = 0; b < 6; ++b) Aug[a][b] = cur.A[a][b];
        double dg = cur.A[a][a];
        dg = dg < 1e-6 ? 1e-6 : (dg > 1e32 ? 1e32 : dg);
        Aug[a][a] += dg / radius;
        rhs[a] = -cur.g[a];
      }
      double dx[6];
      if (!cholesky_solve6(Aug, rhs, dx)) {
        radius /= factor;
        factor *= 2.0;
        continue;
      }
      // model decrease: -(g.dx + 1/2 dx' A dx)
      double gdx = 0, dAd = 0;
      for (int a = 0; a < 6; ++a) {
        gdx += cur.g[a] * dx[a];
        double t = 0;
        for (int b = 0; b < 6; ++b) t += cur.A[a][b] * dx[b];
        dAd += dx[a] * t;
      }
      const double model = -(gdx + 0.5 * dAd);
      double xn = 0, dn = 0;
      if (param == 1) {
        for (int a = 0; a < 6; ++a) xn += aa[a] * aa[a];
      } else {
        for (int a = 0; a < 3; ++a) xn += x.t[a] * x.t[a];
        for (int a = 0; a < 4; ++a) xn += x.q[a] * x.q[a];
      }
      for (int a = 0; a < 6; ++a) dn += dx[a] * dx[a];
      xn = std::sqrt(xn);
      dn = std::sqrt(dn);
      Pose xc;
      double aac[6];
      if (param == 1) {
        for (int a = 0; a < 6; ++a) aac[a] = aa[a] + dx[a];
        for (int a = 0; a < 3; ++a) xc.t[a] = aac[a];
        aa_to_quat(aac + 3, xc.q);
      } else {
        xc = apply_step(x, dx);
        for (int a = 0; a < 6; ++a) aac[a] = aa[a];
      }
      Eval cand;
      evaluate(xc, aac, param, fm, P, n, cyl, pl, delta, &cand);
      const double change = cur.F - cand.F;
      const double q = model > 0 ? change / model : -1.0;
      if (std::isfinite(cand.F) && q > min_rel) {
        x = xc;
        if (param == 1)
          for (int a = 0; a < 6; ++a) aa[a] = aac[a];
        const double Fprev = cur.F;
        cur = cand;
        const double c3 = 2.0 * q - 1.0;
        radius = radius / std::fmax(1.0 / 3.0, 1.0 - c3 * c3 * c3);
        radius = std::fmin(radius, 1e16);
        factor = 2.0;
        double gmax = 0;
        for (int a = 0; a < 6; ++a) gmax = std::fmax(gmax, std::fabs(cur.g[a]));
        if (std::fabs(change) <= ftol * Fprev || gmax <= gtol || dn <= ptol * (xn + ptol)) {
          term = 0;
          if (stop_conv) done = true;
        }
      } else {
        radius /= factor;
        factor *= 2.0;
        if (dn <= ptol * (xn + ptol)) {
          term = 0;
          if (stop_conv) done = true;
        }
      }
    }
    std::memcpy(pose_out + 7 * k, x.t, 3 * sizeof(double));
    std::memcpy(pose_out + 7 * k + 3, x.q, 4 * sizeof(double));
    if (params_out) {
      if (param == 1) {
        std::memcpy(params_out + 6 * k, aa, 6 * sizeof(double));
      } else {
        std::memcpy(params_out + 6 * k, x.t, 3 * sizeof(double));
        quat_to_aa(x.q, params_out + 6 * k + 3);
      }
    }
    status[2 * k] = term;
    status[2 * k + 1] = it;
    cost[2 * k] = F0;
    cost[2 * k + 1] = cur.F;
  }
  return 0;
}

int oracle_left_jacobian(const double w[3], double J[9]) {
  double JJ[3][3];
  left_jacobian(w, JJ);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[3 * i + j] = JJ[i][j];
  return 0;
}

int oracle_aa_quat(const double aa[3], double q[4], double back[3]) {
  aa_to_quat(aa, q);
  quat_to_aa(q, back);
  return 0;
}

}  // extern "C"
