// sloam_dm's sine, cosine, arc tangent and arc cosine evaluated on the GPU and
// on the host for the same arguments: every result must carry the same bits.
// Also counts, for information, how often the GPU's libm and the host's libm
// differ on the same arguments (the reason the pose composition and the
// model gates use sloam_dm).  Built with the library's device flags.
//
//   detmath_gpu [n]     exit 0 when all n x 4 results agree bit for bit
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "detmath.h"

namespace {

constexpr int kOut = 8;  // per argument: 4 sloam_dm results, 4 libm results

__global__ void eval_kernel(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ u, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  sloam_dm::sincos(x[i], &s, &c);
  double* o = out + (size_t)kOut * i;
  o[0] = s;
  o[1] = c;
  o[2] = sloam_dm::atan2(y[i], x[i]);
  o[3] = sloam_dm::acos(u[i]);
  o[4] = ::sin(x[i]);
  o[5] = ::cos(x[i]);
  o[6] = ::atan2(y[i], x[i]);
  o[7] = ::acos(u[i]);
}

bool ok(hipError_t e, const char* what) {
  if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

}  // namespace

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : (1 << 20);
  if (n < 16) return 2;
  std::vector<double> x(n), y(n), u(n);
  std::mt19937_64 g(2024);
  std::uniform_real_distribution<double> small(-4.0, 4.0), wide(-1000.0, 1000.0),
      unit(-1.0, 1.0), expo(-8.0, 8.0);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < n; ++i) {
    x[i] = (i & 1) ? small(g) : (i & 2) ? wide(g) : nd(g) * std::pow(10.0, expo(g));
    y[i] = (i & 4) ? nd(g) : nd(g) * std::pow(10.0, expo(g));
    u[i] = unit(g);
  }
  // exact points: zeros of both signs, axes, +-1 and the ends of acos' range
  const double special[][3] = {{0.0, 0.0, 1.0},   {-0.0, 0.0, -1.0}, {1.0, 0.0, 0.0},
                               {-1.0, 0.0, -0.0}, {0.0, 1.0, 0.5},   {0.0, -1.0, -0.5},
                               {1.0, 1.0, 1.0 - 1e-16}, {-1.0, -1.0, -1.0 + 1e-16},
                               {M_PI, -0.0, 0.9999999}, {M_PI / 2, 1e-300, 1e-300}};
  for (int k = 0; k < 10; ++k) {
    x[k] = special[k][0];
    y[k] = special[k][1];
    u[k] = special[k][2];
  }
  double *dx = nullptr, *dy = nullptr, *du = nullptr, *dout = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n;
  if (!ok(hipMalloc(&dx, bytes), "hipMalloc") || !ok(hipMalloc(&dy, bytes), "hipMalloc") ||
      !ok(hipMalloc(&du, bytes), "hipMalloc") || !ok(hipMalloc(&dout, kOut * bytes), "hipMalloc"))
    return 1;
  if (!ok(hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice), "copy") ||
      !ok(hipMemcpy(dy, y.data(), bytes, hipMemcpyHostToDevice), "copy") ||
      !ok(hipMemcpy(du, u.data(), bytes, hipMemcpyHostToDevice), "copy"))
    return 1;
  const int block = 256, grid = (n + block - 1) / block;
  hipLaunchKernelGGL(eval_kernel, dim3(grid), dim3(block), 0, 0, dx, dy, du, n, dout);
  if (!ok(hipGetLastError(), "launch") || !ok(hipDeviceSynchronize(), "kernel")) return 1;
  std::vector<double> out((size_t)kOut * n);
  if (!ok(hipMemcpy(out.data(), dout, kOut * bytes, hipMemcpyDeviceToHost), "copy back")) return 1;
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipFree(du);
  (void)hipFree(dout);

  const char* names[4] = {"sin", "cos", "atan2", "acos"};
  long long differ[4] = {0, 0, 0, 0}, libm_differ[4] = {0, 0, 0, 0};
  int printed = 0;
  for (int i = 0; i < n; ++i) {
    double s, c;
    sloam_dm::sincos(x[i], &s, &c);
    const double host[4] = {s, c, sloam_dm::atan2(y[i], x[i]), sloam_dm::acos(u[i])};
    const double libm[4] = {std::sin(x[i]), std::cos(x[i]), std::atan2(y[i], x[i]),
                            std::acos(u[i])};
    const double* o = out.data() + (size_t)kOut * i;
    for (int f = 0; f < 4; ++f) {
      if (std::memcmp(&host[f], &o[f], sizeof(double)) != 0) {
        ++differ[f];
        if (printed++ < 10)
          std::fprintf(stderr, "%s differs at x %.17g y %.17g u %.17g: host %.17g device %.17g\n",
                       names[f], x[i], y[i], u[i], host[f], o[f]);
      }
      if (std::memcmp(&libm[f], &o[4 + f], sizeof(double)) != 0) ++libm_differ[f];
    }
  }
  long long total = 0;
  for (int f = 0; f < 4; ++f) {
    std::printf("%s: sloam_dm host/device differ %lld of %d; libm host/device differ %lld\n",
                names[f], differ[f], n, libm_differ[f]);
    total += differ[f];
  }
  return total == 0 ? 0 : 1;
}
