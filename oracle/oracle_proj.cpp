// TEST INFRASTRUCTURE ONLY -- CPU oracle of the range projection and the
// ground / tree partition.
//
// Restates seg::Segmentation (sloam/src/segmentation/inference.cpp):
//   constructor         :7-19    fov_up/down in radians stored as float, fov = |down| + |up|
//   _doProjection       :80-165  per point float range, yaw = -atan2f(y, x),
//                                pitch = asinf(z / range); proj_x = 0.5 (yaw/pi + 1)
//                                and proj_y = 1 - (pitch + |fov_down|) / fov computed
//                                with double literals and stored as float, scaled by
//                                W / H in float, floored and clamped; points written
//                                in decreasing range so the nearest point wins
//   _makeTensor         :167-198 pixel invalid <=> (int)range == 0
//   _mask               :275-300 class 2 -> 255, 1 -> 1, 0 -> 0, invalid -> 0
//   maskCloud           :230-273 ground: compacted input-order list of points whose
//                                pixel mask == 1; tree: dense input-order cloud,
//                                NaN where the pixel mask != 255
//   _destaggerCloud     :200-228 even rows shifted by +32 columns
// With labels given, the class of a pixel is the label of its nearest point.
// Deliberate definitions where the reference is unspecified: points with a
// non-finite coordinate take no part (the reference's sort on NaN ranges is
// undefined behaviour); among equal ranges in one pixel the lowest input index
// wins (the reference's std::sort is unstable).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

extern "C" {

// xyzi: n x 4 floats; label: 0 other, 1 ground, 2 tree.
// Outputs: proj (n x 2 int32: col, row; -1 for skipped points), range image
// (H*W float, 0 = empty), winner (H*W int32, -1 = empty), mask (H*W uint8),
// ground list (xyzi, count returned in *n_ground), tree cloud (n x 4, NaN fill).
int oracle_project_partition(const float* xyzi, const uint8_t* label, int n, int H, int W,
                             float fov_up_deg, float fov_down_deg, int destagger, int32_t* proj,
                             float* range_img, int32_t* winner, uint8_t* mask, float* ground,
                             int32_t* n_ground, float* tree) {
  const float fov_up = fov_up_deg / 180.0 * M_PI;
  const float fov_down = fov_down_deg / 180.0 * M_PI;
  const float fov = std::abs(fov_down) + std::abs(fov_up);
  std::vector<float> ranges(n);
  std::vector<char> valid(n);
  for (int i = 0; i < n; ++i) {
    const float x = xyzi[4 * i], y = xyzi[4 * i + 1], z = xyzi[4 * i + 2];
    valid[i] = std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
    if (!valid[i]) {
      proj[2 * i] = proj[2 * i + 1] = -1;
      ranges[i] = 0.f;
      continue;
    }
    const float range = std::sqrt(x * x + y * y + z * z);
    ranges[i] = range;
    const float yaw = -std::atan2(y, x);
    const float pitch = std::asin(z / range);
    float proj_x = 0.5 * (yaw / M_PI + 1.0);
    float proj_y = 1.0 - (pitch + std::abs(fov_down)) / fov;
    proj_x *= W;
    proj_y *= H;
    proj_x = std::floor(proj_x);
    proj_x = std::min(W - 1.0f, proj_x);
    proj_x = std::max(0.0f, proj_x);
    proj_y = std::floor(proj_y);
    proj_y = std::min(H - 1.0f, proj_y);
    proj_y = std::max(0.0f, proj_y);
    proj[2 * i] = (int32_t)proj_x;
    proj[2 * i + 1] = (int32_t)proj_y;
  }
  // decreasing range, equal ranges by decreasing index: the last write into a
  // pixel is its nearest point with the lowest index
  std::vector<int> order;
  order.reserve(n);
  for (int i = 0; i < n; ++i)
    if (valid[i]) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (ranges[a] != ranges[b]) return ranges[a] > ranges[b];
    return a > b;
  });
  for (int p = 0; p < H * W; ++p) {
    range_img[p] = 0.f;
    winner[p] = -1;
  }
  for (int i : order) {
    const int pix = proj[2 * i + 1] * W + proj[2 * i];
    range_img[pix] = ranges[i];
    winner[pix] = i;
  }
  for (int p = 0; p < H * W; ++p) {
    const bool invalid = (int)range_img[p] == 0;
    uint8_t m = 0;
    if (!invalid) {
      const uint8_t c = label[winner[p]];
      m = c == 2 ? 255 : (c == 1 ? 1 : 0);
    }
    mask[p] = m;
  }
  int ng = 0;
  const float qnan = std::nanf("");
  for (int i = 0; i < n; ++i) {
    uint8_t m = 0;
    if (valid[i]) m = mask[proj[2 * i + 1] * W + proj[2 * i]];
    if (m == 1) {
      std::memcpy(ground + 4 * (size_t)ng, xyzi + 4 * (size_t)i, 16);
      ++ng;
    }
    // dense tree cloud in input order
    float* o = tree + 4 * (size_t)i;
    if (m == 255) {
      std::memcpy(o, xyzi + 4 * (size_t)i, 16);
    } else {
      o[0] = o[1] = o[2] = qnan;
      o[3] = 0.f;
    }
  }
  *n_ground = ng;
  if (destagger && n == H * W) {
    // even rows: out[r][c] = in[r][c - 32] (x, y, z) for c >= 32
    std::vector<float> src(tree, tree + 4 * (size_t)n);
    for (int r = 0; r < H; r += 2)
      for (int c = 32; c < W; ++c)
        for (int k = 0; k < 3; ++k) tree[4 * ((size_t)r * W + c) + k] = src[4 * ((size_t)r * W + c - 32) + k];
  }
  return 0;
}

}  // extern "C"
