// Semantic map of tree landmarks (sloam/src/core/mapManager.cpp).
#include "mapManager.h"

#include <algorithm>
#include <numeric>
#include <utility>

MapManager::MapManager(const float searchRadius) {
  sqSearchRadius = searchRadius * searchRadius;
  landmarks_.reset(new CloudT);
}

void MapManager::updateMap(std::vector<Cylinder>& obs_tms, const std::vector<int>& matches) {
  size_t i = 0;
  for (const auto& tree : obs_tms) {
    PointT pt;
    pt.x = (float)tree.model.root[0];
    pt.y = (float)tree.model.root[1];
    pt.z = (float)tree.model.root[2];
    if (matches[i] == -1) {
      landmarks_->points.push_back(pt);
      treeModels_.push_back(tree);
      treeHits_.push_back(1);
    } else {
      // observation -> map index, recorded by the last getSubmap
      const int matchIdx = matchesMap[matches[i]];
      landmarks_->points[matchIdx] = pt;
      treeModels_[matchIdx] = tree;
      treeHits_[matchIdx] += 1;
    }
    i++;
  }
  landmarks_->width = (uint32_t)landmarks_->points.size();
  landmarks_->height = 1;
  matchesMap.clear();
}

std::vector<Cylinder> MapManager::getMap() {
  std::vector<Cylinder> map;
  for (size_t i = 0; i < treeModels_.size(); ++i)
    if (treeHits_[i] > 2) map.push_back(treeModels_[i]);
  return map;
}

void MapManager::getSubmap(const SE3& pose, std::vector<Cylinder>& submap) {
  const auto& pts = landmarks_->points;
  if (pts.empty()) return;
  // squared distances in float, as PCL's search over PointT
  const float sx = (float)pose.translation()[0], sy = (float)pose.translation()[1], sz = 1.0f;
  std::vector<std::pair<float, int>> d(pts.size());
  for (size_t i = 0; i < pts.size(); ++i) {
    const float dx = pts[i].x - sx, dy = pts[i].y - sy, dz = pts[i].z - sz;
    d[i] = {dx * dx + dy * dy + dz * dz, (int)i};
  }
  const size_t k = std::min<size_t>(100, d.size());
  std::partial_sort(d.begin(), d.begin() + (long)k, d.end());
  const size_t map_size = treeModels_.size();
  int idx_count = 0;
  for (size_t j = 0; j < k; ++j) {
    const int map_idx = d[j].second;
    if (map_size - (size_t)map_idx < 200) {
      matchesMap.insert(std::pair<int, int>(idx_count, map_idx));
      submap.push_back(treeModels_[map_idx]);
      idx_count++;
    }
  }
}
