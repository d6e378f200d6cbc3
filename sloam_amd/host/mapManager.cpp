// Semantic map of tree landmarks (the semantics of
// sloam/src/core/mapManager.cpp:8-71, layout and query of this file's own).
#include "mapManager.h"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <utility>

MapManager::MapManager(const float searchRadius) : sqSearchRadius_(searchRadius * searchRadius) {}

void MapManager::query(const SE3& pose) {
  last_query_.clear();
  const size_t n = records_.size();
  if (n == 0) return;
  // the query point: the pose's x, y at height 1 (mapManager.cpp:48-50)
  const float q[3] = {(float)pose.translation()[0], (float)pose.translation()[1], 1.0f};
  auto dist2 = [&](const Record& r) {
    const float a = r.key[0] - q[0], b = r.key[1] - q[1], c = r.key[2] - q[2];
    return a * a + b * b + c * c;
  };
  rank_.resize(n);
  for (size_t i = 0; i < n; ++i) rank_[i] = {dist2(records_[i]), (uint32_t)i};
  // rank cut: the kNeighbours-th smallest (distance, index) over the whole map
  const size_t k = std::min(kNeighbours, n);
  std::nth_element(rank_.begin(), rank_.begin() + (long)(k - 1), rank_.end());
  const std::pair<float, uint32_t> cut = rank_[k - 1];
  // members: the recent window's records at or below the cut (the window is
  // the records with size - index < kRecent)
  const size_t first = n >= kRecent ? n - kRecent + 1 : 0;
  std::vector<std::pair<float, uint32_t>> members;
  members.reserve(std::min(k, n - first));
  for (size_t i = first; i < n; ++i) {
    const std::pair<float, uint32_t> key{dist2(records_[i]), (uint32_t)i};
    if (!(cut < key)) members.push_back(key);
  }
  std::sort(members.begin(), members.end());
  last_query_.reserve(members.size());
  for (const auto& m : members) last_query_.push_back(m.second);
}

void MapManager::getSubmap(const SE3& pose, std::vector<Cylinder>& submap) {
  query(pose);
  submap.reserve(submap.size() + last_query_.size());
  for (const uint32_t i : last_query_) submap.push_back(records_[i].tree);
}

void MapManager::getSubmap(const SE3& pose, MapView& submap) {
  query(pose);
  submap.reserve(submap.size() + last_query_.size());
  for (const uint32_t i : last_query_) submap.push_back(&records_[i].tree);
}

void MapManager::updateMap(std::vector<Cylinder>& obs_tms, const std::vector<int>& matches) {
  if (matches.size() < obs_tms.size())
    throw std::runtime_error("MapManager::updateMap: " + std::to_string(matches.size()) +
                             " matches for " + std::to_string(obs_tms.size()) + " trees");
  // matched records first: a new record may move the vector (and the trees
  // the last submap view points at)
  for (size_t k = 0; k < obs_tms.size(); ++k) {
    const int j = matches[k];
    if (j == -1) continue;
    if (j < 0 || (size_t)j >= last_query_.size())
      throw std::runtime_error("MapManager::updateMap: match " + std::to_string(j) +
                               " is not an entry of the last submap");
    Record& r = records_[last_query_[j]];
    r.tree = obs_tms[k];
    for (int c = 0; c < 3; ++c) r.key[c] = (float)r.tree.model.root[c];
    ++r.hits;
  }
  for (size_t k = 0; k < obs_tms.size(); ++k) {
    if (matches[k] != -1) continue;
    Record r{obs_tms[k], {0.f, 0.f, 0.f}, 1u};
    for (int c = 0; c < 3; ++c) r.key[c] = (float)r.tree.model.root[c];
    records_.push_back(std::move(r));
  }
  last_query_.clear();
}

std::vector<Cylinder> MapManager::getMap() {
  std::vector<Cylinder> out;
  for (const Record& r : records_)
    if (r.hits > 2) out.push_back(r.tree);
  return out;
}
