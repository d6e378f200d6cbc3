// Semantic map of tree landmarks: the interface of
// sloam/include/core/mapManager.h:5-16 (MapManager).  The submap query is an
// exact k-nearest search over the landmark roots on the host (the reference
// builds a KdTreeFLANN per query), ordered by squared distance and then by
// map index.
#pragma once

#include <cstddef>
#include <map>
#include <vector>

#include "cylinder.h"
#include "definitions.h"

class MapManager {
 public:
  explicit MapManager(const float searchRadius = 50);
  // every landmark observed more than twice (mapManager.cpp:31-40)
  std::vector<Cylinder> getMap();
  // appends to submap the (at most 100) landmarks nearest to (x, y, 1) of
  // the pose that are among the 200 most recently added, nearest first, and
  // remembers which map entry each submap entry is (mapManager.cpp:42-70)
  void getSubmap(const SE3& pose, std::vector<Cylinder>& submap);
  // observations with match -1 become new landmarks; a matched one replaces
  // its map entry (matches index the last submap) and adds a hit
  // (mapManager.cpp:8-29)
  void updateMap(std::vector<Cylinder>& obs_tms, const std::vector<int>& matches);

  size_t size() const { return treeModels_.size(); }

 private:
  float sqSearchRadius;  // kept as in the reference (unused by its queries)
  CloudT::Ptr landmarks_;
  std::vector<Cylinder> treeModels_;
  std::map<int, int> matchesMap;
  std::vector<size_t> treeHits_;
};
