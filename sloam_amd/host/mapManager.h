// Semantic map of tree landmarks with the interface of
// sloam/include/core/mapManager.h:5-16 (MapManager).
//
// Semantics kept from sloam/src/core/mapManager.cpp:8-71: a landmark enters
// getMap() after more than two observations; the submap around a pose holds
// the map trees ranked among the 100 nearest to (x, y, 1) of the pose (by
// their roots, squared distances in float) that are also among the 200 most
// recently added, nearest first; an observation matched to submap entry j
// replaces the tree that entry came from and adds a hit, an unmatched one is
// a new landmark.  Equal distances rank by map index (the reference leaves
// the kd-tree's tie order open).
//
// Layout: one record per landmark (the tree, its root as the float key of
// the queries, its hit count).  A query ranks every landmark by (distance,
// index), cuts at the 100th, and walks only the 200-record recent window
// for the members; the window's records are the only ones an update can
// touch.  getSubmap(pose, MapView&) hands out the trees by reference, so a
// keyframe does not copy each tree's vertices and points.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "cylinder.h"
#include "definitions.h"
#include "sloam.h"

class MapManager {
 public:
  explicit MapManager(const float searchRadius = 50);
  // every landmark observed more than twice, in map order
  std::vector<Cylinder> getMap();
  // appends the submap trees (copies) to `submap`
  void getSubmap(const SE3& pose, std::vector<Cylinder>& submap);
  // the same submap by reference: pointers into the map, valid until the
  // next updateMap
  void getSubmap(const SE3& pose, MapView& submap);
  // obs_tms[i] with matches[i] == -1 becomes a new landmark; matches[i] = j
  // updates the landmark of entry j of the last submap
  void updateMap(std::vector<Cylinder>& obs_tms, const std::vector<int>& matches);

  size_t size() const { return records_.size(); }

 private:
  struct Record {
    Cylinder tree;
    float key[3];  // root as float: the query metric's coordinates
    uint32_t hits;
  };
  // map indices of the submap around pose, nearest first (into last_query_)
  void query(const SE3& pose);

  static constexpr size_t kNeighbours = 100;  // mapManager.cpp:51
  static constexpr size_t kRecent = 200;      // mapManager.cpp:57
  float sqSearchRadius_;  // the reference keeps it; its queries do not read it
  std::vector<Record> records_;
  std::vector<uint32_t> last_query_;  // submap entry -> record index
  std::vector<std::pair<float, uint32_t>> rank_;  // scratch of query()
};
