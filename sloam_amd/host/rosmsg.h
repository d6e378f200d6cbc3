// ROS 1 wire format of the messages SLOAM exchanges, without ROS: the
// sensor_msgs/PointCloud2 scans the nodes subscribe to
// (sloam/src/core/inputNode.cpp:220-250, sloam/src/segmentation/inferenceNode.cpp:104)
// and the sloam_msgs types (sloam_msgs/msg/*.msg) with the geometry_msgs /
// std_msgs types they embed, as plain structs that serialise to and from the
// bytes rosbag and roscpp carry.
//
// ROS 1 serialisation: little-endian fixed-width integers and IEEE floats;
// bool as one byte; string and T[] as a uint32 count followed by the
// elements; T[N] as N elements; time as uint32 seconds, uint32 nanoseconds;
// nested messages inline in field order.
//
// The PointCloud2 conversions follow pcl::toROSMsg / pcl::fromROSMsg for
// pcl::PointXYZI: fields x, y, z at offsets 0, 4, 8 and intensity at 16
// (FLOAT32), point_step 32, row_step = point_step * width; reading matches
// fields by name at any offset and of any numeric datatype, leaves a field
// that is absent at 0, and takes the organisation from height x width.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "cylinder.h"
#include "definitions.h"
#include "plane.h"

namespace sloam_wire {

struct Time {
  uint32_t sec = 0, nsec = 0;
};

// std_msgs/Header
struct Header {
  uint32_t seq = 0;
  Time stamp;
  std::string frame_id;
};

// geometry_msgs/Pose (Point position, Quaternion orientation x y z w)
struct Pose {
  double position[3] = {0, 0, 0};
  double orientation[4] = {0, 0, 0, 1};
};

struct PoseStamped {
  Header header;
  Pose pose;
};

struct PoseWithCovarianceStamped {
  Header header;
  Pose pose;
  std::array<double, 36> covariance{};
};

// sensor_msgs/PointField datatypes
enum PointFieldType : uint8_t {
  INT8 = 1,
  UINT8 = 2,
  INT16 = 3,
  UINT16 = 4,
  INT32 = 5,
  UINT32 = 6,
  FLOAT32 = 7,
  FLOAT64 = 8,
};

struct PointField {
  std::string name;
  uint32_t offset = 0;
  uint8_t datatype = 0;
  uint32_t count = 0;
};

struct PointCloud2 {
  Header header;
  uint32_t height = 0, width = 0;
  std::vector<PointField> fields;
  bool is_bigendian = false;
  uint32_t point_step = 0, row_step = 0;
  std::vector<uint8_t> data;
  bool is_dense = false;
};

// sloam_msgs/ROSCylinder
struct ROSCylinder {
  float root[3] = {0, 0, 0};
  float ray[3] = {0, 0, 0};
  std::vector<double> radii;
  float radius = 0;
  int64_t id = 0;
};

// sloam_msgs/ROSGround
struct ROSGround {
  float coefs[4] = {0, 0, 0, 0};
  PointCloud2 features;
  int64_t id = 0;
};

// sloam_msgs/ROSObservation
struct ROSObservation {
  Header header;
  PoseStamped pose;
  PoseStamped initialGuess;
  PointCloud2 pc;
  std::vector<ROSCylinder> treeModels;
  ROSGround ground;
  std::vector<int32_t> matches;
  bool success = false;
};

// sloam_msgs/ROSScan
struct ROSScan {
  std::vector<ROSCylinder> treeModels;
  PointCloud2 treeFeatures;
  PointCloud2 groundFeatures;
};

// sloam_msgs/ROSSubMap
struct ROSSubMap {
  Header header;
  std::vector<ROSCylinder> treeModels;
  PointCloud2 treeFeatures;
  PointCloud2 groundFeatures;
  PoseWithCovarianceStamped t_map_anchor;
};

// sloam_msgs/ROSSweep
struct ROSSweep {
  Header header;
  std::vector<ROSScan> scans;
  PoseStamped t_anchor_curr;
  PoseStamped t_global_curr;
};

// sloam_msgs/LoopClosure
struct LoopClosure {
  Header header;
  Pose pose;
  int32_t source_idx = 0;
  int32_t target_idx = 0;
};

// sloam_msgs/ObservationPair
struct ObservationPair {
  ROSObservation source;
  ROSObservation target;
  int32_t source_idx = 0;
  int32_t target_idx = 0;
};

// Appends the message's ROS 1 serialisation to out.
#define SLOAM_WIRE_DECL(T)                                  \
  void serialize(const T& m, std::vector<uint8_t>& out);    \
  bool deserialize(const uint8_t* data, size_t size, T& m); \
  size_t serializedLength(const T& m);
SLOAM_WIRE_DECL(Header)
SLOAM_WIRE_DECL(Pose)
SLOAM_WIRE_DECL(PoseStamped)
SLOAM_WIRE_DECL(PoseWithCovarianceStamped)
SLOAM_WIRE_DECL(PointField)
SLOAM_WIRE_DECL(PointCloud2)
SLOAM_WIRE_DECL(ROSCylinder)
SLOAM_WIRE_DECL(ROSGround)
SLOAM_WIRE_DECL(ROSObservation)
SLOAM_WIRE_DECL(ROSScan)
SLOAM_WIRE_DECL(ROSSubMap)
SLOAM_WIRE_DECL(ROSSweep)
SLOAM_WIRE_DECL(LoopClosure)
SLOAM_WIRE_DECL(ObservationPair)
#undef SLOAM_WIRE_DECL
// deserialize(): true when the bytes hold exactly one whole message of the
// type (every count within the remaining bytes, nothing left over).

// pcl::toROSMsg / fromROSMsg for x y z intensity clouds (see above).
// fromROSMsg returns false for a malformed message (data shorter than
// height * row_step, a field past point_step, big-endian data) and fills
// nothing then.
void toROSMsg(const CloudT& cloud, PointCloud2& msg);
bool fromROSMsg(const PointCloud2& msg, CloudT& cloud);

// The model messages of the drop-in objects: a cylinder's root, ray, radius,
// radii and id (the fields sloamNode.cpp:139-147 fills), a ground plane's
// coefficients, features and id.
ROSCylinder toROSCylinder(const Cylinder& c);
ROSGround toROSGround(const Plane& p);
Pose toPose(const SE3& tf);
SE3 fromPose(const Pose& p);

}  // namespace sloam_wire
