// ROS 1 wire format of SLOAM's messages (see rosmsg.h).
#include "rosmsg.h"

#include <cmath>
#include <cstring>
#include <limits>

namespace sloam_wire {
namespace {

// ---- writing ---------------------------------------------------------------
template <typename T>
void put(std::vector<uint8_t>& out, T v) {
  static_assert(std::is_arithmetic<T>::value, "scalar");
  uint8_t b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));  // the hosts this builds for are little-endian
  out.insert(out.end(), b, b + sizeof(T));
}

void put_bool(std::vector<uint8_t>& out, bool v) { out.push_back(v ? 1 : 0); }

void put_string(std::vector<uint8_t>& out, const std::string& s) {
  put<uint32_t>(out, (uint32_t)s.size());
  out.insert(out.end(), s.begin(), s.end());
}

void put_time(std::vector<uint8_t>& out, const Time& t) {
  put(out, t.sec);
  put(out, t.nsec);
}

// ---- reading: a cursor that fails once any read runs past the end ----------
struct Reader {
  const uint8_t* p;
  size_t left;
  bool ok = true;

  bool take(void* dst, size_t n) {
    if (!ok || n > left) {
      ok = false;
      return false;
    }
    std::memcpy(dst, p, n);
    p += n;
    left -= n;
    return true;
  }
  template <typename T>
  T get() {
    T v{};
    take(&v, sizeof(T));
    return v;
  }
  bool get_bool() { return get<uint8_t>() != 0; }
  std::string get_string() {
    const uint32_t n = get<uint32_t>();
    if (!ok || n > left) {
      ok = false;
      return std::string();
    }
    std::string s(reinterpret_cast<const char*>(p), n);
    p += n;
    left -= n;
    return s;
  }
  Time get_time() {
    Time t;
    t.sec = get<uint32_t>();
    t.nsec = get<uint32_t>();
    return t;
  }
  // element count of a T[] whose elements take at least min_bytes each
  uint32_t get_count(size_t min_bytes) {
    const uint32_t n = get<uint32_t>();
    if (ok && min_bytes > 0 && (size_t)n > left / min_bytes) ok = false;
    return ok ? n : 0;
  }
};

// ---- per-type field lists (one function per message, both directions) -----
#define SLOAM_WIRE_FWD(T)                         \
  void write(std::vector<uint8_t>& o, const T& m); \
  void read(Reader& r, T& m);
SLOAM_WIRE_FWD(Header)
SLOAM_WIRE_FWD(Pose)
SLOAM_WIRE_FWD(PoseStamped)
SLOAM_WIRE_FWD(PoseWithCovarianceStamped)
SLOAM_WIRE_FWD(PointField)
SLOAM_WIRE_FWD(PointCloud2)
SLOAM_WIRE_FWD(ROSCylinder)
SLOAM_WIRE_FWD(ROSGround)
SLOAM_WIRE_FWD(ROSObservation)
SLOAM_WIRE_FWD(ROSScan)
SLOAM_WIRE_FWD(ROSSubMap)
SLOAM_WIRE_FWD(ROSSweep)
SLOAM_WIRE_FWD(LoopClosure)
SLOAM_WIRE_FWD(ObservationPair)
#undef SLOAM_WIRE_FWD

void write(std::vector<uint8_t>& o, const Header& m) {
  put(o, m.seq);
  put_time(o, m.stamp);
  put_string(o, m.frame_id);
}
void read(Reader& r, Header& m) {
  m.seq = r.get<uint32_t>();
  m.stamp = r.get_time();
  m.frame_id = r.get_string();
}

void write(std::vector<uint8_t>& o, const Pose& m) {
  for (double v : m.position) put(o, v);
  for (double v : m.orientation) put(o, v);
}
void read(Reader& r, Pose& m) {
  for (double& v : m.position) v = r.get<double>();
  for (double& v : m.orientation) v = r.get<double>();
}

void write(std::vector<uint8_t>& o, const PoseStamped& m) {
  write(o, m.header);
  write(o, m.pose);
}
void read(Reader& r, PoseStamped& m) {
  read(r, m.header);
  read(r, m.pose);
}

void write(std::vector<uint8_t>& o, const PoseWithCovarianceStamped& m) {
  write(o, m.header);
  write(o, m.pose);
  for (double v : m.covariance) put(o, v);
}
void read(Reader& r, PoseWithCovarianceStamped& m) {
  read(r, m.header);
  read(r, m.pose);
  for (double& v : m.covariance) v = r.get<double>();
}

void write(std::vector<uint8_t>& o, const PointField& m) {
  put_string(o, m.name);
  put(o, m.offset);
  put(o, m.datatype);
  put(o, m.count);
}
void read(Reader& r, PointField& m) {
  m.name = r.get_string();
  m.offset = r.get<uint32_t>();
  m.datatype = r.get<uint8_t>();
  m.count = r.get<uint32_t>();
}

void write(std::vector<uint8_t>& o, const PointCloud2& m) {
  write(o, m.header);
  put(o, m.height);
  put(o, m.width);
  put<uint32_t>(o, (uint32_t)m.fields.size());
  for (const auto& f : m.fields) write(o, f);
  put_bool(o, m.is_bigendian);
  put(o, m.point_step);
  put(o, m.row_step);
  put<uint32_t>(o, (uint32_t)m.data.size());
  o.insert(o.end(), m.data.begin(), m.data.end());
  put_bool(o, m.is_dense);
}
void read(Reader& r, PointCloud2& m) {
  read(r, m.header);
  m.height = r.get<uint32_t>();
  m.width = r.get<uint32_t>();
  m.fields.resize(r.get_count(13));  // a field takes at least 13 bytes
  for (auto& f : m.fields) read(r, f);
  m.is_bigendian = r.get_bool();
  m.point_step = r.get<uint32_t>();
  m.row_step = r.get<uint32_t>();
  const uint32_t n = r.get_count(1);
  m.data.resize(n);
  if (n) r.take(m.data.data(), n);
  m.is_dense = r.get_bool();
}

void write(std::vector<uint8_t>& o, const ROSCylinder& m) {
  for (float v : m.root) put(o, v);
  for (float v : m.ray) put(o, v);
  put<uint32_t>(o, (uint32_t)m.radii.size());
  for (double v : m.radii) put(o, v);
  put(o, m.radius);
  put(o, m.id);
}
void read(Reader& r, ROSCylinder& m) {
  for (float& v : m.root) v = r.get<float>();
  for (float& v : m.ray) v = r.get<float>();
  m.radii.resize(r.get_count(8));
  for (double& v : m.radii) v = r.get<double>();
  m.radius = r.get<float>();
  m.id = r.get<int64_t>();
}

void write(std::vector<uint8_t>& o, const ROSGround& m) {
  for (float v : m.coefs) put(o, v);
  write(o, m.features);
  put(o, m.id);
}
void read(Reader& r, ROSGround& m) {
  for (float& v : m.coefs) v = r.get<float>();
  read(r, m.features);
  m.id = r.get<int64_t>();
}

template <typename T>
void write_array(std::vector<uint8_t>& o, const std::vector<T>& v) {
  put<uint32_t>(o, (uint32_t)v.size());
  for (const auto& e : v) write(o, e);
}
template <typename T>
void read_array(Reader& r, std::vector<T>& v, size_t min_bytes) {
  v.resize(r.get_count(min_bytes));
  for (auto& e : v) {
    read(r, e);
    if (!r.ok) break;
  }
}

void write(std::vector<uint8_t>& o, const ROSObservation& m) {
  write(o, m.header);
  write(o, m.pose);
  write(o, m.initialGuess);
  write(o, m.pc);
  write_array(o, m.treeModels);
  write(o, m.ground);
  put<uint32_t>(o, (uint32_t)m.matches.size());
  for (int32_t v : m.matches) put(o, v);
  put_bool(o, m.success);
}
void read(Reader& r, ROSObservation& m) {
  read(r, m.header);
  read(r, m.pose);
  read(r, m.initialGuess);
  read(r, m.pc);
  read_array(r, m.treeModels, 40);  // a cylinder takes at least 40 bytes
  read(r, m.ground);
  m.matches.resize(r.get_count(4));
  for (int32_t& v : m.matches) v = r.get<int32_t>();
  m.success = r.get_bool();
}

void write(std::vector<uint8_t>& o, const ROSScan& m) {
  write_array(o, m.treeModels);
  write(o, m.treeFeatures);
  write(o, m.groundFeatures);
}
void read(Reader& r, ROSScan& m) {
  read_array(r, m.treeModels, 40);
  read(r, m.treeFeatures);
  read(r, m.groundFeatures);
}

void write(std::vector<uint8_t>& o, const ROSSubMap& m) {
  write(o, m.header);
  write_array(o, m.treeModels);
  write(o, m.treeFeatures);
  write(o, m.groundFeatures);
  write(o, m.t_map_anchor);
}
void read(Reader& r, ROSSubMap& m) {
  read(r, m.header);
  read_array(r, m.treeModels, 40);
  read(r, m.treeFeatures);
  read(r, m.groundFeatures);
  read(r, m.t_map_anchor);
}

void write(std::vector<uint8_t>& o, const ROSSweep& m) {
  write(o, m.header);
  write_array(o, m.scans);
  write(o, m.t_anchor_curr);
  write(o, m.t_global_curr);
}
void read(Reader& r, ROSSweep& m) {
  read(r, m.header);
  read_array(r, m.scans, 4);
  read(r, m.t_anchor_curr);
  read(r, m.t_global_curr);
}

void write(std::vector<uint8_t>& o, const LoopClosure& m) {
  write(o, m.header);
  write(o, m.pose);
  put(o, m.source_idx);
  put(o, m.target_idx);
}
void read(Reader& r, LoopClosure& m) {
  read(r, m.header);
  read(r, m.pose);
  m.source_idx = r.get<int32_t>();
  m.target_idx = r.get<int32_t>();
}

void write(std::vector<uint8_t>& o, const ObservationPair& m) {
  write(o, m.source);
  write(o, m.target);
  put(o, m.source_idx);
  put(o, m.target_idx);
}
void read(Reader& r, ObservationPair& m) {
  read(r, m.source);
  read(r, m.target);
  m.source_idx = r.get<int32_t>();
  m.target_idx = r.get<int32_t>();
}

// ---- PointCloud2 field access --------------------------------------------
size_t datatype_size(uint8_t t) {
  switch (t) {
    case INT8:
    case UINT8: return 1;
    case INT16:
    case UINT16: return 2;
    case INT32:
    case UINT32:
    case FLOAT32: return 4;
    case FLOAT64: return 8;
    default: return 0;
  }
}

float read_as_float(const uint8_t* p, uint8_t t) {
  switch (t) {
    case INT8: { int8_t v; std::memcpy(&v, p, 1); return (float)v; }
    case UINT8: return (float)p[0];
    case INT16: { int16_t v; std::memcpy(&v, p, 2); return (float)v; }
    case UINT16: { uint16_t v; std::memcpy(&v, p, 2); return (float)v; }
    case INT32: { int32_t v; std::memcpy(&v, p, 4); return (float)v; }
    case UINT32: { uint32_t v; std::memcpy(&v, p, 4); return (float)v; }
    case FLOAT32: { float v; std::memcpy(&v, p, 4); return v; }
    case FLOAT64: { double v; std::memcpy(&v, p, 8); return (float)v; }
    default: return 0.f;
  }
}

}  // namespace

#define SLOAM_WIRE_DEF(T)                                                      \
  void serialize(const T& m, std::vector<uint8_t>& out) { write(out, m); }     \
  bool deserialize(const uint8_t* data, size_t size, T& m) {                   \
    Reader r{data, size};                                                      \
    T tmp;                                                                     \
    read(r, tmp);                                                              \
    if (!r.ok || r.left != 0) return false;                                    \
    m = std::move(tmp);                                                        \
    return true;                                                               \
  }                                                                            \
  size_t serializedLength(const T& m) {                                        \
    std::vector<uint8_t> b;                                                    \
    write(b, m);                                                               \
    return b.size();                                                           \
  }
SLOAM_WIRE_DEF(Header)
SLOAM_WIRE_DEF(Pose)
SLOAM_WIRE_DEF(PoseStamped)
SLOAM_WIRE_DEF(PoseWithCovarianceStamped)
SLOAM_WIRE_DEF(PointField)
SLOAM_WIRE_DEF(PointCloud2)
SLOAM_WIRE_DEF(ROSCylinder)
SLOAM_WIRE_DEF(ROSGround)
SLOAM_WIRE_DEF(ROSObservation)
SLOAM_WIRE_DEF(ROSScan)
SLOAM_WIRE_DEF(ROSSubMap)
SLOAM_WIRE_DEF(ROSSweep)
SLOAM_WIRE_DEF(LoopClosure)
SLOAM_WIRE_DEF(ObservationPair)
#undef SLOAM_WIRE_DEF

void toROSMsg(const CloudT& cloud, PointCloud2& msg) {
  msg.height = cloud.height;
  msg.width = cloud.width;
  if ((size_t)cloud.width * cloud.height != cloud.points.size()) {
    // an unorganised cloud whose width was not kept up to date
    msg.height = 1;
    msg.width = (uint32_t)cloud.points.size();
  }
  const char* names[4] = {"x", "y", "z", "intensity"};
  const uint32_t offs[4] = {0, 4, 8, 16};
  msg.fields.clear();
  for (int i = 0; i < 4; ++i) msg.fields.push_back(PointField{names[i], offs[i], FLOAT32, 1});
  msg.is_bigendian = false;
  msg.point_step = 32;
  msg.row_step = msg.point_step * msg.width;
  msg.data.assign((size_t)msg.point_step * cloud.points.size(), 0);
  for (size_t i = 0; i < cloud.points.size(); ++i) {
    uint8_t* p = msg.data.data() + i * msg.point_step;
    const PointT& q = cloud.points[i];
    std::memcpy(p + 0, &q.x, 4);
    std::memcpy(p + 4, &q.y, 4);
    std::memcpy(p + 8, &q.z, 4);
    // PCL's PointXYZI pads the xyz block to 16 bytes with a 1.0f
    const float one = 1.0f;
    std::memcpy(p + 12, &one, 4);
    std::memcpy(p + 16, &q.intensity, 4);
  }
  msg.is_dense = cloud.is_dense;
}

bool fromROSMsg(const PointCloud2& msg, CloudT& cloud) {
  if (msg.is_bigendian) return false;
  const size_t n = (size_t)msg.width * msg.height;
  if (msg.height > 0 && msg.row_step < (size_t)msg.point_step * msg.width) return false;
  if (msg.height > 0 && msg.data.size() < (size_t)msg.row_step * (msg.height - 1) +
                                               (size_t)msg.point_step * msg.width)
    return false;
  const char* names[4] = {"x", "y", "z", "intensity"};
  int at[4] = {-1, -1, -1, -1};
  uint8_t type[4] = {0, 0, 0, 0};
  for (const auto& f : msg.fields)
    for (int k = 0; k < 4; ++k)
      if (f.name == names[k] && at[k] < 0) {
        const size_t sz = datatype_size(f.datatype);
        if (sz == 0 || (size_t)f.offset + sz > msg.point_step) return false;
        at[k] = (int)f.offset;
        type[k] = f.datatype;
      }
  CloudT out;
  out.width = msg.width;
  out.height = msg.height;
  out.points.resize(n);
  bool dense = true;
  for (uint32_t row = 0; row < msg.height; ++row)
    for (uint32_t col = 0; col < msg.width; ++col) {
      const uint8_t* p = msg.data.data() + (size_t)row * msg.row_step + (size_t)col * msg.point_step;
      PointT& q = out.points[(size_t)row * msg.width + col];
      float* dst[4] = {&q.x, &q.y, &q.z, &q.intensity};
      for (int k = 0; k < 4; ++k) *dst[k] = at[k] >= 0 ? read_as_float(p + at[k], type[k]) : 0.f;
      if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) dense = false;
    }
  out.is_dense = msg.is_dense && dense;
  cloud = std::move(out);
  return true;
}

ROSCylinder toROSCylinder(const Cylinder& c) {
  ROSCylinder m;
  for (int i = 0; i < 3; ++i) {
    m.root[i] = (float)c.model.root[i];
    m.ray[i] = (float)c.model.ray[i];
  }
  m.radii = c.model.radii;
  m.radius = (float)c.model.radius;
  m.id = (int64_t)c.id;
  return m;
}

ROSGround toROSGround(const Plane& p) {
  ROSGround m;
  for (int i = 0; i < 4; ++i) m.coefs[i] = (float)p.model.plane[i];
  CloudT f;
  for (const auto& q : p.features) f.push_back(q);
  toROSMsg(f, m.features);
  m.id = (int64_t)p.id;
  return m;
}

Pose toPose(const SE3& tf) {
  Pose p;
  for (int i = 0; i < 3; ++i) p.position[i] = tf.translation()[i];
  const Quat& q = tf.unit_quaternion();
  p.orientation[0] = q.x;
  p.orientation[1] = q.y;
  p.orientation[2] = q.z;
  p.orientation[3] = q.w;
  return p;
}

SE3 fromPose(const Pose& p) {
  return SE3(Quat(p.orientation[3], p.orientation[0], p.orientation[1], p.orientation[2]),
             Vector3(p.position[0], p.position[1], p.position[2]));
}

}  // namespace sloam_wire
