// PCD v0.7 reading and writing for x y z intensity clouds.
#include "pcdIO.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace sloam_io {
namespace {

int fail(const std::string& path, const std::string& why) {
  std::fprintf(stderr, "PCD %s: %s\n", path.c_str(), why.c_str());
  return -1;
}

// largest cloud accepted (points, width, height) and values per field
constexpr long long kMaxPoints = 0x7fffffffll;
constexpr int kMaxFieldCount = 1 << 16;

struct Field {
  std::string name;
  int size = 4;
  char type = 'F';
  int count = 1;
  size_t offset = 0;  // bytes into a binary record
  int token = 0;      // first token of the field in an ascii line
};

// one value of a field from a binary record
double binary_value(const unsigned char* p, const Field& f) {
  switch (f.type) {
    case 'F':
      if (f.size == 4) {
        float v;
        std::memcpy(&v, p, 4);
        return v;
      } else {
        double v;
        std::memcpy(&v, p, 8);
        return v;
      }
    case 'I':
      switch (f.size) {
        case 1: return (double)*(const int8_t*)p;
        case 2: { int16_t v; std::memcpy(&v, p, 2); return v; }
        case 4: { int32_t v; std::memcpy(&v, p, 4); return v; }
        default: { int64_t v; std::memcpy(&v, p, 8); return (double)v; }
      }
    default:
      switch (f.size) {
        case 1: return (double)*p;
        case 2: { uint16_t v; std::memcpy(&v, p, 2); return v; }
        case 4: { uint32_t v; std::memcpy(&v, p, 4); return v; }
        default: { uint64_t v; std::memcpy(&v, p, 8); return (double)v; }
      }
  }
}

}  // namespace

int loadPCDFile(const std::string& path, CloudT& cloud) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(path, "cannot open");
  std::vector<Field> fields;
  constexpr size_t kMaxFields = 1024;
  std::vector<int> sizes, counts;
  std::vector<char> types;
  long long width = -1, height = 1, points = -1;
  std::string data, line;
  while (data.empty() && std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    std::string tok;
    if (key == "FIELDS" || key == "COLUMNS") {
      while (ss >> tok) {
        if (fields.size() >= kMaxFields) return fail(path, "too many FIELDS");
        Field f;
        f.name = tok;
        fields.push_back(f);
      }
    } else if (key == "SIZE") {
      while (ss >> tok) sizes.push_back(std::atoi(tok.c_str()));
    } else if (key == "TYPE") {
      while (ss >> tok) types.push_back(tok.empty() ? '?' : tok[0]);
    } else if (key == "COUNT") {
      while (ss >> tok) counts.push_back(std::atoi(tok.c_str()));
    } else if (key == "WIDTH") {
      ss >> width;
    } else if (key == "HEIGHT") {
      ss >> height;
    } else if (key == "POINTS") {
      ss >> points;
    } else if (key == "DATA") {
      ss >> data;
    }
  }
  if (data.empty()) return fail(path, "no DATA line");
  if (data != "ascii" && data != "binary")
    return fail(path, "DATA " + data + " is not supported (ascii or binary)");
  if (fields.empty()) return fail(path, "no FIELDS");
  if (sizes.size() != fields.size() || types.size() != fields.size())
    return fail(path, "SIZE / TYPE do not match FIELDS");
  if (counts.empty()) counts.assign(fields.size(), 1);
  if (counts.size() != fields.size()) return fail(path, "COUNT does not match FIELDS");
  size_t rec = 0;
  int token = 0;
  for (size_t i = 0; i < fields.size(); ++i) {
    Field& f = fields[i];
    f.size = sizes[i];
    f.type = types[i];
    f.count = counts[i];
    const bool ok_type = (f.type == 'F' && (f.size == 4 || f.size == 8)) ||
                         ((f.type == 'I' || f.type == 'U') &&
                          (f.size == 1 || f.size == 2 || f.size == 4 || f.size == 8));
    if (!ok_type || f.count < 1 || f.count > kMaxFieldCount)
      return fail(path, "unsupported field " + f.name);
    f.offset = rec;
    f.token = token;
    rec += (size_t)f.size * f.count;
    token += f.count;
  }
  // bounds before any product or allocation: a header may claim any size
  if (width < 0 || height < 1 || width > kMaxPoints || height > kMaxPoints)
    return fail(path, "bad WIDTH / HEIGHT");
  if (width * height > kMaxPoints) return fail(path, "more than 2^31 - 1 points");
  if (points < 0) points = width * height;
  if (points != width * height) return fail(path, "POINTS != WIDTH * HEIGHT");
  // the data must fit in what is left of the file (binary: rec bytes per
  // point; ascii: at least one byte per value), checked before allocating
  const std::streampos here = in.tellg();
  in.seekg(0, std::ios::end);
  const std::streampos end = in.tellg();
  in.seekg(here);
  if (here < 0 || end < here) return fail(path, "cannot size the data section");
  const unsigned long long left = (unsigned long long)(end - here);
  const unsigned long long per_point = data == "binary" ? (unsigned long long)rec
                                                        : (unsigned long long)token;
  if ((unsigned long long)points > 0 && per_point > left / (unsigned long long)points)
    return fail(path, "data section shorter than POINTS records");
  const Field* fx = nullptr;
  const Field* fy = nullptr;
  const Field* fz = nullptr;
  const Field* fi = nullptr;
  for (const Field& f : fields) {
    if (f.name == "x") fx = &f;
    if (f.name == "y") fy = &f;
    if (f.name == "z") fz = &f;
    if (f.name == "intensity") fi = &f;
  }
  if (!fx || !fy || !fz) return fail(path, "x, y and z are required");
  CloudT out;
  out.points.resize((size_t)points);
  out.width = (uint32_t)width;
  out.height = (uint32_t)height;
  bool dense = true;
  if (data == "ascii") {
    std::vector<double> vals((size_t)token);
    for (long long p = 0; p < points; ++p) {
      for (int t = 0; t < token; ++t) {
        std::string s;
        if (!(in >> s)) return fail(path, "short ascii data");
        // strtod reads nan / inf in any case
        vals[(size_t)t] = std::strtod(s.c_str(), nullptr);
      }
      PointT& q = out.points[(size_t)p];
      q.x = (float)vals[(size_t)fx->token];
      q.y = (float)vals[(size_t)fy->token];
      q.z = (float)vals[(size_t)fz->token];
      q.intensity = fi ? (float)vals[(size_t)fi->token] : 0.f;
    }
  } else if (data == "binary") {
    std::vector<unsigned char> buf(rec * (size_t)points);
    if (!buf.empty() && !in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)buf.size()))
      return fail(path, "short binary data");
    for (long long p = 0; p < points; ++p) {
      const unsigned char* r = buf.data() + rec * (size_t)p;
      PointT& q = out.points[(size_t)p];
      q.x = (float)binary_value(r + fx->offset, *fx);
      q.y = (float)binary_value(r + fy->offset, *fy);
      q.z = (float)binary_value(r + fz->offset, *fz);
      q.intensity = fi ? (float)binary_value(r + fi->offset, *fi) : 0.f;
    }
  } else {
    return fail(path, "DATA " + data + " is not supported (ascii or binary)");
  }
  for (const PointT& q : out.points)
    if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) dense = false;
  out.is_dense = dense;
  cloud = std::move(out);
  return 0;
}

namespace {

int save(const std::string& path, const CloudT& cloud, bool binary) {
  const size_t n = cloud.points.size();
  const uint32_t width = cloud.width, height = cloud.height;
  if ((size_t)width * height != n) return fail(path, "width * height != points");
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return fail(path, "cannot open for writing");
  std::fprintf(f,
               "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z intensity\n"
               "SIZE 4 4 4 4\nTYPE F F F F\nCOUNT 1 1 1 1\nWIDTH %u\nHEIGHT %u\n"
               "VIEWPOINT 0 0 0 1 0 0 0\nPOINTS %zu\nDATA %s\n",
               width, height, n, binary ? "binary" : "ascii");
  bool ok = true;
  if (binary) {
    static_assert(sizeof(PointT) == 16, "PointT is four floats");
    ok = n == 0 || std::fwrite(cloud.points.data(), sizeof(PointT), n, f) == n;
  } else {
    // nine significant digits: every float reads back exactly
    for (const PointT& q : cloud.points)
      ok = ok && std::fprintf(f, "%.9g %.9g %.9g %.9g\n", q.x, q.y, q.z, q.intensity) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : fail(path, "write failed");
}

}  // namespace

int savePCDFileASCII(const std::string& path, const CloudT& cloud) {
  return save(path, cloud, false);
}

int savePCDFileBinary(const std::string& path, const CloudT& cloud) {
  return save(path, cloud, true);
}

}  // namespace sloam_io
