// Point cloud files in the PCD v0.7 format for x y z intensity clouds: the
// calls the reference makes through PCL (pcl::io::loadPCDFile<PointXYZI>,
// savePCDFileASCII; sloam/src/segmentation/test_inference.cpp:61-104), with
// PCL's return convention (0 on success, -1 on failure, a message on stderr).
//
// Reading: FIELDS are matched by name (x, y, z required; intensity optional,
// 0 when absent; other fields skipped), TYPE F (SIZE 4 or 8), I or U (SIZE
// 1, 2, 4 or 8), COUNT >= 1 (the first element of a named field is used),
// DATA ascii or binary (binary_compressed is refused); WIDTH x HEIGHT gives
// the organisation, is_dense is false when a point has a non-finite
// coordinate.
#pragma once

#include <string>

#include "definitions.h"

namespace sloam_io {

int loadPCDFile(const std::string& path, CloudT& cloud);
int savePCDFileASCII(const std::string& path, const CloudT& cloud);
// DATA binary: four little-endian float32 per point
int savePCDFileBinary(const std::string& path, const CloudT& cloud);

}  // namespace sloam_io
