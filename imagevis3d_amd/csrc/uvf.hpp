// uvf.hpp -- reader for UVF v2 files (the container Tuvok's UVFDataset
// opens in renderer.loadDataset, ImageVis3D/UI/RenderWindowGL.cpp:111).
// Layout per doc/import.adoc:904-1102; the tail of the RasterDataBlock
// header (element signedness, data offset) follows DemoData/c60.uvf.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace tvk {

struct UvfRaster {
  uint64_t dims[3] = {0, 0, 0};    // x, y, z
  uint64_t brick_size[3] = {0, 0, 0};
  uint64_t overlap[3] = {0, 0, 0};
  uint64_t lod_levels = 0;
  uint32_t bits = 0;
  bool is_signed = false;
  double transform[16] = {0};
  uint64_t data_offset = 0;        // file offset of the voxel payload
};

struct UvfFile {
  uint64_t version = 0;
  uint64_t checksum_type = 0;
  std::vector<uint8_t> checksum;
  bool checksum_ok = false;        // valid only when verified
  std::vector<UvfRaster> rasters;
  std::vector<uint64_t> histogram_1d;
  std::vector<uint64_t> histogram_2d;
  uint64_t histogram_2d_dims[2] = {0, 0};
  float max_gradient = 0.f;
};

// Parse `path`; with verify the MD5 over the bytes after the checksum is
// compared with the stored one.  Returns 0, or a tvk_status error code with
// a message in *err.
int uvf_read(const std::string& path, bool verify, UvfFile* out, std::string* err);

// Load the voxels of a single-brick, single-LOD raster block (x fastest).
int uvf_read_voxels(const std::string& path, const UvfRaster& r, std::vector<uint8_t>* out,
                    std::string* err);

// MD5 (RFC 1321) of a byte range.
void md5(const uint8_t* data, uint64_t n, uint8_t digest[16]);

}  // namespace tvk
