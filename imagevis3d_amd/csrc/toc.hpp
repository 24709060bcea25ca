// toc.hpp -- bricked multi-resolution volumes in a UVF version-3 file with a
// table-of-contents block: the role of Tuvok's TOCBlock / ExtendedOctree
// (UVFReader/DataSource.h:475-491 builds one with FlatDataToBrickedLOD;
// UVFReader/BlockInfo.h:38-96 queries its LODs, brick counts and per-brick
// compression).
//
// ExtendedOctree's own byte layout is not in the reference tree, so the TOC
// block's payload here is this library's (documented in toc.cpp); the UVF
// master and block headers follow doc/import.adoc:904-1102.
#pragma once
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/tvk_hip.h"

namespace tvk {

struct TocLod {
  uint64_t dims[3] = {0, 0, 0};
  uint64_t grid[3] = {0, 0, 0};
  uint64_t first = 0;  // index of the level's first entry
};

struct TocEntry {
  uint64_t offset = 0;  // absolute file offset of the stored brick
  uint64_t stored = 0;  // stored (compressed) bytes
  uint64_t raw = 0;     // bytes after decompression: S^3 * bpv
  uint32_t compression = 0;
  uint32_t reserved = 0;
  float mn = 0.f, mx = 0.f;  // value range over the stored voxels
};

class TocVolume {
 public:
  ~TocVolume();
  int Open(const std::string& path);
  int ReadBrick(uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out, uint64_t bytes);
  void Info(tvk_toc_info* out) const;

  uint64_t dims[3] = {0, 0, 0};
  uint32_t bits = 0, max_brick = 0, overlap = 0, payload = 0, n_lod = 0;
  uint32_t median = 0, clamp = 0, compression = 0, level = 0, layout = 0;
  std::vector<TocLod> lods;
  std::vector<TocEntry> entries;
  uint64_t file_bytes = 0;
  // Histogram1D / Histogram2D blocks of the file (empty when absent)
  std::vector<uint64_t> hist1d;
  std::vector<uint64_t> hist2d;  // hist2d_x (value, slowest) x hist2d_y
  uint64_t hist2d_x = 0, hist2d_y = 0;
  float max_grad = 0.f;

 private:
  int fd_ = -1;
};

// Writes the TOC block followed by a Histogram1D block (2^bits bins) and a
// Histogram2D block (256 value x 256 gradient bins, the dataset's maximum
// gradient), computed on the device with the renderer's histogram rules.
int toc_build(int device, const void* voxels, const uint64_t dims[3], uint32_t bits,
              uint32_t max_brick, uint32_t overlap, bool median, bool clamp, uint32_t compression,
              uint32_t level, uint32_t layout, const std::string& path, tvk_toc_info* info);

// Bricks of a (gx, gy, gz) grid in the on-disk order of `layout`
// (scanline, Morton, Hilbert, seeded random); returns scanline indices.
std::vector<uint64_t> toc_layout_order(uint32_t layout, const uint64_t grid[3], uint32_t lod);

// Codec round trip used by the builder and the reader (exposed for tests).
int toc_compress(uint32_t codec, uint32_t level, const void* src, uint64_t n, std::vector<uint8_t>* out);
int toc_decompress(uint32_t codec, const void* src, uint64_t n, void* dst, uint64_t raw);

}  // namespace tvk
