// host_volume.hpp -- host-side pyramid, brick ranges and brick extraction
// for paged (out-of-core) volumes; same rules as the GPU kernels.
#pragma once
#include <stdint.h>

namespace tvk {

// dst dims = ceil(dims / 2); each voxel = floor(sum / count) of its
// existing 2x2x2 children.  Host threads over rows.
void host_downsample(const void* src, int bpv, const int64_t dims[3], void* dst);

// One (B+2a)^3 slot of brick (bx,by,bz), clamp-to-edge outside the level.
void host_extract_brick(const void* level, int bpv, const int64_t dims[3], int B, int apron, int bx,
                        int by, int bz, void* slot);

// (min, max) per brick over its slot (payload + apron, clamped): 2 floats
// per brick in brick-grid order.
void host_brick_ranges(const void* level, int bpv, const int64_t dims[3], int B, int apron,
                       const int grid[3], float* out);

}  // namespace tvk
