// toc.cpp -- TOC (UVF version 3) volume files: builder, reader, codecs and
// on-disk brick orders.
//
// File layout (little endian):
//   UVF master header (doc/import.adoc:904-975): "UVF-DATA", endianness 0,
//   version 3, checksum type 0 with 0 checksum bytes, 0 bytes to skip.
//   One data block (doc/import.adoc:977-1018): u64 length + ASCII type
//   string, u64 semantic 9 (the TOC block; the in-tree table stops at 8),
//   u64 block compression 0, u64 offset to the next block 0 (last block),
//   then the TOC payload:
//     char[8] "TVKTOC01"; u32 bits, components (1);
//     u64 dims[3];
//     u32 max_brick, overlap, payload, n_lod;
//     u32 use_median, clamp_to_edge, compression, level, layout, reserved;
//     per LOD: u64 dims[3], grid[3], first entry;
//     per brick (LOD-major, then z, y, x): u64 offset (absolute), stored
//       bytes, raw bytes; u32 compression, reserved; f32 min, max;
//     the stored bricks, in the layout's order.
//   Then a Histogram1D block (semantic 5: u64 n, n x u64) and a
//   Histogram2D block (semantic 6: f32 max gradient, u64 elems_x, u64
//   elems_y, elems_x x elems_y u64 with x = value slowest), laid out as
//   doc/import.adoc:1021-1056 documents them; the TOC block's next-block
//   offset points at the first of them.
#include "toc.hpp"

#include <dlfcn.h>
#include <math.h>
#include <fcntl.h>
#include <hip/hip_runtime_api.h>
#include <string.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <thread>

#include "errors.hpp"
#include "kernels.hpp"
#include "renderer.hpp"

namespace tvk {
namespace {

constexpr char kTocMagic[8] = {'T', 'V', 'K', 'T', 'O', 'C', '0', '1'};
constexpr uint64_t kTocSemantic = 9;
constexpr uint32_t kCodecs = 5;  // none, zlib, lzma, lz4, bzip2

int fail(int code, const std::string& msg) {
  set_last_error(msg);
  return code;
}

#define TOC_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
      return e_ == hipErrorOutOfMemory ? TVK_ERR_OOM : TVK_ERR_HIP;        \
    }                                                                      \
  } while (0)

// ---------------------------------------------------------------- codecs
// zlib is linked; xz (liblzma), lz4 and bzip2 are loaded on first use from
// the system's shared libraries through their public C APIs.
struct OptionalCodecs {
  // liblzma
  int (*xz_encode)(uint32_t, int, const void*, const uint8_t*, size_t, uint8_t*, size_t*, size_t) = nullptr;
  int (*xz_decode)(uint64_t*, uint32_t, const void*, const uint8_t*, size_t*, size_t, uint8_t*, size_t*,
                   size_t) = nullptr;
  size_t (*xz_bound)(size_t) = nullptr;
  // liblz4
  int (*lz4_bound)(int) = nullptr;
  int (*lz4_encode)(const char*, char*, int, int, int) = nullptr;
  int (*lz4_decode)(const char*, char*, int, int) = nullptr;
  // libbz2
  int (*bz_encode)(char*, unsigned*, char*, unsigned, int, int, int) = nullptr;
  int (*bz_decode)(char*, unsigned*, char*, unsigned, int, int) = nullptr;

  OptionalCodecs() {
    if (void* h = dlopen("liblzma.so.5", RTLD_NOW | RTLD_LOCAL)) {
      xz_encode = (decltype(xz_encode))dlsym(h, "lzma_easy_buffer_encode");
      xz_decode = (decltype(xz_decode))dlsym(h, "lzma_stream_buffer_decode");
      xz_bound = (decltype(xz_bound))dlsym(h, "lzma_stream_buffer_bound");
    }
    if (void* h = dlopen("liblz4.so.1", RTLD_NOW | RTLD_LOCAL)) {
      lz4_bound = (decltype(lz4_bound))dlsym(h, "LZ4_compressBound");
      lz4_encode = (decltype(lz4_encode))dlsym(h, "LZ4_compress_fast");
      lz4_decode = (decltype(lz4_decode))dlsym(h, "LZ4_decompress_safe");
    }
    if (void* h = dlopen("libbz2.so.1", RTLD_NOW | RTLD_LOCAL)) {
      bz_encode = (decltype(bz_encode))dlsym(h, "BZ2_bzBuffToBuffCompress");
      bz_decode = (decltype(bz_decode))dlsym(h, "BZ2_bzBuffToBuffDecompress");
    }
  }
  bool has(uint32_t c) const {
    switch (c) {
      case 0:
      case 1: return true;
      case 2: return xz_encode && xz_decode && xz_bound;
      case 3: return lz4_bound && lz4_encode && lz4_decode;
      case 4: return bz_encode && bz_decode;
      default: return false;
    }
  }
};

const OptionalCodecs& codecs() {
  static const OptionalCodecs c;
  return c;
}

const char* codec_name(uint32_t c) {
  static const char* names[] = {"none", "zlib", "lzma", "lz4", "bzip2"};
  return c < kCodecs ? names[c] : "unknown";
}

// ------------------------------------------------------------ brick orders
uint64_t morton3(uint64_t x, uint64_t y, uint64_t z) {
  uint64_t m = 0;
  for (int b = 0; b < 21; ++b)
    m |= ((x >> b) & 1ull) << (3 * b) | ((y >> b) & 1ull) << (3 * b + 1) | ((z >> b) & 1ull) << (3 * b + 2);
  return m;
}

// Index along the 3-D Hilbert curve of a 2^bits grid, by J. Skilling's
// transpose method ("Programming the Hilbert curve", AIP Conf. Proc. 707,
// 2004): undo the curve's rotations/reflections axis by axis, Gray-encode,
// then interleave the transposed bits (axis 0 most significant).
uint64_t hilbert3(uint64_t x, uint64_t y, uint64_t z, int bits) {
  uint64_t X[3] = {x, y, z};
  const uint64_t M = 1ull << (bits - 1);
  for (uint64_t Q = M; Q > 1; Q >>= 1) {
    const uint64_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint64_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint64_t t = 0;
  for (uint64_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t h = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1ull);
  return h;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t x = (s += 0x9e3779b97f4a7c15ull);
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// ------------------------------------------------------------- file I/O
struct Writer {
  FILE* f = nullptr;
  bool ok = true;
  void raw(const void* p, size_t n) { ok = ok && fwrite(p, 1, n, f) == n; }
  void u64(uint64_t v) { raw(&v, 8); }
  void u32(uint32_t v) { raw(&v, 4); }
  void f32(float v) { raw(&v, 4); }
};

struct Reader {
  int fd;
  uint64_t pos;
  uint64_t size;
  bool ok = true;
  void raw(void* p, size_t n) {
    if (!ok || pos + n > size || pread(fd, p, n, (off_t)pos) != (ssize_t)n) {
      ok = false;
      return;
    }
    pos += n;
  }
  uint64_t u64() {
    uint64_t v = 0;
    raw(&v, 8);
    return v;
  }
  uint32_t u32() {
    uint32_t v = 0;
    raw(&v, 4);
    return v;
  }
  float f32() {
    float v = 0;
    raw(&v, 4);
    return v;
  }
};

constexpr uint64_t kEntryBytes = 40;
const char kBlockName[] = "Tuvok MI355X TOC volume";
const char kHist1Name[] = "1D Histogram";
const char kHist2Name[] = "2D Histogram";
constexpr uint64_t kMasterBytes = 8 + 1 + 8 + 8 + 8 + 8;

uint64_t header_bytes(uint32_t n_lod) {
  const uint64_t master = 8 + 1 + 8 + 8 + 8 + 8;
  const uint64_t block = 8 + (sizeof(kBlockName) - 1) + 8 + 8 + 8;
  const uint64_t toc = 8 + 4 + 4 + 24 + 16 + 24 + (uint64_t)n_lod * 56;
  return master + block + toc;
}

}  // namespace

// ------------------------------------------------------------- codecs API
int toc_compress(uint32_t codec, uint32_t level, const void* src, uint64_t n, std::vector<uint8_t>* out) {
  const OptionalCodecs& C = codecs();
  if (!C.has(codec)) return fail(TVK_ERR_ARG, std::string("compression ") + codec_name(codec) + " is not available");
  const int lv = (int)std::min<uint32_t>(std::max<uint32_t>(level, 1), 10);
  if (n > (1ull << 31)) return fail(TVK_ERR_ARG, "brick too large to compress");
  switch (codec) {
    case 0:
      out->assign((const uint8_t*)src, (const uint8_t*)src + n);
      return TVK_OK;
    case 1: {
      uLongf len = compressBound((uLong)n);
      out->resize(len);
      if (compress2(out->data(), &len, (const Bytef*)src, (uLong)n, std::min(lv, 9)) != Z_OK)
        return fail(TVK_ERR_IO, "zlib compression failed");
      out->resize(len);
      return TVK_OK;
    }
    case 2: {
      size_t pos = 0;
      out->resize(C.xz_bound((size_t)n));
      // preset 0..9 from level 1..10, CRC64 check (LZMA_CHECK_CRC64 = 4)
      if (C.xz_encode((uint32_t)(lv - 1), 4, nullptr, (const uint8_t*)src, (size_t)n, out->data(), &pos,
                      out->size()) != 0)
        return fail(TVK_ERR_IO, "lzma compression failed");
      out->resize(pos);
      return TVK_OK;
    }
    case 3: {
      out->resize((size_t)C.lz4_bound((int)n));
      // acceleration 11 - level: level 10 = default effort, 1 = fastest
      const int len = C.lz4_encode((const char*)src, (char*)out->data(), (int)n, (int)out->size(), 11 - lv);
      if (len <= 0) return fail(TVK_ERR_IO, "lz4 compression failed");
      out->resize((size_t)len);
      return TVK_OK;
    }
    case 4: {
      unsigned len = (unsigned)(n + n / 100 + 601);
      out->resize(len);
      if (C.bz_encode((char*)out->data(), &len, (char*)src, (unsigned)n, std::min(lv, 9), 0, 0) != 0)
        return fail(TVK_ERR_IO, "bzip2 compression failed");
      out->resize(len);
      return TVK_OK;
    }
  }
  return fail(TVK_ERR_ARG, "unknown compression");
}

int toc_decompress(uint32_t codec, const void* src, uint64_t n, void* dst, uint64_t raw) {
  const OptionalCodecs& C = codecs();
  if (!C.has(codec)) return fail(TVK_ERR_FORMAT, std::string("compression ") + codec_name(codec) + " is not available");
  switch (codec) {
    case 0:
      if (n != raw) return fail(TVK_ERR_FORMAT, "stored brick size mismatch");
      memcpy(dst, src, n);
      return TVK_OK;
    case 1: {
      uLongf len = (uLongf)raw;
      if (uncompress((Bytef*)dst, &len, (const Bytef*)src, (uLong)n) != Z_OK || len != raw)
        return fail(TVK_ERR_FORMAT, "zlib brick is corrupt");
      return TVK_OK;
    }
    case 2: {
      uint64_t memlimit = UINT64_MAX;
      size_t in_pos = 0, out_pos = 0;
      if (C.xz_decode(&memlimit, 0, nullptr, (const uint8_t*)src, &in_pos, (size_t)n, (uint8_t*)dst, &out_pos,
                      (size_t)raw) != 0 ||
          out_pos != raw)
        return fail(TVK_ERR_FORMAT, "lzma brick is corrupt");
      return TVK_OK;
    }
    case 3: {
      const int len = C.lz4_decode((const char*)src, (char*)dst, (int)n, (int)raw);
      if (len < 0 || (uint64_t)len != raw) return fail(TVK_ERR_FORMAT, "lz4 brick is corrupt");
      return TVK_OK;
    }
    case 4: {
      unsigned len = (unsigned)raw;
      if (C.bz_decode((char*)dst, &len, (char*)src, (unsigned)n, 0, 0) != 0 || len != raw)
        return fail(TVK_ERR_FORMAT, "bzip2 brick is corrupt");
      return TVK_OK;
    }
  }
  return fail(TVK_ERR_FORMAT, "unknown compression");
}

std::vector<uint64_t> toc_layout_order(uint32_t layout, const uint64_t grid[3], uint32_t lod) {
  const uint64_t n = grid[0] * grid[1] * grid[2];
  std::vector<uint64_t> order(n);
  for (uint64_t i = 0; i < n; ++i) order[i] = i;
  if (layout == 1 || layout == 2) {
    int bits = 1;
    while ((1ull << bits) < std::max(grid[0], std::max(grid[1], grid[2]))) ++bits;
    std::vector<uint64_t> key(n);
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t x = i % grid[0], y = (i / grid[0]) % grid[1], z = i / (grid[0] * grid[1]);
      key[i] = layout == 1 ? morton3(x, y, z) : hilbert3(x, y, z, bits);
    }
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return key[a] < key[b]; });
  } else if (layout == 3) {
    uint64_t s = 0x70c5eedull ^ ((uint64_t)lod << 32);
    for (uint64_t i = n; i > 1; --i) std::swap(order[i - 1], order[splitmix(s) % i]);
  }
  return order;
}

// ---------------------------------------------------------------- builder
int toc_build(int device, const void* voxels, const uint64_t dims[3], uint32_t bits,
              uint32_t max_brick, uint32_t overlap, bool median, bool clamp, uint32_t compression,
              uint32_t level, uint32_t layout, const std::string& path, tvk_toc_info* info) {
  if (!voxels || (bits != 8 && bits != 16) || !dims[0] || !dims[1] || !dims[2])
    return fail(TVK_ERR_ARG, "bad volume for the TOC builder");
  if (dims[0] > (1u << 20) || dims[1] > (1u << 20) || dims[2] > (1u << 20))
    return fail(TVK_ERR_ARG, "volume dimension too large");
  if (max_brick > 1024 || 2 * overlap >= max_brick || max_brick - 2 * overlap < 2)
    return fail(TVK_ERR_ARG, "max brick size must exceed twice the overlap (payload >= 2) and be <= 1024");
  if (compression >= kCodecs) return fail(TVK_ERR_ARG, "compression must be 0..4");
  if (layout > 3) return fail(TVK_ERR_ARG, "layout must be 0..3");
  if (!codecs().has(compression))
    return fail(TVK_ERR_ARG, std::string("compression ") + codec_name(compression) + " is not available");
  const int bpv = (int)bits / 8;
  const uint32_t P = max_brick - 2 * overlap;
  const int64_t S = max_brick;
  const uint64_t brick_raw = (uint64_t)S * S * S * bpv;

  // pyramid shape: halve (ceil) until the largest dimension fits one payload
  std::vector<TocLod> lods;
  {
    TocLod L;
    for (int i = 0; i < 3; ++i) L.dims[i] = dims[i];
    while (true) {
      for (int i = 0; i < 3; ++i) L.grid[i] = (L.dims[i] + P - 1) / P;
      lods.push_back(L);
      if (std::max(L.dims[0], std::max(L.dims[1], L.dims[2])) <= P || lods.size() >= 24) break;
      for (int i = 0; i < 3; ++i) L.dims[i] = (L.dims[i] + 1) / 2;
    }
  }
  uint64_t total = 0;
  for (auto& L : lods) {
    L.first = total;
    total += L.grid[0] * L.grid[1] * L.grid[2];
  }

  TOC_HIP(hipSetDevice(device));
  hipStream_t stream = nullptr;
  TOC_HIP(hipStreamCreate(&stream));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{stream};

  // levels on the device
  std::vector<DevBuf> levels(lods.size());
  for (size_t l = 0; l < lods.size(); ++l)
    TOC_HIP(levels[l].alloc(lods[l].dims[0] * lods[l].dims[1] * lods[l].dims[2] * bpv));
  TOC_HIP(hipMemcpyAsync(levels[0].p, voxels, levels[0].bytes, hipMemcpyHostToDevice, stream));
  for (size_t l = 1; l < lods.size(); ++l) {
    const int64_t sd[3] = {(int64_t)lods[l - 1].dims[0], (int64_t)lods[l - 1].dims[1], (int64_t)lods[l - 1].dims[2]};
    TOC_HIP(launch_downsample(levels[l - 1].p, bpv, sd, levels[l].p, stream, median));
  }

  // bricks in batches of up to 256 MiB, in the layout's order
  const int batch = (int)std::max<uint64_t>(1, std::min<uint64_t>(4096, (256ull << 20) / brick_raw));
  DevBuf d_bricks, d_jobs, d_mm;
  TOC_HIP(d_bricks.alloc((size_t)batch * brick_raw));
  TOC_HIP(d_jobs.alloc((size_t)batch * sizeof(BrickJob)));
  TOC_HIP(d_mm.alloc((size_t)batch * 8));
  uint8_t* h_bricks = nullptr;
  TOC_HIP(hipHostMalloc((void**)&h_bricks, (size_t)batch * brick_raw, hipHostMallocDefault));
  std::unique_ptr<uint8_t, void (*)(uint8_t*)> h_guard(h_bricks, [](uint8_t* p) { (void)hipHostFree(p); });
  std::vector<uint32_t> h_mm((size_t)batch * 2);
  std::vector<BrickJob> jobs((size_t)batch);

  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return fail(TVK_ERR_IO, "cannot create " + path);
  std::unique_ptr<FILE, int (*)(FILE*)> f_guard(f, fclose);
  const uint64_t table_at = header_bytes((uint32_t)lods.size());
  const uint64_t data_at = table_at + total * kEntryBytes;
  if (fseeko(f, (off_t)data_at, SEEK_SET) != 0) return fail(TVK_ERR_IO, "cannot write " + path);
  std::vector<TocEntry> entries(total);
  uint64_t offset = data_at, stored_total = 0;
  const unsigned n_threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<uint8_t>> packed((size_t)batch);

  for (size_t l = 0; l < lods.size(); ++l) {
    const TocLod& L = lods[l];
    const std::vector<uint64_t> order = toc_layout_order(layout, L.grid, (uint32_t)l);
    const int64_t ld[3] = {(int64_t)L.dims[0], (int64_t)L.dims[1], (int64_t)L.dims[2]};
    for (size_t b0 = 0; b0 < order.size(); b0 += (size_t)batch) {
      const int nb = (int)std::min<size_t>((size_t)batch, order.size() - b0);
      for (int j = 0; j < nb; ++j) {
        const uint64_t i = order[b0 + j];
        jobs[j] = BrickJob{(int32_t)(i % L.grid[0]), (int32_t)((i / L.grid[0]) % L.grid[1]),
                           (int32_t)(i / (L.grid[0] * L.grid[1])), (int32_t)(L.first + i), (int64_t)j};
      }
      TOC_HIP(hipMemcpyAsync(d_jobs.p, jobs.data(), (size_t)nb * sizeof(BrickJob), hipMemcpyHostToDevice, stream));
      TOC_HIP(launch_extract_linear_bricks(levels[l].p, bpv, ld, (const BrickJob*)d_jobs.p, nb, d_bricks.p, (int)P,
                                           (int)overlap, clamp, (uint32_t*)d_mm.p, stream));
      TOC_HIP(hipMemcpyAsync(h_bricks, d_bricks.p, (size_t)nb * brick_raw, hipMemcpyDeviceToHost, stream));
      TOC_HIP(hipMemcpyAsync(h_mm.data(), d_mm.p, (size_t)nb * 8, hipMemcpyDeviceToHost, stream));
      TOC_HIP(hipStreamSynchronize(stream));
      // compress on host threads
      std::atomic<int> next{0}, err{TVK_OK};
      std::string first_error;
      std::mutex err_mu;
      auto work = [&]() {
        for (int j; (j = next.fetch_add(1)) < nb;) {
          const int e = toc_compress(compression, level, h_bricks + (size_t)j * brick_raw, brick_raw, &packed[j]);
          if (e != TVK_OK) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (err.load() == TVK_OK) first_error = tvk_last_error();
            err = e;
          }
        }
      };
      std::vector<std::thread> pool;
      for (unsigned t = 1; t < std::min<unsigned>(n_threads, (unsigned)nb); ++t) pool.emplace_back(work);
      work();
      for (auto& t : pool) t.join();
      if (err.load() != TVK_OK) return fail(err.load(), first_error);
      Writer w{f};
      for (int j = 0; j < nb; ++j) {
        TocEntry& E = entries[(size_t)jobs[j].brick_id];
        E.offset = offset;
        E.stored = packed[j].size();
        E.raw = brick_raw;
        E.compression = compression;
        E.mn = (float)h_mm[2 * j];
        E.mx = (float)h_mm[2 * j + 1];
        w.raw(packed[j].data(), packed[j].size());
        offset += packed[j].size();
        stored_total += packed[j].size();
      }
      if (!w.ok) return fail(TVK_ERR_IO, "write failed on " + path);
    }
  }

  // histogram blocks after the bricks (the renderer's histogram kernels)
  const uint64_t n1 = 1ull << bits;
  const int n2v = 256, n2g = 256;
  std::vector<uint64_t> h1(n1), h2((size_t)n2v * n2g);
  unsigned long long kmax = 0;
  {
    DevBuf d_h1, d_h2, d_km;
    TOC_HIP(d_h1.alloc(n1 * 8));
    TOC_HIP(d_h2.alloc(h2.size() * 8));
    TOC_HIP(d_km.alloc(8));
    const int64_t d0[3] = {(int64_t)dims[0], (int64_t)dims[1], (int64_t)dims[2]};
    TOC_HIP(launch_histogram_1d(levels[0].p, bpv, d0[0] * d0[1] * d0[2], (unsigned long long*)d_h1.p, stream));
    TOC_HIP(launch_gradient_kmax(levels[0].p, bpv, d0, (unsigned long long*)d_km.p, stream));
    TOC_HIP(launch_histogram_2d(levels[0].p, bpv, d0, (const unsigned long long*)d_km.p, n2v, n2g,
                                (unsigned long long*)d_h2.p, stream));
    TOC_HIP(hipMemcpyAsync(h1.data(), d_h1.p, n1 * 8, hipMemcpyDeviceToHost, stream));
    TOC_HIP(hipMemcpyAsync(h2.data(), d_h2.p, h2.size() * 8, hipMemcpyDeviceToHost, stream));
    TOC_HIP(hipMemcpyAsync(&kmax, d_km.p, 8, hipMemcpyDeviceToHost, stream));
    TOC_HIP(hipStreamSynchronize(stream));
  }
  const float max_grad = (float)(sqrt((double)kmax) / (2.0 * (double)((1u << bits) - 1)));
  const uint64_t h1_at = offset;
  const uint64_t h1_bytes = 8 + (sizeof(kHist1Name) - 1) + 24 + 8 + n1 * 8;
  {
    Writer wh{f};
    wh.u64(sizeof(kHist1Name) - 1);
    wh.raw(kHist1Name, sizeof(kHist1Name) - 1);
    wh.u64(5);
    wh.u64(0);
    wh.u64(h1_bytes);  // the 2D histogram follows
    wh.u64(n1);
    wh.raw(h1.data(), n1 * 8);
    wh.u64(sizeof(kHist2Name) - 1);
    wh.raw(kHist2Name, sizeof(kHist2Name) - 1);
    wh.u64(6);
    wh.u64(0);
    wh.u64(0);  // last block
    wh.f32(max_grad);
    wh.u64((uint64_t)n2v);
    wh.u64((uint64_t)n2g);
    wh.raw(h2.data(), h2.size() * 8);
    if (!wh.ok) return fail(TVK_ERR_IO, "write failed on " + path);
  }

  // headers and the table
  if (fseeko(f, 0, SEEK_SET) != 0) return fail(TVK_ERR_IO, "cannot write " + path);
  Writer w{f};
  w.raw("UVF-DATA", 8);
  const uint8_t little = 0;
  w.raw(&little, 1);
  w.u64(3);  // version
  w.u64(0);  // checksum: none
  w.u64(0);  // checksum bytes
  w.u64(0);  // bytes to skip before the first block
  w.u64(sizeof(kBlockName) - 1);
  w.raw(kBlockName, sizeof(kBlockName) - 1);
  w.u64(kTocSemantic);
  w.u64(0);            // block compression: none
  w.u64(h1_at - kMasterBytes);  // next block: the 1D histogram
  w.raw(kTocMagic, 8);
  w.u32(bits);
  w.u32(1);
  for (int i = 0; i < 3; ++i) w.u64(dims[i]);
  w.u32(max_brick);
  w.u32(overlap);
  w.u32(P);
  w.u32((uint32_t)lods.size());
  w.u32(median ? 1 : 0);
  w.u32(clamp ? 1 : 0);
  w.u32(compression);
  w.u32(level);
  w.u32(layout);
  w.u32(0);
  for (const auto& L : lods) {
    for (int i = 0; i < 3; ++i) w.u64(L.dims[i]);
    for (int i = 0; i < 3; ++i) w.u64(L.grid[i]);
    w.u64(L.first);
  }
  for (const auto& E : entries) {
    w.u64(E.offset);
    w.u64(E.stored);
    w.u64(E.raw);
    w.u32(E.compression);
    w.u32(0);
    w.f32(E.mn);
    w.f32(E.mx);
  }
  if (!w.ok || (uint64_t)ftello(f) != data_at) return fail(TVK_ERR_IO, "write failed on " + path);
  if (info) {
    memset(info, 0, sizeof(*info));
    for (int i = 0; i < 3; ++i) info->dims[i] = dims[i];
    info->bits = bits;
    info->max_brick = max_brick;
    info->overlap = overlap;
    info->n_lod = (uint32_t)lods.size();
    info->use_median = median ? 1 : 0;
    info->clamp_to_edge = clamp ? 1 : 0;
    info->compression = compression;
    info->level = level;
    info->layout = layout;
    info->n_bricks = total;
    info->stored_bytes = stored_total;
    info->raw_bytes = total * brick_raw;
  }
  return TVK_OK;
}

// ----------------------------------------------------------------- reader
TocVolume::~TocVolume() {
  if (fd_ >= 0) close(fd_);
}

int TocVolume::Open(const std::string& path) {
  fd_ = open(path.c_str(), O_RDONLY);
  if (fd_ < 0) return fail(TVK_ERR_IO, "cannot open " + path);
  const off_t end = lseek(fd_, 0, SEEK_END);
  if (end <= 0) return fail(TVK_ERR_IO, "cannot read " + path);
  file_bytes = (uint64_t)end;
  Reader r{fd_, 0, file_bytes};
  char magic[8];
  r.raw(magic, 8);
  uint8_t endian = 1;
  r.raw(&endian, 1);
  if (!r.ok || memcmp(magic, "UVF-DATA", 8) != 0) return fail(TVK_ERR_FORMAT, path + " is not a UVF file");
  if (endian != 0) return fail(TVK_ERR_FORMAT, "big-endian UVF files are not supported");
  const uint64_t version = r.u64();
  r.u64();  // checksum type
  const uint64_t ncs = r.u64();
  if (!r.ok || version != 3) return fail(TVK_ERR_FORMAT, "a TOC volume is a UVF version 3 file");
  if (ncs > 64) return fail(TVK_ERR_FORMAT, "bad checksum length");
  r.pos += ncs;
  const uint64_t skip = r.u64();
  r.pos += skip;
  // walk the blocks: the TOC block and the histogram blocks
  uint64_t toc_at = 0, h1_at = 0, h2_at = 0;
  for (int guard = 0; guard < 4096; ++guard) {
    const uint64_t block_at = r.pos;
    const uint64_t name_len = r.u64();
    if (!r.ok || name_len > 4096) return fail(TVK_ERR_FORMAT, "bad UVF block header");
    r.pos += name_len;
    const uint64_t semantic = r.u64();
    const uint64_t block_comp = r.u64();
    const uint64_t next = r.u64();
    if (!r.ok) return fail(TVK_ERR_FORMAT, "truncated UVF block header");
    if (block_comp == 0) {
      if (semantic == kTocSemantic && !toc_at) toc_at = r.pos;
      if (semantic == 5 && !h1_at) h1_at = r.pos;
      if (semantic == 6 && !h2_at) h2_at = r.pos;
    }
    if (next == 0) break;
    if (next > file_bytes - block_at) return fail(TVK_ERR_FORMAT, "bad UVF block offset");
    r.pos = block_at + next;
  }
  if (!toc_at) return fail(TVK_ERR_FORMAT, "no TOC block in " + path);
  if (h1_at) {
    Reader h{fd_, h1_at, file_bytes};
    const uint64_t n = h.u64();
    if (!h.ok || n == 0 || n > (1u << 16) || n * 8 > file_bytes - h.pos) return fail(TVK_ERR_FORMAT, "bad 1D histogram block");
    hist1d.resize(n);
    h.raw(hist1d.data(), n * 8);
    if (!h.ok) return fail(TVK_ERR_FORMAT, "truncated 1D histogram block");
  }
  if (h2_at) {
    Reader h{fd_, h2_at, file_bytes};
    max_grad = h.f32();
    hist2d_x = h.u64();
    hist2d_y = h.u64();
    if (!h.ok || hist2d_x == 0 || hist2d_y == 0 || hist2d_x > 65536 || hist2d_y > 65536 ||
        hist2d_x * hist2d_y * 8 > file_bytes - h.pos)
      return fail(TVK_ERR_FORMAT, "bad 2D histogram block");
    hist2d.resize(hist2d_x * hist2d_y);
    h.raw(hist2d.data(), hist2d.size() * 8);
    if (!h.ok) return fail(TVK_ERR_FORMAT, "truncated 2D histogram block");
  }
  r.pos = toc_at;
  char tmagic[8];
  r.raw(tmagic, 8);
  if (!r.ok || memcmp(tmagic, kTocMagic, 8) != 0) return fail(TVK_ERR_FORMAT, "unknown TOC block payload");
  bits = r.u32();
  const uint32_t comps = r.u32();
  for (int i = 0; i < 3; ++i) dims[i] = r.u64();
  max_brick = r.u32();
  overlap = r.u32();
  payload = r.u32();
  n_lod = r.u32();
  median = r.u32();
  clamp = r.u32();
  compression = r.u32();
  level = r.u32();
  layout = r.u32();
  r.u32();
  if (!r.ok || (bits != 8 && bits != 16) || comps != 1 || n_lod < 1 || n_lod > 24 || max_brick < 2 ||
      max_brick > 1024 || 2 * (uint64_t)overlap >= max_brick || payload != max_brick - 2 * overlap ||
      compression >= kCodecs)
    return fail(TVK_ERR_FORMAT, "bad TOC header");
  for (int i = 0; i < 3; ++i)
    if (dims[i] == 0 || dims[i] > (1ull << 32)) return fail(TVK_ERR_FORMAT, "bad TOC volume size");
  lods.resize(n_lod);
  uint64_t total = 0;
  for (size_t l = 0; l < lods.size(); ++l) {
    TocLod& L = lods[l];
    for (int i = 0; i < 3; ++i) L.dims[i] = r.u64();
    for (int i = 0; i < 3; ++i) L.grid[i] = r.u64();
    L.first = r.u64();
    // level 0 is the volume; each further level halves it, rounding up
    for (int i = 0; i < 3; ++i) {
      const uint64_t want = l == 0 ? dims[i] : (lods[l - 1].dims[i] + 1) / 2;
      if (L.dims[i] != want || L.grid[i] != (L.dims[i] + payload - 1) / payload)
        return fail(TVK_ERR_FORMAT, "bad TOC level");
    }
    if (L.first != total) return fail(TVK_ERR_FORMAT, "bad TOC level table");
    total += L.grid[0] * L.grid[1] * L.grid[2];
  }
  if (!r.ok || total * kEntryBytes > file_bytes) return fail(TVK_ERR_FORMAT, "truncated TOC table");
  const uint64_t raw = (uint64_t)max_brick * max_brick * max_brick * (bits / 8);
  entries.resize(total);
  for (auto& E : entries) {
    E.offset = r.u64();
    E.stored = r.u64();
    E.raw = r.u64();
    E.compression = r.u32();
    E.reserved = r.u32();
    E.mn = r.f32();
    E.mx = r.f32();
    if (r.ok && (E.raw != raw || E.compression >= kCodecs || E.offset > file_bytes ||
                 E.stored > file_bytes - E.offset))
      return fail(TVK_ERR_FORMAT, "bad TOC entry");
  }
  if (!r.ok) return fail(TVK_ERR_FORMAT, "truncated TOC table");
  return TVK_OK;
}

int TocVolume::ReadBrick(uint32_t lod, uint32_t bx, uint32_t by, uint32_t bz, void* out, uint64_t bytes) {
  if (lod >= n_lod) return fail(TVK_ERR_ARG, "LOD index out of range");
  const TocLod& L = lods[lod];
  if (bx >= L.grid[0] || by >= L.grid[1] || bz >= L.grid[2]) return fail(TVK_ERR_ARG, "brick index out of range");
  const TocEntry& E = entries[L.first + ((uint64_t)bz * L.grid[1] + by) * L.grid[0] + bx];
  if (!out || bytes < E.raw) return fail(TVK_ERR_ARG, "output buffer too small for a brick");
  if (E.compression == 0) {
    if (E.stored != E.raw) return fail(TVK_ERR_FORMAT, "stored brick size mismatch");
    if (pread(fd_, out, E.stored, (off_t)E.offset) != (ssize_t)E.stored) return fail(TVK_ERR_IO, "brick read failed");
    return TVK_OK;
  }
  std::vector<uint8_t> buf(E.stored);
  if (pread(fd_, buf.data(), E.stored, (off_t)E.offset) != (ssize_t)E.stored) return fail(TVK_ERR_IO, "brick read failed");
  return toc_decompress(E.compression, buf.data(), E.stored, out, E.raw);
}

void TocVolume::Info(tvk_toc_info* out) const {
  memset(out, 0, sizeof(*out));
  for (int i = 0; i < 3; ++i) out->dims[i] = dims[i];
  out->bits = bits;
  out->max_brick = max_brick;
  out->overlap = overlap;
  out->n_lod = n_lod;
  out->use_median = median;
  out->clamp_to_edge = clamp;
  out->compression = compression;
  out->level = level;
  out->layout = layout;
  out->n_bricks = entries.size();
  for (const auto& E : entries) {
    out->stored_bytes += E.stored;
    out->raw_bytes += E.raw;
  }
}

}  // namespace tvk
