// uvf.cpp -- UVF v2 reader and MD5 (see uvf.hpp).
#include "uvf.hpp"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../include/tvk_hip.h"

namespace tvk {
namespace {

// ------------------------------------------------------------------- MD5
// RFC 1321: 64-step compression over 512-bit blocks; the additive constants
// are floor(2^32 * |sin(i + 1)|), the per-round rotations and message
// schedules are those of the RFC.
struct Md5 {
  uint32_t h[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
  uint64_t len = 0;
  uint8_t buf[64];
  uint32_t nbuf = 0;
  uint32_t K[64];
  Md5() {
    for (int i = 0; i < 64; ++i) K[i] = (uint32_t)(uint64_t)floor(fabs(sin((double)(i + 1))) * 4294967296.0);
  }
  static uint32_t rotl(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }
  void block(const uint8_t* p) {
    static const int R[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    uint32_t m[16];
    for (int i = 0; i < 16; ++i)
      m[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
             ((uint32_t)p[4 * i + 3] << 24);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
    for (int i = 0; i < 64; ++i) {
      uint32_t f;
      int g;
      const int round = i >> 4;
      if (round == 0) { f = (b & c) | (~b & d); g = i; }
      else if (round == 1) { f = (b & d) | (c & ~d); g = (5 * i + 1) & 15; }
      else if (round == 2) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
      else { f = c ^ (b | ~d); g = (7 * i) & 15; }
      const uint32_t t = d;
      d = c;
      c = b;
      b = b + rotl(a + f + K[i] + m[g], R[round][i & 3]);
      a = t;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d;
  }
  void update(const uint8_t* p, uint64_t n) {
    len += n;
    while (n > 0) {
      if (nbuf == 0 && n >= 64) {
        block(p);
        p += 64;
        n -= 64;
        continue;
      }
      const uint32_t take = (uint32_t)(n < 64 - nbuf ? n : 64 - nbuf);
      memcpy(buf + nbuf, p, take);
      nbuf += take;
      p += take;
      n -= take;
      if (nbuf == 64) {
        block(buf);
        nbuf = 0;
      }
    }
  }
  void final(uint8_t out[16]) {
    const uint64_t bits = len * 8;
    const uint8_t pad0 = 0x80, zero = 0;
    update(&pad0, 1);
    while (nbuf != 56) update(&zero, 1);
    uint8_t l[8];
    for (int i = 0; i < 8; ++i) l[i] = (uint8_t)(bits >> (8 * i));
    update(l, 8);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h[i] >> (8 * j));
  }
};

struct Reader {
  FILE* f = nullptr;
  uint64_t size = 0;
  ~Reader() {
    if (f) fclose(f);
  }
  bool open(const std::string& p) {
    f = fopen(p.c_str(), "rb");
    if (!f) return false;
    fseeko(f, 0, SEEK_END);
    size = (uint64_t)ftello(f);
    fseeko(f, 0, SEEK_SET);
    return true;
  }
  bool seek(uint64_t o) { return o <= size && fseeko(f, (off_t)o, SEEK_SET) == 0; }
  bool bytes(void* p, uint64_t n) { return fread(p, 1, n, f) == n; }
  bool u64(uint64_t* v) { return bytes(v, 8); }
  uint64_t tell() { return (uint64_t)ftello(f); }
};

int fail(std::string* err, int code, const std::string& msg) {
  if (err) *err = msg;
  return code;
}

}  // namespace

void md5(const uint8_t* data, uint64_t n, uint8_t digest[16]) {
  Md5 m;
  m.update(data, n);
  m.final(digest);
}

int uvf_read(const std::string& path, bool verify, UvfFile* out, std::string* err) {
  Reader r;
  if (!r.open(path)) return fail(err, TVK_ERR_IO, "cannot open " + path);
  char magic[8];
  uint8_t endian = 0;
  if (!r.bytes(magic, 8) || memcmp(magic, "UVF-DATA", 8) != 0)
    return fail(err, TVK_ERR_FORMAT, "not a UVF file (magic)");
  if (!r.bytes(&endian, 1)) return fail(err, TVK_ERR_FORMAT, "truncated header");
  if (endian != 0) return fail(err, TVK_ERR_FORMAT, "big-endian UVF is not supported");
  uint64_t ncs = 0;
  if (!r.u64(&out->version) || !r.u64(&out->checksum_type) || !r.u64(&ncs) || ncs > 64)
    return fail(err, TVK_ERR_FORMAT, "truncated header");
  if (out->version != 2) return fail(err, TVK_ERR_FORMAT, "only UVF version 2 is supported");
  out->checksum.resize(ncs);
  if (ncs && !r.bytes(out->checksum.data(), ncs)) return fail(err, TVK_ERR_FORMAT, "truncated checksum");
  const uint64_t checked_from = r.tell();
  uint64_t seek = 0;
  if (!r.u64(&seek)) return fail(err, TVK_ERR_FORMAT, "truncated header");
  const uint64_t first_block = r.tell() + seek;

  if (verify && out->checksum_type == 2) {
    if (!r.seek(checked_from)) return fail(err, TVK_ERR_FORMAT, "bad checksum offset");
    Md5 m;
    std::vector<uint8_t> chunk(1 << 20);
    uint64_t left = r.size - checked_from;
    while (left) {
      const uint64_t n = left < chunk.size() ? left : chunk.size();
      if (!r.bytes(chunk.data(), n)) return fail(err, TVK_ERR_IO, "read error");
      m.update(chunk.data(), n);
      left -= n;
    }
    uint8_t dg[16];
    m.final(dg);
    out->checksum_ok = ncs == 16 && memcmp(dg, out->checksum.data(), 16) == 0;
    if (!out->checksum_ok) return fail(err, TVK_ERR_CHECKSUM, "UVF MD5 checksum mismatch");
  }

  uint64_t o = first_block;
  for (int guard = 0; guard < 4096; ++guard) {
    if (!r.seek(o)) return fail(err, TVK_ERR_FORMAT, "block offset beyond end of file");
    const uint64_t start = o;
    uint64_t n = 0, sem = 0, comp = 0, next = 0;
    if (!r.u64(&n) || n > 65536) return fail(err, TVK_ERR_FORMAT, "bad block header");
    std::string desc(n, '\0');
    if (n && !r.bytes(&desc[0], n)) return fail(err, TVK_ERR_FORMAT, "bad block header");
    if (!r.u64(&sem) || !r.u64(&comp) || !r.u64(&next)) return fail(err, TVK_ERR_FORMAT, "bad block header");
    if (comp != 0) return fail(err, TVK_ERR_FORMAT, "compressed UVF blocks are not supported");
    if (sem == 1) {
      UvfRaster ras;
      uint64_t nd = 0;
      if (!r.u64(&nd) || nd != 3) return fail(err, TVK_ERR_FORMAT, "only 3D raster blocks are supported");
      uint64_t tmp;
      for (uint64_t i = 0; i < nd; ++i) r.u64(&tmp);  // domain semantics
      if (!r.bytes(ras.transform, 16 * sizeof(double))) return fail(err, TVK_ERR_FORMAT, "truncated raster header");
      uint64_t lodf[3], lodg[3];
      bool ok = true;
      for (int i = 0; i < 3; ++i) ok &= r.u64(&ras.dims[i]);
      for (int i = 0; i < 3; ++i) ok &= r.u64(&ras.brick_size[i]);
      for (int i = 0; i < 3; ++i) ok &= r.u64(&ras.overlap[i]);
      for (int i = 0; i < 3; ++i) ok &= r.u64(&lodf[i]);
      for (int i = 0; i < 3; ++i) ok &= r.u64(&lodg[i]);
      uint64_t ngroups = 1;
      for (int i = 0; i < 3; ++i) ngroups = lodg[i] + 1 > ngroups ? lodg[i] + 1 : ngroups;
      if (!ok || ngroups > 3) return fail(err, TVK_ERR_FORMAT, "truncated raster header");
      for (uint64_t g = 0; g < ngroups; ++g) {
        uint64_t c = 0;
        ok &= r.u64(&c);
        if (g == 0) ras.lod_levels = c;
      }
      uint64_t nel = 0;
      ok &= r.u64(&nel);
      if (!ok || nel != 1) return fail(err, TVK_ERR_FORMAT, "only scalar raster blocks are supported");
      uint64_t es = 0;
      ok &= r.u64(&es);
      if (!ok || es != 1) return fail(err, TVK_ERR_FORMAT, "only scalar raster blocks are supported");
      uint64_t esem = 0, ebits = 0, emant = 0;
      ok &= r.u64(&esem) && r.u64(&ebits) && r.u64(&emant);
      uint8_t sgn = 0;
      ok &= r.bytes(&sgn, 1);
      uint64_t doff = 0;
      ok &= r.u64(&doff);
      if (!ok) return fail(err, TVK_ERR_FORMAT, "truncated raster header");
      ras.bits = (uint32_t)ebits;
      ras.is_signed = sgn != 0;
      ras.data_offset = r.tell() + doff;
      out->rasters.push_back(ras);
    } else if (sem == 5) {
      uint64_t hn = 0;
      if (!r.u64(&hn) || hn > (1ull << 32)) return fail(err, TVK_ERR_FORMAT, "bad 1D histogram");
      out->histogram_1d.resize(hn);
      if (!r.bytes(out->histogram_1d.data(), hn * 8)) return fail(err, TVK_ERR_FORMAT, "bad 1D histogram");
    } else if (sem == 6) {
      uint64_t ex = 0, ey = 0;
      if (!r.bytes(&out->max_gradient, 4) || !r.u64(&ex) || !r.u64(&ey) || ex * ey > (1ull << 28))
        return fail(err, TVK_ERR_FORMAT, "bad 2D histogram");
      out->histogram_2d_dims[0] = ex;
      out->histogram_2d_dims[1] = ey;
      out->histogram_2d.resize(ex * ey);
      if (!r.bytes(out->histogram_2d.data(), ex * ey * 8)) return fail(err, TVK_ERR_FORMAT, "bad 2D histogram");
    }
    if (next == 0) break;
    o = start + next;
  }
  if (out->rasters.empty()) return fail(err, TVK_ERR_FORMAT, "UVF file holds no raster block");
  return 0;
}

int uvf_read_voxels(const std::string& path, const UvfRaster& ras, std::vector<uint8_t>* out,
                    std::string* err) {
  if (ras.bits != 8 && ras.bits != 16) return fail(err, TVK_ERR_FORMAT, "only 8/16-bit voxels are supported");
  if (ras.is_signed) return fail(err, TVK_ERR_FORMAT, "signed voxels are not supported");
  for (int i = 0; i < 3; ++i)
    if (ras.dims[i] == 0 || ras.dims[i] > ras.brick_size[i])
      return fail(err, TVK_ERR_FORMAT, "multi-brick raster blocks are not supported by this reader");
  if (ras.lod_levels != 1) return fail(err, TVK_ERR_FORMAT, "multi-LOD raster blocks are not supported by this reader");
  Reader r;
  if (!r.open(path)) return fail(err, TVK_ERR_IO, "cannot open " + path);
  const uint64_t n = ras.dims[0] * ras.dims[1] * ras.dims[2] * (ras.bits / 8);
  if (ras.data_offset + n > r.size) return fail(err, TVK_ERR_FORMAT, "raster payload truncated");
  out->resize(n);
  if (!r.seek(ras.data_offset) || !r.bytes(out->data(), n)) return fail(err, TVK_ERR_IO, "read error");
  return 0;
}

}  // namespace tvk
