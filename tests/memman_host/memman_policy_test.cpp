// Policies of the device-wide memory manager (imagevis3d_amd/csrc/memman.cpp)
// on the host, with hip_stub.cpp standing in for the device: budgets,
// cross-volume LRU eviction, this frame's slots, slots being filled, chunk
// reclaim across slot sizes, re-registration after the last volume of a
// size went away, the host brick cache and the worker pool.  Built with
// AddressSanitizer by tests/test_memman_host.py.  TEST INFRASTRUCTURE ONLY.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <vector>

#include "../../imagevis3d_amd/csrc/memman.hpp"

using tvk::MemMan;

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      exit(1);                                                           \
    }                                                                    \
  } while (0)

struct Client : tvk::PoolClient {
  std::vector<int> evicted;
  void PoolEvict(int gid) override { evicted.push_back(gid); }
};

static const uint64_t MiB = 1ull << 20;

static size_t acquire(MemMan& mm, int id, int first, size_t n, std::vector<int64_t>* slots = nullptr,
                      bool commit = true) {
  std::vector<int> gids(n);
  for (size_t i = 0; i < n; ++i) gids[i] = first + (int)i;
  std::vector<int64_t> s(n);
  std::vector<void*> a(n);
  const size_t got = mm.Acquire(id, gids.data(), n, s.data(), a.data());
  for (size_t i = 0; i < got; ++i) CHECK(a[i] != nullptr);
  if (commit) mm.Commit(id, s.data(), got);
  if (slots) *slots = std::vector<int64_t>(s.begin(), s.begin() + got);
  return got;
}

static uint64_t used(MemMan& mm) {
  tvk::MemManStats st;
  mm.Stats(&st);
  return st.gpu_used;
}

int main() {
  MemMan& mm = MemMan::ForDevice(0);
  const uint64_t base = used(mm);

  {  // two volumes share 100 slots; the second pages the first's oldest out
    mm.SetGPUBudget(base + 100 * MiB);
    Client a, b;
    const int ia = mm.Register(&a, MiB, 64), ib = mm.Register(&b, MiB, 64);
    std::vector<int64_t> sa;
    CHECK(acquire(mm, ia, 0, 64, &sa) == 64);
    mm.BeginFrame(ib);
    CHECK(acquire(mm, ib, 0, 64) == 64);
    CHECK(a.evicted.size() == 28 && b.evicted.empty());
    for (int g : a.evicted) CHECK(g >= 0 && g < 64);
    CHECK(used(mm) <= base + 100 * MiB);
    // a's next frame samples its remaining 36 bricks: they are protected
    // from a's own request, b's are not -- a takes b's least recent ones
    mm.BeginFrame(ia);
    mm.Touch(ia, sa.data(), sa.size());
    CHECK(acquire(mm, ia, 100, 20) == 20);
    CHECK(b.evicted.size() == 20 && a.evicted.size() == 28);
    // least recent first across both volumes: b's bricks (stamped before
    // a's last frame) go before a's, whose frame this no longer is
    mm.BeginFrame(ia);
    CHECK(acquire(mm, ia, 200, 8) == 8 && b.evicted.size() == 28 && a.evicted.size() == 28);
    mm.Unregister(ia);
    mm.Unregister(ib);
    CHECK(used(mm) == base);  // the last volumes of this size returned the chunks
    printf("OK share\n");
  }
  {  // a volume's own share: this frame's bricks are never its victims
    mm.SetGPUBudget(base + 100 * MiB);
    Client a;
    const int ia = mm.Register(&a, MiB, 10);
    CHECK(acquire(mm, ia, 0, 10) == 10);
    CHECK(acquire(mm, ia, 10, 5) == 0);  // all 10 sampled in this frame
    mm.BeginFrame(ia);
    CHECK(acquire(mm, ia, 10, 5) == 5 && a.evicted.size() == 5);
    CHECK(mm.ClientSlots(ia) == 10);
    mm.Unregister(ia);
    printf("OK own share\n");
  }
  {  // slots being filled are taken by no one; committed ones are
    mm.SetGPUBudget(base + 100 * MiB);
    Client a, b;
    const int ia = mm.Register(&a, MiB, 0), ib = mm.Register(&b, MiB, 0);
    std::vector<int64_t> sa;
    CHECK(acquire(mm, ia, 0, 80, &sa, false) == 80);
    mm.BeginFrame(ib);
    CHECK(acquire(mm, ib, 0, 50) == 20 && a.evicted.empty());
    mm.Commit(ia, sa.data(), sa.size());
    CHECK(acquire(mm, ib, 20, 30) == 30 && a.evicted.size() == 30);
    mm.Unregister(ia);
    mm.Unregister(ib);
    printf("OK filling\n");
  }
  {  // other device memory: the pool gives way, chunk by chunk
    mm.SetGPUBudget(base + 100 * MiB);
    Client a;
    const int ia = mm.Register(&a, MiB, 0);
    CHECK(acquire(mm, ia, 0, 100) == 100);
    CHECK(used(mm) == base + 100 * MiB);
    CHECK(mm.ReserveGPU(10 * MiB));  // frees the (only) chunk
    CHECK(a.evicted.size() == 100 && used(mm) == base + 10 * MiB);
    CHECK(acquire(mm, ia, 200, 100) == 90);  // the rest of the budget
    CHECK(!mm.ReserveGPU(200 * MiB));         // more than the budget: refused
    mm.ReleaseGPU(10 * MiB);
    mm.Unregister(ia);
    CHECK(used(mm) == base);
    printf("OK reserve\n");
  }
  {  // another slot size in need takes a whole chunk of the first size
    mm.SetGPUBudget(base + 64 * MiB);
    Client a, c;
    const int ia = mm.Register(&a, 2 * MiB, 0), ic = mm.Register(&c, 512 * 1024, 0);
    CHECK(acquire(mm, ia, 0, 32) == 32);
    mm.BeginFrame(ic);
    CHECK(acquire(mm, ic, 0, 40) == 40);
    CHECK(a.evicted.size() == 32);
    tvk::MemManStats st;
    mm.Stats(&st);
    CHECK(st.chunks_freed >= 1 && st.gpu_used <= st.gpu_budget);
    mm.Unregister(ia);
    mm.Unregister(ic);
    printf("OK cross size\n");
  }
  {  // the last volume of a size goes, a new one of that size comes
    mm.SetGPUBudget(base + 100 * MiB);
    for (int round = 0; round < 3; ++round) {
      Client a;
      const int ia = mm.Register(&a, MiB, 0);
      CHECK(acquire(mm, ia, 0, 50) == 50);
      CHECK(acquire(mm, ia, 50, 50) == 50);
      mm.Unregister(ia);
      CHECK(used(mm) == base);
    }
    printf("OK re-register\n");
  }
  {  // host brick cache within the CPU budget
    Client a;
    const int ia = mm.Register(&a, MiB, 0);
    std::vector<uint8_t> brick(4096, 7), out(4096);
    tvk::MemManStats st;
    mm.Stats(&st);
    mm.SetCPUBudget(st.cpu_used + 3 * 4096);
    for (int g = 0; g < 5; ++g) {
      brick[0] = (uint8_t)g;
      mm.CachePut(ia, g, brick.data(), brick.size());
    }
    mm.Stats(&st);
    CHECK(st.cache_bytes == 3 * 4096);  // the two oldest gave way
    CHECK(!mm.CacheGet(ia, 0, out.data(), out.size()));
    CHECK(mm.CacheGet(ia, 4, out.data(), out.size()) && out[0] == 4 && out[1] == 7);
    CHECK(mm.ReserveCPU(2 * 4096));  // other host memory pushes the cache out
    mm.Stats(&st);
    CHECK(st.cache_bytes <= 4096);
    mm.ReleaseCPU(2 * 4096);
    mm.Unregister(ia);
    mm.Stats(&st);
    CHECK(st.cache_bytes == 0);
    mm.SetCPUBudget(1ull << 40);
    printf("OK cache\n");
  }
  {  // worker pool: every index exactly once, generation after generation
    for (int rep = 0; rep < 200; ++rep) {
      const size_t n = 1 + (size_t)(rep * 37) % 300;
      std::vector<std::atomic<int>> hits(n);
      for (auto& h : hits) h = 0;
      mm.ParallelFor(n, 1 + rep % 9, [&](size_t i) { hits[i].fetch_add(1); });
      for (auto& h : hits) CHECK(h.load() == 1);
    }
    printf("OK parallel for\n");
  }
  printf("ALL OK\n");
  return 0;
}
