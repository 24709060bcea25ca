// memman.cpp -- the device-wide memory manager (see memman.hpp).
#include "memman.hpp"

#include <string.h>
#include <unistd.h>

#include <algorithm>

namespace tvk {

namespace {

constexpr uint64_t kChunkTarget = 512ull << 20;   // largest paged-pool chunk
constexpr uint64_t kMinChunkBytes = 16ull << 20;  // smallest one the budget allows
constexpr int kMaxWorkers = 31;
constexpr int kGenShift = 40;  // ParallelFor ticket: generation << 40 | index

// Runs with `device` current and restores the caller's device.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int device) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    (void)hipSetDevice(device);
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace

MemMan& MemMan::ForDevice(int device) {
  static std::mutex m;
  static std::map<int, MemMan*> all;  // one per device for the life of the process
  std::lock_guard<std::mutex> g(m);
  auto it = all.find(device);
  if (it == all.end()) it = all.emplace(device, new MemMan(device)).first;
  return *it->second;
}

MemMan::MemMan(int device) : device_(device) {
  size_t free_b = 0, total_b = 0;
  {
    DeviceScope ds(device);
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) total_b = 0;
  }
  gpu_total_ = total_b;
  // default budget: 90% of the device (the rest for other users of the GPU)
  gpu_budget_ = total_b ? total_b / 10 * 9 : (1ull << 62);
  const long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGE_SIZE);
  cpu_total_ = pages > 0 && psz > 0 ? (uint64_t)pages * (uint64_t)psz : 0;
  cpu_budget_ = cpu_total_ ? cpu_total_ / 2 : (1ull << 62);
}

void MemMan::SetGPUBudget(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  gpu_budget_ = bytes;
  // free chunks until the pool is back within the budget
  while (gpu_used_ > gpu_budget_ && reclaim_chunk(-1, gpu_used_ - gpu_budget_)) {
  }
}

void MemMan::SetCPUBudget(uint64_t bytes) {
  {
    std::lock_guard<std::mutex> g(mu_);
    cpu_budget_ = bytes;
  }
  std::lock_guard<std::mutex> g(cache_mu_);
  cache_trim_locked(0);
}

uint64_t MemMan::GPUBudget() {
  std::lock_guard<std::mutex> g(mu_);
  return gpu_budget_;
}

void MemMan::Stats(MemManStats* out) {
  memset(out, 0, sizeof(*out));
  {
    std::lock_guard<std::mutex> g(mu_);
    out->gpu_total = gpu_total_;
    out->gpu_budget = gpu_budget_;
    out->gpu_used = gpu_used_;
    out->cpu_total = cpu_total_;
    out->cpu_budget = cpu_budget_;
    out->cpu_used = cpu_used_;
    for (const SlotClass& C : classes_) {
      for (size_t c = 0; c < C.chunks.size(); ++c)
        if (C.chunks[c]) {
          out->pool_bytes += (uint64_t)C.chunk_slots[c] * C.slot_bytes;
          out->pool_slots += C.chunk_slots[c];
        }
      for (const Slot& s : C.slots) out->pool_resident += (s.live && s.client >= 0) ? 1 : 0;
    }
    out->pool_evictions = evictions_;
    out->chunks_freed = chunks_freed_;
  }
  std::lock_guard<std::mutex> g(cache_mu_);
  out->cache_bytes = cache_bytes_;
  out->cpu_used += cache_bytes_;
  out->cache_hits = cache_hits_;
  out->cache_misses = cache_misses_;
}

bool MemMan::ReserveGPU(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  while (gpu_used_ + bytes > gpu_budget_)
    if (!reclaim_chunk(-1, gpu_used_ + bytes - gpu_budget_)) return false;
  gpu_used_ += bytes;
  return true;
}

void MemMan::ReleaseGPU(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  gpu_used_ -= std::min(gpu_used_, bytes);
}

bool MemMan::ReserveCPU(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  {
    std::lock_guard<std::mutex> c(cache_mu_);
    if (cpu_used_ + bytes > cpu_budget_) return false;
    // the cache gives way to other host memory
    const uint64_t room = cpu_budget_ - cpu_used_ - bytes;
    while (cache_bytes_ > room && !cache_lru_.empty()) {
      auto it = cache_.find(cache_lru_.back());
      cache_bytes_ -= it->second.data.size();
      cache_.erase(it);
      cache_lru_.pop_back();
    }
  }
  cpu_used_ += bytes;
  return true;
}

void MemMan::ReleaseCPU(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  cpu_used_ -= std::min(cpu_used_.load(), bytes);
}

// ------------------------------------------------------------- paged pool
int MemMan::class_for(uint64_t slot_bytes) {
  for (size_t k = 0; k < classes_.size(); ++k)
    if (classes_[k].slot_bytes == slot_bytes) return (int)k;
  SlotClass C;
  C.slot_bytes = slot_bytes;
  C.per_chunk = (uint32_t)std::max<uint64_t>(1, kChunkTarget / slot_bytes);
  classes_.push_back(std::move(C));
  return (int)classes_.size() - 1;
}

void* MemMan::slot_addr(const SlotClass& C, int64_t s) const {
  const uint64_t c = (uint64_t)s / C.per_chunk, i = (uint64_t)s % C.per_chunk;
  return (uint8_t*)C.chunks[c] + i * C.slot_bytes;
}

bool MemMan::protected_for(const Slot& sl, int client) const {
  return sl.client >= 0 && sl.client == client && sl.stamp >= clients_[client].frame_start;
}

bool MemMan::grow(int cls, uint64_t want) {
  SlotClass& C = classes_[cls];
  const uint64_t room = gpu_budget_ > gpu_used_ ? gpu_budget_ - gpu_used_ : 0;
  const uint64_t min_slots = std::max<uint64_t>(1, kMinChunkBytes / C.slot_bytes);
  const uint64_t n = std::min<uint64_t>({(uint64_t)C.per_chunk, room / C.slot_bytes, std::max(want, min_slots)});
  if (n == 0) return false;
  void* p = nullptr;
  {
    DeviceScope ds(device_);
    if (hipMalloc(&p, n * C.slot_bytes) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
  }
  gpu_used_ += n * C.slot_bytes;
  size_t c = 0;
  while (c < C.chunks.size() && C.chunks[c]) ++c;
  if (c == C.chunks.size()) {
    C.chunks.push_back(nullptr);
    C.chunk_slots.push_back(0);
  }
  if (C.slots.size() < (c + 1) * (size_t)C.per_chunk) C.slots.resize((c + 1) * (size_t)C.per_chunk);
  C.chunks[c] = p;
  C.chunk_slots[c] = (uint32_t)n;
  for (uint64_t i = 0; i < C.per_chunk; ++i) {
    Slot& sl = C.slots[c * C.per_chunk + i];
    sl = Slot{};
    sl.live = i < n;
  }
  for (uint64_t i = n; i-- > 0;) C.free_list.push_back((int64_t)(c * C.per_chunk + i));
  return true;
}

void MemMan::evict_slot(SlotClass& C, int64_t s) {
  Slot& sl = C.slots[s];
  if (sl.client >= 0) {
    Client& K = clients_[sl.client];
    if (K.alive && K.c) K.c->PoolEvict(sl.gid);
    K.held -= std::min<uint64_t>(K.held, 1);
    ++evictions_;
  }
  sl.client = -1;
  sl.gid = -1;
  sl.stamp = 0;
  C.free_list.push_back(s);
}

// Empty and free the least recently used chunk of another slot size (free
// chunks first) that holds no slot being filled.
bool MemMan::reclaim_chunk(int exclude_cls, uint64_t need_bytes) {
  (void)need_bytes;
  int best_k = -1;
  size_t best_c = 0;
  uint64_t best_stamp = UINT64_MAX;
  for (size_t k = 0; k < classes_.size(); ++k) {
    if ((int)k == exclude_cls) continue;
    SlotClass& C = classes_[k];
    for (size_t c = 0; c < C.chunks.size(); ++c) {
      if (!C.chunks[c]) continue;
      uint64_t newest = 0;
      bool busy = false;
      for (uint32_t i = 0; i < C.chunk_slots[c] && !busy; ++i) {
        const Slot& sl = C.slots[c * C.per_chunk + i];
        if (sl.client < 0) continue;
        busy = sl.filling;  // another size's slots: their owners page them in again
        newest = std::max(newest, sl.stamp);
      }
      if (!busy && newest < best_stamp) {
        best_stamp = newest;
        best_k = (int)k;
        best_c = c;
      }
    }
  }
  if (best_k < 0) return false;
  SlotClass& C = classes_[best_k];
  const uint64_t base = best_c * C.per_chunk;
  for (uint32_t i = 0; i < C.chunk_slots[best_c]; ++i)
    if (C.slots[base + i].client >= 0) evict_slot(C, (int64_t)(base + i));
  C.free_list.erase(std::remove_if(C.free_list.begin(), C.free_list.end(),
                                   [&](int64_t s) { return (uint64_t)s / C.per_chunk == best_c; }),
                    C.free_list.end());
  {
    DeviceScope ds(device_);
    (void)hipFree(C.chunks[best_c]);
  }
  gpu_used_ -= std::min<uint64_t>(gpu_used_, (uint64_t)C.chunk_slots[best_c] * C.slot_bytes);
  for (uint32_t i = 0; i < C.per_chunk; ++i) C.slots[base + i] = Slot{-1, -1, 0, false, false};
  C.chunks[best_c] = nullptr;
  C.chunk_slots[best_c] = 0;
  ++chunks_freed_;
  return true;
}

int MemMan::Register(PoolClient* c, uint64_t slot_bytes, uint64_t max_slots) {
  std::lock_guard<std::mutex> g(mu_);
  Client K;
  K.c = c;
  K.cls = class_for(slot_bytes);
  K.max_slots = max_slots;
  K.frame_start = ++clock_;
  K.alive = true;
  for (size_t i = 0; i < clients_.size(); ++i)
    if (!clients_[i].alive && !clients_[i].c) {
      clients_[i] = K;
      return (int)i;
    }
  clients_.push_back(K);
  return (int)clients_.size() - 1;
}

void MemMan::Unregister(int client) {
  {
    std::lock_guard<std::mutex> g(mu_);
    if (client < 0 || client >= (int)clients_.size()) return;
    Client& K = clients_[client];
    SlotClass& C = classes_[K.cls];
    for (size_t s = 0; s < C.slots.size(); ++s)
      if (C.slots[s].live && C.slots[s].client == client) {
        C.slots[s] = Slot{};
        C.free_list.push_back((int64_t)s);
      }
    const int cls = K.cls;
    K = Client{};
    // the last volume of this slot size is gone: give its chunks back
    bool used = false;
    for (const Client& o : clients_) used |= o.alive && o.cls == cls;
    if (!used) {
      DeviceScope ds(device_);
      for (size_t c = 0; c < C.chunks.size(); ++c)
        if (C.chunks[c]) {
          (void)hipFree(C.chunks[c]);
          gpu_used_ -= std::min<uint64_t>(gpu_used_, (uint64_t)C.chunk_slots[c] * C.slot_bytes);
          C.chunks[c] = nullptr;
          C.chunk_slots[c] = 0;
        }
      C.chunks.clear();
      C.chunk_slots.clear();
      C.slots.clear();
      C.free_list.clear();
    }
  }
  std::lock_guard<std::mutex> g(cache_mu_);
  for (auto it = cache_lru_.begin(); it != cache_lru_.end();) {
    if ((int)(*it >> 32) == client) {
      auto e = cache_.find(*it);
      cache_bytes_ -= e->second.data.size();
      cache_.erase(e);
      it = cache_lru_.erase(it);
    } else {
      ++it;
    }
  }
}

void MemMan::BeginFrame(int client) {
  std::lock_guard<std::mutex> g(mu_);
  clients_[client].frame_start = ++clock_;
}

size_t MemMan::Acquire(int client, const int* gids, size_t n, int64_t* slots, void** addrs) {
  std::lock_guard<std::mutex> g(mu_);
  Client& K = clients_[client];
  SlotClass& C = classes_[K.cls];
  // evict the m least recently sampled slots that `ok` admits
  auto evict_lru = [&](size_t m, auto ok) {
    std::vector<std::pair<uint64_t, int64_t>> cand;
    for (size_t s = 0; s < C.slots.size(); ++s) {
      const Slot& sl = C.slots[s];
      if (sl.live && sl.client >= 0 && !sl.filling && !protected_for(sl, client) && ok(sl))
        cand.emplace_back(sl.stamp, (int64_t)s);
    }
    m = std::min(m, cand.size());
    std::partial_sort(cand.begin(), cand.begin() + m, cand.end());
    for (size_t i = 0; i < m; ++i) evict_slot(C, cand[i].second);
    return m;
  };
  size_t k = 0;
  while (k < n) {
    // the client's own share is full: reuse its least recent slots
    if (K.max_slots && K.held >= K.max_slots) {
      if (evict_lru(n - k, [&](const Slot& sl) { return sl.client == client; }) == 0) break;
    }
    size_t want = n - k;
    if (K.max_slots) want = std::min<uint64_t>(want, K.max_slots - std::min(K.max_slots, K.held));
    want = std::max<size_t>(want, 1);
    if (C.free_list.empty() && !grow(K.cls, want)) {
      // take back the least recently sampled slots of this size, else a
      // chunk of another size
      if (evict_lru(want, [](const Slot&) { return true; }) == 0) {
        if (!reclaim_chunk(K.cls, C.slot_bytes) || !grow(K.cls, want)) break;
      }
    }
    while (k < n && !C.free_list.empty() && (!K.max_slots || K.held < K.max_slots)) {
      const int64_t s = C.free_list.back();
      C.free_list.pop_back();
      Slot& sl = C.slots[s];
      sl.client = client;
      sl.gid = gids[k];
      sl.stamp = clock_;
      sl.filling = true;
      ++K.held;
      slots[k] = s;
      addrs[k] = slot_addr(C, s);
      ++k;
    }
  }
  return k;
}

void MemMan::Commit(int client, const int64_t* slots, size_t n) {
  std::lock_guard<std::mutex> g(mu_);
  SlotClass& C = classes_[clients_[client].cls];
  for (size_t i = 0; i < n; ++i)
    if (slots[i] >= 0 && (size_t)slots[i] < C.slots.size() && C.slots[slots[i]].client == client)
      C.slots[slots[i]].filling = false;
}

void MemMan::Touch(int client, const int64_t* slots, size_t n) {
  std::lock_guard<std::mutex> g(mu_);
  SlotClass& C = classes_[clients_[client].cls];
  const uint64_t t = ++clock_;
  for (size_t i = 0; i < n; ++i) {
    if (slots[i] < 0 || (size_t)slots[i] >= C.slots.size()) continue;
    Slot& sl = C.slots[slots[i]];
    if (sl.client == client) sl.stamp = t;
  }
}

void MemMan::Free(int client, int64_t s) {
  std::lock_guard<std::mutex> g(mu_);
  Client& K = clients_[client];
  SlotClass& C = classes_[K.cls];
  if (s < 0 || (size_t)s >= C.slots.size() || C.slots[s].client != client) return;
  C.slots[s] = Slot{};
  C.free_list.push_back(s);
  K.held -= std::min<uint64_t>(K.held, 1);
}

uint64_t MemMan::ClientSlots(int client) {
  std::lock_guard<std::mutex> g(mu_);
  return clients_[client].held;
}

// ------------------------------------------------------------- host cache
void MemMan::cache_trim_locked(uint64_t room) {
  const uint64_t budget = cpu_budget_.load(), used = cpu_used_.load();
  const uint64_t cap = budget > used ? budget - used : 0;
  while (cache_bytes_ + room > cap && !cache_lru_.empty()) {
    auto it = cache_.find(cache_lru_.back());
    cache_bytes_ -= it->second.data.size();
    cache_.erase(it);
    cache_lru_.pop_back();
  }
}

bool MemMan::CacheGet(int client, int gid, void* dst, uint64_t bytes) {
  std::lock_guard<std::mutex> g(cache_mu_);
  const uint64_t key = ((uint64_t)(uint32_t)client << 32) | (uint32_t)gid;
  auto it = cache_.find(key);
  if (it == cache_.end() || it->second.data.size() != bytes) {
    ++cache_misses_;
    return false;
  }
  memcpy(dst, it->second.data.data(), bytes);
  cache_lru_.splice(cache_lru_.begin(), cache_lru_, it->second.lru);
  ++cache_hits_;
  return true;
}

void MemMan::CachePut(int client, int gid, const void* src, uint64_t bytes) {
  std::lock_guard<std::mutex> g(cache_mu_);
  const uint64_t key = ((uint64_t)(uint32_t)client << 32) | (uint32_t)gid;
  if (cache_.count(key)) return;
  cache_trim_locked(bytes);
  const uint64_t budget = cpu_budget_.load(), used = cpu_used_.load();
  const uint64_t cap = budget > used ? budget - used : 0;
  if (cache_bytes_ + bytes > cap) return;
  CacheEntry e;
  try {
    e.data.assign((const uint8_t*)src, (const uint8_t*)src + bytes);
  } catch (const std::bad_alloc&) {
    return;
  }
  cache_lru_.push_front(key);
  e.lru = cache_lru_.begin();
  cache_bytes_ += bytes;
  cache_.emplace(key, std::move(e));
}

// ----------------------------------------------------------- worker pool
// The next index of the job tagged `tag`, if that job is still current and
// has indices left.  A compare-exchange, so that a worker that wakes late
// from an earlier job never consumes an index of the next one.
bool MemMan::take_index(uint64_t tag, size_t n, size_t* i) {
  size_t t = job_next_.load();
  for (;;) {
    if ((t >> kGenShift) != tag || (t & ((1ull << kGenShift) - 1)) >= n) return false;
    if (job_next_.compare_exchange_weak(t, t + 1)) {
      *i = t & ((1ull << kGenShift) - 1);
      return true;
    }
  }
}

void MemMan::worker() {
  uint64_t seen = 0;
  for (;;) {
    const std::function<void(size_t)>* fn;
    size_t n;
    uint64_t gen;
    {
      std::unique_lock<std::mutex> lk(wmu_);
      wcv_.wait(lk, [&] { return stop_ || job_gen_ != seen; });
      if (stop_) return;
      seen = gen = job_gen_;
      fn = job_;
      n = job_n_;
      ++job_active_;
    }
    const uint64_t tag = gen & ((1ull << (64 - kGenShift)) - 1);
    for (size_t i; take_index(tag, n, &i);) (*fn)(i);
    std::lock_guard<std::mutex> lk(wmu_);
    --job_active_;
    wdone_.notify_all();
  }
}

void MemMan::ParallelFor(size_t n, int threads, const std::function<void(size_t)>& fn) {
  threads = std::max(1, std::min(threads, kMaxWorkers + 1));
  if (threads == 1 || n <= 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::lock_guard<std::mutex> pg(par_mu_);
  {
    std::lock_guard<std::mutex> lk(wmu_);
    while ((int)workers_.size() < threads - 1) workers_.emplace_back([this] { worker(); });
    ++job_gen_;
    job_ = &fn;
    job_n_ = n;
    job_next_.store((size_t)(job_gen_ & ((1ull << (64 - kGenShift)) - 1)) << kGenShift);
  }
  wcv_.notify_all();
  const uint64_t tag = job_gen_ & ((1ull << (64 - kGenShift)) - 1);
  for (size_t i; take_index(tag, n, &i);) fn(i);
  // every index has been taken; wait for the workers still running one
  std::unique_lock<std::mutex> lk(wmu_);
  wdone_.wait(lk, [&] { return job_active_ == 0; });
  job_ = nullptr;
}

}  // namespace tvk
