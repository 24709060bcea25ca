// memman.hpp -- the device-wide memory manager: the role of Tuvok's GPUMemMan
// with the SysInfo budgets ImageVis3D sets (SetMaxUsableGPUMem /
// SetMaxUsableCPUMem + MemSizesChanged, ImageVis3D/UI/ImageVis3D_Settings.cpp
// :556-559) and the accounting its debug window reads (GetGPUMem,
// GetAllocatedGPUMem, GetCPUMem, GetAllocatedCPUMem,
// ImageVis3D_DebugWindow.cpp:172-206).
//
// One MemMan per device, shared by every renderer (and the render server)
// on it:
//  - a GPU budget over all device memory the library allocates there
//    (every DevBuf is accounted), and a CPU budget over host memory held
//    for bricks (the host brick cache, host copies of paged levels);
//  - the paged brick pool: slots of every paged volume of every renderer
//    come from chunked arenas (one arena per slot size, chunks allocated
//    on demand within the budget, sized to the request that needs them:
//    at least 16 MiB, at most 512 MiB, so the pool does not take the
//    budget's headroom before it needs it).  Page tables hold the
//    slots' device addresses, so a brick can live in any chunk.  When a
//    request cannot be met from free slots or budget, the least recently
//    sampled slot of any volume of that size is taken back (its owner
//    first waits for its queued device work, then drops the brick from its
//    page table); when no slot of that size may go, a whole chunk of
//    another size is emptied and freed for it.  A volume's own requests
//    never take the slots it sampled in its current frame (the frame needs
//    them); another volume's may, after the owner's queued work is done,
//    and the owner pages them in again.  Slots being filled are taken by
//    no one.  Each volume may also cap its own share (its pool size),
//    evicting its own least recent bricks;
//  - a host brick cache (LRU, within the CPU budget) of brick contents as
//    the sources produced them, so a brick paged out and needed again does
//    not run its source (disk read, decompression) a second time;
//  - worker threads on which brick sources run in parallel.
//
// Locking: the MemMan's lock is taken by its own calls only.  A client's
// PoolEvict runs under it, so clients never call the MemMan while holding
// a lock that PoolEvict takes.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <list>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tvk {

// A paged volume holding slots of the shared pool.
class PoolClient {
 public:
  virtual ~PoolClient() = default;
  // The pool takes back the slot of brick `gid`.  Before returning, the
  // client waits for its queued device work that may read the slot and
  // drops the brick from its page table.  Runs under the MemMan lock.
  virtual void PoolEvict(int gid) = 0;
};

struct MemManStats {
  uint64_t gpu_total, gpu_budget, gpu_used;  // device memory, budget, bytes allocated
  uint64_t cpu_total, cpu_budget, cpu_used;  // host memory, budget, bytes held
  uint64_t pool_bytes, pool_slots, pool_resident;  // paged pool chunks, their slots, slots in use
  uint64_t pool_evictions, chunks_freed;
  uint64_t cache_bytes, cache_hits, cache_misses;
};

class MemMan {
 public:
  static MemMan& ForDevice(int device);

  void SetGPUBudget(uint64_t bytes);
  void SetCPUBudget(uint64_t bytes);
  uint64_t GPUBudget();
  void Stats(MemManStats* out);

  // Device memory outside the paged pool (accounted by DevBuf).  Reserve
  // frees paged-pool chunks if that makes room; false when it cannot.
  bool ReserveGPU(uint64_t bytes);
  void ReleaseGPU(uint64_t bytes);
  // Host memory held for bricks outside the cache (the cache evicts itself).
  bool ReserveCPU(uint64_t bytes);
  void ReleaseCPU(uint64_t bytes);

  // Paged pool.  max_slots caps the client's share (0: no cap).
  int Register(PoolClient* c, uint64_t slot_bytes, uint64_t max_slots);
  void Unregister(int client);
  // From now on the slots the client touches or receives are its current
  // frame's: never taken from it until its next BeginFrame.
  void BeginFrame(int client);
  // Slots for bricks gids[0..n): returns how many were given (the first
  // k bricks, k <= n), with their slot ids and device addresses.  They stay
  // pinned -- no one may take them -- until Commit: the client fills them
  // and enters them in its page table in between.
  size_t Acquire(int client, const int* gids, size_t n, int64_t* slots, void** addrs);
  void Commit(int client, const int64_t* slots, size_t n);
  // The client sampled these slots in its current frame (LRU stamps).
  void Touch(int client, const int64_t* slots, size_t n);
  // A slot the client no longer needs (its brick left the page table).
  void Free(int client, int64_t slot);
  uint64_t ClientSlots(int client);

  // Host brick cache, keyed by (client, brick).
  bool CacheGet(int client, int gid, void* dst, uint64_t bytes);
  void CachePut(int client, int gid, const void* src, uint64_t bytes);

  // Runs fn(0..n) on up to `threads` workers (the caller works too) and
  // returns when all are done.
  void ParallelFor(size_t n, int threads, const std::function<void(size_t)>& fn);

 private:
  explicit MemMan(int device);
  struct Slot {
    int client = -1;
    int gid = -1;
    uint64_t stamp = 0;
    bool live = true;     // its chunk is allocated
    bool filling = false; // acquired, not committed yet
  };
  struct SlotClass {
    uint64_t slot_bytes = 0;
    uint32_t per_chunk = 1;
    std::vector<void*> chunks;     // nullptr: freed
    std::vector<uint32_t> chunk_slots;  // slots in each chunk
    std::vector<Slot> slots;       // chunk c's slots at [c * per_chunk, ...)
    std::vector<int64_t> free_list;
  };
  struct Client {
    PoolClient* c = nullptr;
    int cls = -1;
    uint64_t max_slots = 0;
    uint64_t held = 0;
    uint64_t frame_start = 0;
    bool alive = false;
  };
  struct CacheEntry {
    std::vector<uint8_t> data;
    std::list<uint64_t>::iterator lru;
  };

  int class_for(uint64_t slot_bytes);
  bool grow(int cls, uint64_t want);        // one more chunk within the budget
  bool reclaim_chunk(int exclude_cls, uint64_t need_bytes);  // free another class's chunk
  void evict_slot(SlotClass& C, int64_t s);
  bool protected_for(const Slot& sl, int client) const;
  void cache_trim_locked(uint64_t room);
  void* slot_addr(const SlotClass& C, int64_t s) const;

  int device_;
  std::mutex mu_;
  uint64_t gpu_total_ = 0, gpu_budget_ = 0, gpu_used_ = 0;
  uint64_t cpu_total_ = 0;
  std::atomic<uint64_t> cpu_budget_{0}, cpu_used_{0};  // also read by the cache without mu_
  uint64_t clock_ = 1;
  uint64_t evictions_ = 0, chunks_freed_ = 0;
  std::vector<SlotClass> classes_;
  std::vector<Client> clients_;

  std::mutex cache_mu_;
  std::unordered_map<uint64_t, CacheEntry> cache_;
  std::list<uint64_t> cache_lru_;  // front = most recent
  uint64_t cache_bytes_ = 0, cache_hits_ = 0, cache_misses_ = 0;

  // worker pool
  void worker();
  bool take_index(uint64_t tag, size_t n, size_t* i);
  std::vector<std::thread> workers_;
  std::mutex wmu_;
  std::condition_variable wcv_, wdone_;
  const std::function<void(size_t)>* job_ = nullptr;
  size_t job_n_ = 0;
  std::atomic<size_t> job_next_{0};
  size_t job_active_ = 0;
  int job_threads_ = 0;
  uint64_t job_gen_ = 0;
  bool stop_ = false;
  std::mutex par_mu_;  // one ParallelFor at a time
};

}  // namespace tvk
