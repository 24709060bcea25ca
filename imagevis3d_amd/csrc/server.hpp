// server.hpp -- RenderServer: continuous batching of render requests over
// several resident volumes (tvk_server_* in tvk_hip.h).
//
// A worker thread owns `lanes` HIP streams, each with its own job table.
// Requests wait in arrival order; whenever a lane is free the oldest
// request's kernel variant picks the launch, and every waiting request of
// that variant whose output fits the arenas joins it (up to kMaxJobs, and
// up to kLaunchCostCap estimated samples after the first request, so that a
// heavy request does not hold light ones until it finishes).  A request
// heavier than the part cost renders in parts: its tile slots split evenly
// over several launches, one part on the GPU at a time, so that the
// requests behind it run between its parts instead of waiting for the whole
// image.  A launch renders into a device arena block per request and copies
// each finished image to a pinned host arena block on the lane's stream;
// the lane is retired when its event completes, which frees the device
// blocks of finished requests and wakes the waiters.  Lanes run
// concurrently, so requests that arrive while others are on the GPU start
// as soon as a lane and memory free up.
#pragma once
#include <hip/hip_runtime_api.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/tvk_hip.h"
#include "renderer.hpp"

namespace tvk {

// Estimated samples (tile-rect rays x the volume's longest axis in steps)
// one launch may carry beyond its first request: about 5 ms of raycasting.
constexpr uint64_t kLaunchCostCap = 1000000000ull;
// Default part cost: a request of more estimated samples renders in parts of
// at most this many (about 20 ms of raycasting each), and at most kMaxParts.
constexpr uint64_t kPartCost = 4000000000ull;
constexpr uint32_t kMaxParts = 64;

// First-fit allocator over [0, size) in 256-byte units, with coalescing.
class Arena {
 public:
  void reset(uint64_t size);
  bool alloc(uint64_t bytes, uint64_t* off);
  void free(uint64_t off, uint64_t bytes);
  uint64_t size() const { return size_; }

 private:
  uint64_t size_ = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
};

class RenderServer {
 public:
  ~RenderServer();
  int Init(int device, uint64_t arena_bytes, uint32_t lanes);
  int AddVolume(HipRenderer* r, uint32_t* id);
  int Submit(const tvk_request& q, uint64_t* ticket);
  int Wait(uint64_t ticket, uint32_t timeout_ms, tvk_request_result* out);
  int Stats(tvk_server_stats* out);
  int SetPartCost(uint64_t samples);

 private:
  struct Request {
    tvk_request q;
    uint64_t key = 0;        // kernel variant
    double t_submit = 0, t_launch = 0, t_done = 0;
    uint64_t dev_off = 0, host_off = 0, bytes = 0;
    uint64_t cost = 0;       // samples if every ray of its tile rect ran the longest chord
    uint64_t evaluated = 0, nominal = 0;  // summed over its parts
    uint32_t launch_jobs = 0;
    int tiles = 0;           // its tile slots
    uint32_t parts = 1;      // launches it renders in: part p takes slots [p T / parts, (p+1) T / parts)
    uint32_t parts_launched = 0, parts_retired = 0;
    int status = TVK_OK;
    std::string error;       // why it failed (status != TVK_OK)
    bool done = false;
    bool host_held = false;  // owns a pinned host arena block
    bool paged = false;      // its volume pages through the device pool
  };
  // A paged volume: its requests are rendered in batches by a worker of its
  // own (the renderer's multi-pass paging loop), admitted when their
  // working set fits the device pool beside the batches already running.
  struct PagedVolume {
    uint32_t id = 0;
    uint64_t need = 0;  // pool bytes a batch may hold
    std::deque<uint64_t> pending;
    std::condition_variable cv;
    std::thread worker;
  };
  struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // one job table (and captured graph) per kernel variant
    std::map<uint64_t, std::unique_ptr<LaunchSlot>> slots;
    LaunchSlot* cur = nullptr;  // the slot of the launch in flight
    bool busy = false;
    std::vector<uint64_t> tickets;
  };

  void loop();
  void paged_loop(PagedVolume* pv);
  uint64_t pool_capacity() const;  // device pool bytes available to served batches
  bool retire_locked();  // true when a lane finished
  bool launch_locked(Lane& lane);
  void drop_done_locked();  // finished (or failed) requests leave the queue
  static uint64_t variant_key(const LaunchInfo& li);

  int device_ = -1;
  std::vector<HipRenderer*> volumes_;
  std::map<uint32_t, std::unique_ptr<PagedVolume>> paged_;
  uint64_t pool_reserved_ = 0;  // working sets of the paged batches in flight
  std::condition_variable pool_cv_;
  std::vector<Lane> lanes_;
  DevBuf dev_arena_mem_;
  uint8_t* host_arena_mem_ = nullptr;  // pinned
  Arena dev_arena_, host_arena_;

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<uint64_t> pending_;
  std::unordered_map<uint64_t, Request> reqs_;
  uint64_t next_ticket_ = 1;
  bool memory_wait_ = false;  // the last launch was put back for device memory
  uint64_t part_cost_ = kPartCost;
  bool stop_ = false;
  std::thread worker_;
  tvk_server_stats stats_{};
};

}  // namespace tvk
