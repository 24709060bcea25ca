// server.cpp -- RenderServer (see server.hpp).
#include "server.hpp"

#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>

#include "errors.hpp"
#include "kernels.hpp"
#include "memman.hpp"

namespace tvk {

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int fail(int code, const std::string& msg) {
  set_last_error(msg);
  return code;
}

constexpr uint64_t kAlign = 256;

}  // namespace

// ------------------------------------------------------------------ arena
void Arena::reset(uint64_t size) {
  size_ = size & ~(kAlign - 1);
  free_.clear();
  if (size_) free_[0] = size_;
}

bool Arena::alloc(uint64_t bytes, uint64_t* off) {
  bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    *off = it->first;
    const uint64_t rest = it->second - bytes, at = it->first + bytes;
    free_.erase(it);
    if (rest) free_[at] = rest;
    return true;
  }
  return false;
}

void Arena::free(uint64_t off, uint64_t bytes) {
  bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
  auto it = free_.emplace(off, bytes).first;
  auto next = std::next(it);
  if (next != free_.end() && it->first + it->second == next->first) {
    it->second += next->second;
    free_.erase(next);
  }
  if (it != free_.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      free_.erase(it);
    }
  }
}

// ----------------------------------------------------------------- server
RenderServer::~RenderServer() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (worker_.joinable()) worker_.join();
  for (auto& kv : paged_) {
    kv.second->cv.notify_all();
    pool_cv_.notify_all();
    if (kv.second->worker.joinable()) kv.second->worker.join();
  }
  if (device_ >= 0) {
    (void)hipSetDevice(device_);
    for (Lane& L : lanes_) {
      if (L.stream) (void)hipStreamSynchronize(L.stream);
      for (auto& kv : L.slots) kv.second->release();
      if (L.done) (void)hipEventDestroy(L.done);
      if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    dev_arena_mem_.reset();
    if (host_arena_mem_) (void)hipHostFree(host_arena_mem_);
  }
}

int RenderServer::Init(int device, uint64_t arena_bytes, uint32_t lanes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return fail(TVK_ERR_ARG, "device ordinal out of range");
  if (lanes < 1 || lanes > 16) return fail(TVK_ERR_ARG, "lanes must be in [1, 16]");
  if (arena_bytes < (1u << 20)) return fail(TVK_ERR_ARG, "arena must hold at least 1 MiB");
  device_ = device;
  if (hipSetDevice(device_) != hipSuccess) return fail(TVK_ERR_HIP, "hipSetDevice failed");
  lanes_.resize(lanes);
  for (Lane& L : lanes_) {
    if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess)
      return fail(TVK_ERR_HIP, "stream / event creation failed");
  }
  if (dev_arena_mem_.alloc(arena_bytes) != hipSuccess)
    return fail(TVK_ERR_OOM, "device output arena allocation failed");
  if (hipHostMalloc((void**)&host_arena_mem_, arena_bytes, hipHostMallocDefault) != hipSuccess)
    return fail(TVK_ERR_OOM, "pinned host output arena allocation failed");
  dev_arena_.reset(arena_bytes);
  host_arena_.reset(arena_bytes);
  worker_ = std::thread([this] { loop(); });
  return TVK_OK;
}

int RenderServer::AddVolume(HipRenderer* r, uint32_t* id) {
  if (!r || !id) return fail(TVK_ERR_ARG, "null renderer or id");
  if (r->Device() != device_) return fail(TVK_ERR_ARG, "renderer is on another device");
  std::lock_guard<std::mutex> lk(mu_);
  volumes_.push_back(r);
  *id = (uint32_t)(volumes_.size() - 1);
  if (r->Paged()) {
    auto pv = std::make_unique<PagedVolume>();
    pv->id = *id;
    pv->need = r->PagedWorkingSetBytes();
    PagedVolume* p = pv.get();
    paged_[*id] = std::move(pv);
    p->worker = std::thread([this, p] { paged_loop(p); });
  }
  return TVK_OK;
}

uint64_t RenderServer::pool_capacity() const {
  MemManStats m;
  MemMan::ForDevice(device_).Stats(&m);
  const uint64_t outside = m.gpu_used > m.pool_bytes ? m.gpu_used - m.pool_bytes : 0;
  return m.gpu_budget > outside ? m.gpu_budget - outside : 0;
}

// Batches of one paged volume's requests, one render mode each, admitted
// by brick-pool memory and rendered with the renderer's paging loop.
void RenderServer::paged_loop(PagedVolume* pv) {
  (void)hipSetDevice(device_);
  HipRenderer* r = volumes_[pv->id];
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    pv->cv.wait(lk, [&] { return stop_ || !pv->pending.empty(); });
    if (stop_) break;
    const int mode = reqs_[pv->pending.front()].q.render_mode;
    std::vector<uint64_t> pick;
    for (auto it = pv->pending.begin(); it != pv->pending.end() && (int)pick.size() < kMaxJobs;) {
      Request& q = reqs_[*it];
      if (q.q.render_mode != mode) {
        ++it;
        continue;
      }
      uint64_t hoff = 0;
      if (!host_arena_.alloc(q.bytes, &hoff)) break;  // output memory, in order
      q.host_off = hoff;
      q.host_held = true;
      pick.push_back(*it);
      it = pv->pending.erase(it);
    }
    if (pick.empty()) {
      // host output memory is held by finished requests not collected yet
      cv_done_.notify_all();
      pv->cv.wait_for(lk, std::chrono::milliseconds(1));
      continue;
    }
    // admission by brick-pool memory: this batch's working set beside the
    // paged batches in flight (one batch alone always runs)
    bool waited = false;
    while (!stop_ && pool_reserved_ > 0 && pool_reserved_ + pv->need > pool_capacity()) {
      waited = true;
      pool_cv_.wait(lk);
    }
    if (waited) ++stats_.pool_waits;
    pool_reserved_ += pv->need;
    stats_.max_pool_reserved = std::max(stats_.max_pool_reserved, pool_reserved_);
    const double t = now_ms();
    std::vector<tvk_view> views(pick.size());
    for (size_t k = 0; k < pick.size(); ++k) {
      Request& q = reqs_[pick[k]];
      q.t_launch = t;
      q.launch_jobs = (uint32_t)pick.size();
      views[k] = q.q.view;
      views[k].rgba = host_arena_mem_ + q.host_off;
    }
    lk.unlock();
    std::vector<tvk_view_stats> vs(pick.size());
    const int rc = r->ServeBatch(views.data(), (uint32_t)views.size(), mode, vs.data());
    const std::string err = rc == TVK_OK ? std::string() : last_error();
    lk.lock();
    pool_reserved_ -= pv->need;
    pool_cv_.notify_all();
    const double td = now_ms();
    for (size_t k = 0; k < pick.size(); ++k) {
      Request& q = reqs_[pick[k]];
      q.done = true;
      q.t_done = td;
      q.status = rc;
      q.error = err;
      if (rc == TVK_OK) {
        q.evaluated = vs[k].evaluated_samples;
        q.nominal = vs[k].nominal_samples;
        stats_.nominal_samples += q.nominal;
      }
      ++stats_.completed;
    }
    ++stats_.launches;
    ++stats_.paged_launches;
    stats_.max_jobs_per_launch = std::max<uint64_t>(stats_.max_jobs_per_launch, pick.size());
    cv_done_.notify_all();
  }
  // requests never rendered fail with a state error
  for (uint64_t t : pv->pending) {
    Request& q = reqs_[t];
    q.status = TVK_ERR_STATE;
    q.error = "the server stopped before the request ran";
    q.done = true;
  }
  pv->pending.clear();
  cv_done_.notify_all();
}

uint64_t RenderServer::variant_key(const LaunchInfo& li) {
  return (uint64_t)li.bpv | (uint64_t)li.interp << 4 | (uint64_t)li.mode << 8 |
         (uint64_t)li.lighting << 12 | (uint64_t)li.batch << 16 | (uint64_t)li.regions << 24 |
         (uint64_t)li.stage << 28;
}

int RenderServer::Submit(const tvk_request& q, uint64_t* ticket) {
  if (!ticket) return fail(TVK_ERR_ARG, "null ticket");
  std::unique_lock<std::mutex> lk(mu_);
  if (stop_) return fail(TVK_ERR_STATE, "server is stopping");
  if (q.volume >= volumes_.size()) return fail(TVK_ERR_ARG, "unknown volume id");
  HipRenderer* r = volumes_[q.volume];
  if (r->Paged()) {
    if (int e = r->ServeCheck(q.view, q.render_mode)) return e;
    Request req;
    req.q = q;
    req.paged = true;
    req.bytes = (uint64_t)q.view.width * q.view.height * 4;
    if (req.bytes > host_arena_.size()) return fail(TVK_ERR_OOM, "image larger than the output arena");
    req.t_submit = now_ms();
    const uint64_t t = next_ticket_++;
    reqs_.emplace(t, req);
    PagedVolume* pv = paged_[q.volume].get();
    pv->pending.push_back(t);
    ++stats_.submitted;
    *ticket = t;
    lk.unlock();
    pv->cv.notify_one();
    return TVK_OK;
  }
  if (int e = r->ServePrepare(q.view.tf_slot, q.render_mode)) return e;
  Request req;
  req.q = q;
  RaycastParams probe;
  LaunchInfo li;
  if (int e = r->ServeJob(q.view, q.render_mode, nullptr, &probe, &li)) return e;
  req.key = variant_key(li);
  {
    const float longest = std::max(probe.box[0], std::max(probe.box[1], probe.box[2]));
    req.cost = (uint64_t)probe.rect_w * (uint64_t)probe.rect_h * 64ull *
               (uint64_t)std::max(1.0f, longest / probe.step);
  }
  req.tiles = job_tiles(probe);
  if (req.cost > part_cost_ && req.tiles > 1) {
    const uint64_t n = (req.cost + part_cost_ - 1) / part_cost_;
    req.parts = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(n, kMaxParts), (uint64_t)req.tiles);
    ++stats_.split_requests;
  }
  req.bytes = (uint64_t)q.view.width * q.view.height * 4;
  if (req.bytes > dev_arena_.size()) return fail(TVK_ERR_OOM, "image larger than the output arena");
  req.t_submit = now_ms();
  const uint64_t t = next_ticket_++;
  reqs_.emplace(t, req);
  pending_.push_back(t);
  ++stats_.submitted;
  *ticket = t;
  lk.unlock();
  cv_work_.notify_one();
  return TVK_OK;
}

bool RenderServer::retire_locked() {
  bool any = false;
  for (Lane& L : lanes_) {
    if (!L.busy) continue;
    const hipError_t e = hipEventQuery(L.done);
    if (e == hipErrorNotReady) continue;
    const double t = now_ms();
    bool failed = false;
    for (size_t k = 0; k < L.tickets.size(); ++k) {
      Request& r = reqs_[L.tickets[k]];
      ++r.parts_retired;
      if (e == hipSuccess) {
        const uint64_t nominal = L.cur->counters_h[kQueues + kJobCounters * k + 1];
        r.evaluated += L.cur->counters_h[kQueues + kJobCounters * k];
        r.nominal += nominal;
        stats_.nominal_samples += nominal;
      } else if (r.status == TVK_OK) {
        r.status = TVK_ERR_HIP;
        r.error = std::string("launch failed: ") + hipGetErrorString(e);
        failed = true;
      }
      // done after its last part (whose launch copied the image out), or
      // at the first part that failed
      if (r.parts_retired == r.parts || r.status != TVK_OK) {
        r.done = true;
        r.t_done = t;
        dev_arena_.free(r.dev_off, r.bytes);
        ++stats_.completed;
      }
    }
    // a failed request's parts not yet launched never run
    if (failed) drop_done_locked();
    L.tickets.clear();
    L.busy = false;
    any = true;
  }
  if (any) cv_done_.notify_all();
  return any;
}

bool RenderServer::launch_locked(Lane& L) {
  // The launch's kernel variant is that of the oldest request without a
  // part on the GPU; every waiting request of that variant joins it, oldest
  // first, with its next part: after the first, no more than
  // kLaunchCostCap estimated samples per launch, so that light requests
  // are not held until a heavy one in the same launch finishes (a skipped
  // request keeps its place and leads a later launch).
  std::vector<uint64_t> pick;
  std::vector<uint32_t> part;
  uint64_t key = 0, cost = 0;
  bool keyed = false;
  for (uint64_t tk : pending_) {
    if ((int)pick.size() >= kMaxJobs) break;
    Request& r = reqs_[tk];
    if (r.parts_launched > r.parts_retired) continue;  // its previous part is on the GPU
    if (!keyed) {
      key = r.key;
      keyed = true;
    }
    const uint64_t c = r.cost / r.parts;
    if (r.key != key || (!pick.empty() && cost + c > kLaunchCostCap)) continue;
    if (r.parts_launched == 0) {
      uint64_t doff = 0, hoff = 0;
      if (!dev_arena_.alloc(r.bytes, &doff)) break;  // admission by output memory, in order
      if (!host_arena_.alloc(r.bytes, &hoff)) {
        dev_arena_.free(doff, r.bytes);
        break;
      }
      r.dev_off = doff;
      r.host_off = hoff;
      r.host_held = true;
    }
    pick.push_back(tk);
    part.push_back(r.parts_launched);
    cost += c;
  }
  if (pick.empty()) return false;
  std::vector<RaycastParams> jobs(pick.size());
  LaunchInfo li;
  bool ok = true;
  for (size_t k = 0; k < pick.size() && ok; ++k) {
    Request& r = reqs_[pick[k]];
    uint32_t* fb = (uint32_t*)((uint8_t*)dev_arena_mem_.p + r.dev_off);
    ok = volumes_[r.q.volume]->ServeJob(r.q.view, r.q.render_mode, fb, &jobs[k], &li) == TVK_OK;
    if (ok && r.parts > 1) {
      const int64_t first = (int64_t)part[k] * r.tiles / r.parts;
      jobs[k].tile_first = (int32_t)first;
      jobs[k].tile_count = (int32_t)((int64_t)(part[k] + 1) * r.tiles / r.parts - first);
    }
  }
  // a 1D TF shared by every job of the launch is staged in LDS
  bool shared = ok && li.mode == 0 && jobs[0].n_tf <= 4096;
  for (size_t k = 1; k < jobs.size() && shared; ++k)
    shared = jobs[k].tf == jobs[0].tf && jobs[k].n_tf == jobs[0].n_tf;
  li.n_tf_lds = shared ? jobs[0].n_tf : 0;
  const uint64_t slot_key = variant_key(li) ^ ((uint64_t)li.n_tf_lds << 32);
  auto& slot = L.slots[slot_key];
  std::string why = ok ? std::string() : last_error();
  if (ok && !slot) {
    std::unique_ptr<LaunchSlot> fresh(new LaunchSlot());
    const int rc = fresh->alloc();
    if (rc == TVK_OK) {
      slot = std::move(fresh);
    } else {
      bool in_flight = pool_reserved_ > 0;
      for (const Lane& x : lanes_) in_flight |= x.busy;
      if (rc == TVK_ERR_OOM && in_flight) {
        // device memory is held by work in flight (a paged batch filling
        // pool slots, other lanes): the requests keep their places and
        // launch once that work is done
        for (size_t k = 0; k < pick.size(); ++k) {
          if (part[k] != 0) continue;
          Request& r = reqs_[pick[k]];
          dev_arena_.free(r.dev_off, r.bytes);
          host_arena_.free(r.host_off, r.bytes);
          r.host_held = false;
        }
        memory_wait_ = true;
        return false;
      }
      ok = false;
      why = last_error();
    }
  }
  bool enqueued = false;  // device work of this launch may be in flight
  if (ok) {
    L.cur = slot.get();
    enqueued = true;
    ok = launch_job_table(*slot, jobs.data(), (int)jobs.size(), li, 4, 0, L.stream, true) == TVK_OK;
    if (!ok) why = last_error();
  }
  // the images finished by this launch (a request's last part) go to host memory
  for (size_t k = 0; k < pick.size() && ok; ++k) {
    Request& r = reqs_[pick[k]];
    if (part[k] + 1 < r.parts) continue;
    ok = hipMemcpyAsync(host_arena_mem_ + r.host_off, (uint8_t*)dev_arena_mem_.p + r.dev_off, r.bytes,
                        hipMemcpyDeviceToHost, L.stream) == hipSuccess;
  }
  if (ok) ok = hipEventRecord(L.done, L.stream) == hipSuccess;
  const double t = now_ms();
  if (!ok) {
    // the kernel or a copy may still write the picked arena blocks: let the
    // lane's stream drain before they go back to the arena
    if (enqueued) (void)hipStreamSynchronize(L.stream);
    for (uint64_t tk : pick) {
      Request& r = reqs_[tk];
      r.status = TVK_ERR_HIP;
      r.error = why.empty() ? "launch could not be enqueued" : why;
      r.done = true;
      r.t_done = t;
      if (r.t_launch == 0) r.t_launch = t;
      dev_arena_.free(r.dev_off, r.bytes);
      ++stats_.completed;
    }
    drop_done_locked();
    cv_done_.notify_all();
    return true;
  }
  for (size_t k = 0; k < pick.size(); ++k) {
    Request& r = reqs_[pick[k]];
    if (part[k] == 0) r.t_launch = t;
    r.launch_jobs = (uint32_t)pick.size();
    ++r.parts_launched;
  }
  // requests whose last part is on the GPU leave the queue
  pending_.erase(std::remove_if(pending_.begin(), pending_.end(),
                                [&](uint64_t tk) {
                                  const Request& r = reqs_[tk];
                                  return r.parts_launched == r.parts;
                                }),
                 pending_.end());
  stats_.parts_launched += pick.size();
  L.tickets = pick;
  L.busy = true;
  ++stats_.launches;
  stats_.max_jobs_per_launch = std::max<uint64_t>(stats_.max_jobs_per_launch, pick.size());
  uint64_t busy = 0;
  for (const Lane& x : lanes_) busy += x.busy ? 1 : 0;
  stats_.max_lanes_busy = std::max(stats_.max_lanes_busy, busy);
  return true;
}

void RenderServer::loop() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    bool progress = retire_locked();
    for (Lane& L : lanes_)
      if (!L.busy && !pending_.empty() && !stop_) progress |= launch_locked(L);
    bool busy = false;
    for (const Lane& L : lanes_) busy |= L.busy;
    if (stop_ && !busy) break;
    if (progress) continue;
    if (memory_wait_) {
      // a launch waits for device memory held by work in flight
      memory_wait_ = false;
      cv_work_.wait_for(lk, std::chrono::microseconds(500));
    } else if (busy) {
      // lanes in flight: poll their events at a short interval
      cv_work_.wait_for(lk, std::chrono::microseconds(50));
    } else {
      cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
    }
  }
  // requests never launched fail with a state error
  for (uint64_t t : pending_) {
    Request& r = reqs_[t];
    r.status = TVK_ERR_STATE;
    r.error = "the server stopped before the request ran";
    r.done = true;
  }
  pending_.clear();
  cv_done_.notify_all();
}

int RenderServer::Wait(uint64_t ticket, uint32_t timeout_ms, tvk_request_result* out) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(ticket);
  if (it == reqs_.end()) return fail(TVK_ERR_ARG, "unknown or retired ticket");
  auto ready = [&] { return reqs_[ticket].done; };
  if (timeout_ms == 0) {
    cv_done_.wait(lk, ready);
  } else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return fail(TVK_ERR_STATE, "request not finished within the time-out");
  }
  Request r = reqs_[ticket];
  reqs_.erase(ticket);
  const bool copy = r.status == TVK_OK && r.q.view.rgba;
  lk.unlock();
  // the host block stays reserved until the copy into the caller's buffer
  if (copy) memcpy(r.q.view.rgba, host_arena_mem_ + r.host_off, r.bytes);
  if (r.host_held) {
    std::lock_guard<std::mutex> g(mu_);
    host_arena_.free(r.host_off, r.bytes);
  }
  cv_work_.notify_one();  // output memory was freed
  {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& kv : paged_) kv.second->cv.notify_one();
  }
  if (out) {
    *out = tvk_request_result{};
    out->status = r.status;
    out->launch_jobs = r.launch_jobs;
    out->parts = r.parts;
    out->evaluated_samples = r.evaluated;
    out->nominal_samples = r.nominal;
    out->queue_ms = r.t_launch > 0 ? r.t_launch - r.t_submit : 0.0;
    out->latency_ms = (copy ? now_ms() : r.t_done) - r.t_submit;
  }
  if (r.status != TVK_OK) return fail(r.status, "request failed" + (r.error.empty() ? "" : ": " + r.error));
  return TVK_OK;
}

void RenderServer::drop_done_locked() {
  pending_.erase(std::remove_if(pending_.begin(), pending_.end(), [&](uint64_t tk) { return reqs_[tk].done; }),
                 pending_.end());
}

int RenderServer::SetPartCost(uint64_t samples) {
  std::lock_guard<std::mutex> lk(mu_);
  part_cost_ = samples == 0 ? kPartCost : samples;
  return TVK_OK;
}

int RenderServer::Stats(tvk_server_stats* out) {
  if (!out) return fail(TVK_ERR_ARG, "null stats");
  std::lock_guard<std::mutex> lk(mu_);
  *out = stats_;
  return TVK_OK;
}

}  // namespace tvk
