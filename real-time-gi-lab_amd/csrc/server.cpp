// Request server: continuous batching of GI updates over resident worlds under a device-memory budget.
//
// Threads: the caller's (add_world / submit / poll), a scheduler that forms and starts batched launches, and a
// completion thread that waits for them in launch order. One mutex guards the queue, the world table, the slots and
// the completion list; HIP work (world creation, copies, launches, waits) runs with it released.
//
// A slot is one stream with its own descriptor and block-prefix buffers (device and pinned host), so a batched launch
// never waits for another slot's. A world is in at most one launch at a time; requests of one world keep their order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "rtgi.h"
#include "rtgi_internal.h"
#include "world.h"

namespace {

using Clock = std::chrono::steady_clock;

// Host copy of a registered world's scene in pinned memory: an admission uploads it by DMA at full link speed.
template <class T>
struct PinnedVec {
  T* p = nullptr;
  PinnedVec() = default;
  PinnedVec(const PinnedVec&) = delete;
  PinnedVec& operator=(const PinnedVec&) = delete;
  ~PinnedVec() {
    if (p) (void)hipHostFree(p);
  }
  bool assign(const T* src, size_t n) {
    if (hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    if (n) std::memcpy(p, src, n * sizeof(T));
    return true;
  }
  T* data() { return p; }
};

struct SrvWorld {
  rtgi_vox_params vp{};
  rtgi_rc_params rp{};
  PinnedVec<float> pos, nrm, mat;
  PinnedVec<uint32_t> tri;
  uint32_t nverts = 0, ntris = 0, nmat = 0;
  uint64_t bytes = 0;     // device bytes once resident
  rtgi_world* w = nullptr;  // resident world, or null
  bool busy = false;      // in a launch (or being admitted)
  uint64_t last_use = 0;  // launch counter at its last launch (LRU order)
  int queued = 0;         // its requests in the queue
  uint64_t skip_pass = 0; // gather pass in which one of its requests was passed over
  int blocked = 0;        // consecutive passes in which it could not be admitted
  WorldState saved;       // temporal state (SH history, bounce volume) kept over an eviction
};

struct Req {
  rtgi_request r;
  double t_submit = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  TraceK* d_desc = nullptr;
  uint32_t* d_start = nullptr;
  uint32_t* d_queue = nullptr;  // the slot's work-queue counter
  TraceK* h_desc = nullptr;    // pinned
  uint32_t* h_start = nullptr;  // pinned
  bool busy = false;
  std::vector<Req> reqs;
  double t_launch = 0;
  int status = RTGI_OK;
};

}  // namespace

struct rtgi_server {
  int device = 0;
  rtgi_server_params p{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_poll;
  std::list<Req> queue;
  std::vector<std::unique_ptr<SrvWorld>> worlds;
  std::vector<Slot> slots;
  std::deque<int> inflight;  // slot indices in launch order
  std::deque<rtgi_completion> done;
  rtgi_completion_fn fn = nullptr;
  void* user = nullptr;
  rtgi_server_stats st{};
  uint64_t launches = 0, passes = 0;
  bool stop = false;
  std::thread sched, reaper;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void slot_free(Slot& s) {
  if (s.h_desc) (void)hipHostFree(s.h_desc);
  if (s.h_start) (void)hipHostFree(s.h_start);
  if (s.d_desc) (void)hipFree(s.d_desc);
  if (s.d_start) (void)hipFree(s.d_start);
  if (s.d_queue) (void)hipFree(s.d_queue);
  if (s.done) (void)hipEventDestroy(s.done);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = Slot{};
}

// Creates the device world of sw from its host copy and restores the temporal state it had when it was evicted
// (called without the lock).
int make_resident(SrvWorld& sw) {
  rtgi_world* w = nullptr;
  int e = rtgi_world_create(&sw.vp, &sw.rp, &w);
  if (!e)
    e = rtgi_world_upload_scene(w, sw.pos.data(), sw.nrm.data(), sw.nverts, sw.tri.data(), sw.ntris, sw.mat.data(),
                                sw.nmat);
  if (!e) e = world_restore_state(w, sw.saved);
  if (e) {
    if (w) rtgi_world_destroy(w);
    return e;
  }
  sw.w = w;
  return RTGI_OK;
}

// Makes room for `need` bytes by evicting idle resident worlds: first those with no queued request, each group least
// recently used first. Returns whether it fits. Caller holds the lock; the evicted worlds are handed back to have
// their temporal state saved and be freed without it.
using Victim = std::pair<SrvWorld*, rtgi_world*>;
bool make_room(rtgi_server* s, uint64_t need, std::vector<Victim>& victims) {
  uint64_t idle = 0;
  for (auto& x : s->worlds)
    if (x->w && !x->busy) idle += x->bytes;
  if (s->st.resident_bytes - idle + need > s->st.budget) return false;  // evicts nothing unless it then fits
  while (s->st.resident_bytes + need > s->st.budget) {
    SrvWorld* v = nullptr;
    for (auto& x : s->worlds) {
      if (!x->w || x->busy) continue;
      if (!v || (x->queued == 0) > (v->queued == 0) ||
          ((x->queued == 0) == (v->queued == 0) && x->last_use < v->last_use))
        v = x.get();
    }
    if (!v) return false;
    victims.push_back(Victim(v, v->w));
    v->w = nullptr;
    s->st.resident_bytes -= v->bytes;
    s->st.resident_worlds--;
    s->st.evictions++;
  }
  return true;
}

void sched_loop(rtgi_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  const int nslots = (int)s->slots.size();
  uint64_t seen = ~0ull;  // completion count at the last pass that found nothing to run
  for (;;) {
    int si = -1;
    for (int i = 0; i < nslots; i++)
      if (!s->slots[i].busy) si = si < 0 ? i : si;
    if (s->stop) break;
    if (si < 0 || s->queue.empty() || seen == s->st.completed + s->st.submitted) {
      s->cv_sched.wait(lk);
      continue;
    }
    // Gather: FIFO over the queue, one request per idle world. A world that is not resident is admitted when its
    // bytes fit (evicting idle worlds); once one cannot be admitted, no later world is, so freed memory goes to it.
    std::vector<Req> pick;
    std::vector<SrvWorld*> pw;
    bool admit_closed = false;
    const uint64_t pass = ++s->passes;
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)pick.size() < s->p.max_batch;) {
      SrvWorld* w = s->worlds[it->r.world].get();
      // Busy, or an earlier request of the world was passed over in this pass (its requests keep their order).
      if (w->busy || w->skip_pass == pass) {
        w->skip_pass = pass;
        ++it;
        continue;
      }
      if (!w->w) {
        std::vector<Victim> victims;
        if (admit_closed || !make_room(s, w->bytes, victims)) {
          w->skip_pass = pass;
          // Memory held by running launches: after a few tries, start nothing behind it until they end, so that
          // launches of resident worlds cannot keep it waiting for ever.
          if (!admit_closed && ++w->blocked > 2) break;
          admit_closed = true;
          ++it;
          continue;
        }
        w->blocked = 0;
        w->busy = true;  // reserved while it is created
        s->st.resident_bytes += w->bytes;
        s->st.peak_resident_bytes = std::max(s->st.peak_resident_bytes, s->st.resident_bytes);
        lk.unlock();
        for (const Victim& v : victims) {
          // An evicted world keeps its history: a failed save drops it (its next update starts a new history).
          if (world_save_state(v.second, &v.first->saved)) v.first->saved = WorldState{};
          rtgi_world_destroy(v.second);
        }
        int e = make_resident(*w);
        lk.lock();
        w->busy = false;
        if (e) {
          s->st.resident_bytes -= w->bytes;
          // The request fails; report it like a finished update.
          rtgi_completion c{it->r.tag, it->r.world, e, 0, it->t_submit, s->now(), s->now()};
          s->done.push_back(c);
          s->st.completed++;
          w->queued--;
          it = s->queue.erase(it);
          s->cv_poll.notify_all();
          s->cv_done.notify_all();
          continue;
        }
        s->st.resident_worlds++;
        s->st.admissions++;
      }
      w->busy = true;
      w->last_use = ++s->launches;
      pick.push_back(*it);
      pw.push_back(w);
      w->queued--;
      it = s->queue.erase(it);
    }
    if (pick.empty()) {
      seen = s->st.completed + s->st.submitted;
      continue;
    }
    seen = ~0ull;
    Slot& sl = s->slots[si];
    sl.busy = true;
    sl.reqs = pick;
    sl.t_launch = s->now();
    lk.unlock();

    // Launch: frames, descriptors and block prefixes through the slot's pinned buffers, then one launch per stage.
    // A request whose frame its world refuses (submit checks frames, so this does not happen in practice) completes
    // alone with the error; the others launch.
    int e = RTGI_OK;
    std::vector<rtgi_world*> ws;
    std::vector<rtgi_completion> refused;
    for (size_t i = 0; i < pick.size();) {
      rtgi_world* w = pw[i]->w;
      int fe = rtgi_world_set_frame(w, &pick[i].r.frame);
      if (fe) {
        refused.push_back(rtgi_completion{pick[i].r.tag, pick[i].r.world, fe, 0, pick[i].t_submit, sl.t_launch,
                                          s->now()});
        pick.erase(pick.begin() + (long)i);
        pw.erase(pw.begin() + (long)i);
        continue;
      }
      if (hipStreamSynchronize(w->stream) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "world stream");
      ws.push_back(w);
      i++;
    }
    const int n = (int)pick.size();
    if (!refused.empty()) {
      lk.lock();
      for (size_t i = 0; i < refused.size(); i++) s->worlds[refused[i].world]->busy = false;
      for (const rtgi_completion& c : refused) s->done.push_back(c);
      s->st.completed += refused.size();
      sl.reqs = pick;
      s->cv_poll.notify_all();
      if (n == 0) {
        sl.busy = false;
        s->cv_sched.notify_all();
        continue;
      }
      lk.unlock();
    }
    BatchPlan plan{};
    if (!e) {
      plan_blocks(ws.data(), n, sl.h_start, plan);
      for (int i = 0; i < n; i++) batch_world_begin(ws[i], plan, &sl.h_desc[i]);
      plan.sh_pairs = plan_sh_pairs(ws.data(), sl.h_desc, n);
      if (hipMemcpyAsync(sl.d_desc, sl.h_desc, sizeof(TraceK) * n, hipMemcpyHostToDevice, sl.stream) != hipSuccess ||
          hipMemcpyAsync(sl.d_start, sl.h_start, sizeof(uint32_t) * BL_COUNT * (n + 1), hipMemcpyHostToDevice,
                         sl.stream) != hipSuccess)
        e = rtgi_fail(RTGI_E_HIP, "descriptor upload failed");
    }
    if (!e) {
      plan.w = sl.d_desc;
      plan.queue = sl.d_queue;
      for (int l = 0; l < BL_COUNT; l++) plan.start[l] = sl.d_start + (size_t)l * (n + 1);
      e = plan_launch(plan, sl.stream, nullptr);
    }
    if (!e && hipEventRecord(sl.done, sl.stream) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "hipEventRecord");
    if (!e)
      for (rtgi_world* w : ws) w->ext_done = sl.done;
    lk.lock();
    sl.status = e;
    s->st.batches++;
    s->st.max_batch_seen = std::max(s->st.max_batch_seen, n);
    s->inflight.push_back(si);
    s->cv_done.notify_all();
  }
}

void reap_loop(rtgi_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_done.wait(lk, [&] { return !s->inflight.empty() || s->stop; });
    if (s->inflight.empty()) break;  // stop with nothing in flight
    int si = s->inflight.front();
    s->inflight.pop_front();
    Slot& sl = s->slots[si];
    std::vector<SrvWorld*> ws;
    for (const Req& q : sl.reqs) ws.push_back(s->worlds[q.r.world].get());
    lk.unlock();
    int e = sl.status;
    if (!e && hipEventSynchronize(sl.done) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "batched update failed");
    std::vector<rtgi_completion> cs;
    for (size_t i = 0; i < sl.reqs.size(); i++) {
      const rtgi_request& r = sl.reqs[i].r;
      rtgi_world* w = ws[i]->w;
      if (!e) w->ext_done = nullptr;  // complete: later waits on this world need not wait for the slot
      int re = e;
      if (!re && r.sh_out) re = rtgi_world_read_sh(w, r.sh_out);
      if (!re && r.counters_out) re = rtgi_world_read_counters(w, r.counters_out);
      cs.push_back(rtgi_completion{r.tag, r.world, re, (int32_t)sl.reqs.size(), sl.reqs[i].t_submit, sl.t_launch,
                                   s->now()});
    }
    lk.lock();
    for (SrvWorld* w : ws) w->busy = false;
    sl.busy = false;
    sl.reqs.clear();
    for (const rtgi_completion& c : cs) s->done.push_back(c);
    s->st.completed += cs.size();
    rtgi_completion_fn fn = s->fn;
    void* user = s->user;
    s->cv_sched.notify_all();
    s->cv_poll.notify_all();
    if (fn) {
      lk.unlock();
      for (const rtgi_completion& c : cs) fn(user, &c);
      lk.lock();
    }
  }
}

}  // namespace

extern "C" int rtgi_server_create(const rtgi_server_params* p, rtgi_server** out) {
  if (!out) return rtgi_fail(RTGI_E_INVALID, "rtgi_server_create: null argument");
  *out = nullptr;
  rtgi_server_params q = p ? *p : rtgi_server_params{};
  if (q.max_batch < 0 || q.slots < 0 || q.slots > 8) return rtgi_fail(RTGI_E_INVALID, "max_batch >= 0, slots in 0..8");
  if (q.max_batch == 0) q.max_batch = 64;
  if (q.slots == 0) q.slots = 2;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return rtgi_fail(RTGI_E_NODEVICE, "no HIP device visible");
  rtgi_server* s = new (std::nothrow) rtgi_server();
  if (!s) return rtgi_fail(RTGI_E_NOMEM, "out of host memory");
  (void)hipGetDevice(&s->device);
  if (q.hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return rtgi_fail(RTGI_E_HIP, "hipMemGetInfo failed");
    }
    q.hbm_budget = (uint64_t)(fr * 0.9);
  }
  s->p = q;
  s->st.budget = q.hbm_budget;
  s->t0 = Clock::now();
  s->slots.resize(q.slots);
  int e = RTGI_OK;
  for (Slot& sl : s->slots) {
    if (!e && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "hipStreamCreate");
    if (!e && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "hipEventCreate");
    if (!e && hipMalloc((void**)&sl.d_desc, sizeof(TraceK) * q.max_batch) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc");
    if (!e && hipMalloc((void**)&sl.d_start, sizeof(uint32_t) * BL_COUNT * (q.max_batch + 1)) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc");
    if (!e && hipMalloc((void**)&sl.d_queue, sizeof(uint32_t) * RTGI_QUEUE_WORDS) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc");
    if (!e && hipHostMalloc((void**)&sl.h_desc, sizeof(TraceK) * q.max_batch, hipHostMallocDefault) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipHostMalloc");
    if (!e && hipHostMalloc((void**)&sl.h_start, sizeof(uint32_t) * BL_COUNT * (q.max_batch + 1),
                            hipHostMallocDefault) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipHostMalloc");
  }
  if (e) {
    for (Slot& sl : s->slots) slot_free(sl);
    delete s;
    return e;
  }
  s->sched = std::thread(sched_loop, s);
  s->reaper = std::thread(reap_loop, s);
  *out = s;
  return RTGI_OK;
}

extern "C" int rtgi_server_add_world(rtgi_server* s, const rtgi_vox_params* vp, const rtgi_rc_params* rp,
                                     const float* pos, const float* nrm, uint32_t nverts, const uint32_t* tri,
                                     uint32_t ntris, const float* materials, uint32_t nmat, int32_t* id) {
  if (!s || !vp || !rp || !id) return rtgi_fail(RTGI_E_INVALID, "rtgi_server_add_world: null argument");
  if ((nverts && (!pos || !nrm)) || (ntris && !tri) || (nmat && !materials))
    return rtgi_fail(RTGI_E_INVALID, "rtgi_server_add_world: null array");
  for (uint32_t t = 0; t < ntris; t++) {
    if (tri[4 * t] >= nverts || tri[4 * t + 1] >= nverts || tri[4 * t + 2] >= nverts)
      return rtgi_fail(RTGI_E_INVALID, "triangle %u indexes a vertex out of range", t);
    if (tri[4 * t + 3] >= nmat) return rtgi_fail(RTGI_E_INVALID, "triangle %u indexes a material out of range", t);
  }
  uint64_t bytes = 0;
  int e = rtgi_world_estimate_bytes(vp, rp, nverts, ntris, nmat, &bytes);
  if (e) return e;
  if (bytes > s->p.hbm_budget)
    return rtgi_fail(RTGI_E_NOMEM, "world needs %llu bytes, over the whole budget of %llu", (unsigned long long)bytes,
                     (unsigned long long)s->p.hbm_budget);
  std::unique_ptr<SrvWorld> w(new (std::nothrow) SrvWorld());
  if (!w) return rtgi_fail(RTGI_E_NOMEM, "out of host memory");
  w->vp = *vp;
  w->rp = *rp;
  if (!w->pos.assign(pos, (size_t)nverts * 3) || !w->nrm.assign(nrm, (size_t)nverts * 3) ||
      !w->tri.assign(tri, (size_t)ntris * 4) || !w->mat.assign(materials, (size_t)nmat * 4))
    return rtgi_fail(RTGI_E_NOMEM, "hipHostMalloc for the scene copy failed");
  w->nverts = nverts;
  w->ntris = ntris;
  w->nmat = nmat;
  w->bytes = bytes;
  std::lock_guard<std::mutex> lk(s->mu);
  *id = (int32_t)s->worlds.size();
  s->worlds.push_back(std::move(w));
  s->st.worlds = (int32_t)s->worlds.size();
  return RTGI_OK;
}

extern "C" int rtgi_server_submit(rtgi_server* s, const rtgi_request* r) {
  if (!s || !r) return rtgi_fail(RTGI_E_INVALID, "rtgi_server_submit: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->world < 0 || r->world >= (int32_t)s->worlds.size()) return rtgi_fail(RTGI_E_INVALID, "unknown world id");
  {  // the frame check of rtgi_world_set_frame, so that a request's frame cannot fail its batch later
    int origin[RTGI_MAX_LEVELS][3], G0[3];
    const SrvWorld& sw = *s->worlds[r->world];
    int e = frame_origins(&sw.vp, &sw.rp, &r->frame, origin, G0);
    if (e) return e;
  }
  if (s->stop) return rtgi_fail(RTGI_E_STATE, "server is stopping");
  Req q;
  q.r = *r;
  q.t_submit = s->now();
  s->queue.push_back(q);
  s->worlds[r->world]->queued++;
  s->st.submitted++;
  s->cv_sched.notify_all();
  return RTGI_OK;
}

extern "C" int rtgi_server_poll(rtgi_server* s, rtgi_completion* out, int32_t max, int32_t wait_ms, int32_t* n) {
  if (!s || !n || (max > 0 && !out)) return rtgi_fail(RTGI_E_INVALID, "rtgi_server_poll: bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return !s->done.empty(); };
  if (wait_ms < 0)
    s->cv_poll.wait(lk, ready);
  else if (wait_ms > 0)
    s->cv_poll.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
  int k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return RTGI_OK;
}

extern "C" int rtgi_server_set_callback(rtgi_server* s, rtgi_completion_fn fn, void* user) {
  if (!s) return rtgi_fail(RTGI_E_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->fn = fn;
  s->user = user;
  return RTGI_OK;
}

extern "C" int rtgi_server_drain(rtgi_server* s) {
  if (!s) return rtgi_fail(RTGI_E_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_poll.wait(lk, [&] { return s->st.completed == s->st.submitted; });
  return RTGI_OK;
}

extern "C" int rtgi_server_get_stats(rtgi_server* s, rtgi_server_stats* out) {
  if (!s || !out) return rtgi_fail(RTGI_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->queued = (int32_t)s->queue.size();
  return RTGI_OK;
}

extern "C" int rtgi_server_world(rtgi_server* s, int32_t id, rtgi_world** out) {
  if (!s || !out) return rtgi_fail(RTGI_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (id < 0 || id >= (int32_t)s->worlds.size()) return rtgi_fail(RTGI_E_INVALID, "unknown world id");
  if (!s->worlds[id]->w) return rtgi_fail(RTGI_E_STATE, "world %d is not resident", id);
  *out = s->worlds[id]->w;
  return RTGI_OK;
}

extern "C" int rtgi_server_destroy(rtgi_server* s) {
  if (!s) return RTGI_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_poll.wait(lk, [&] { return s->st.completed == s->st.submitted; });
    s->stop = true;
    s->cv_sched.notify_all();
    s->cv_done.notify_all();
  }
  s->sched.join();
  s->reaper.join();
  (void)hipSetDevice(s->device);
  for (auto& w : s->worlds)
    if (w->w) rtgi_world_destroy(w->w);
  for (Slot& sl : s->slots) slot_free(sl);
  delete s;
  return RTGI_OK;
}
