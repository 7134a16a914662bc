// pm_server.hip -- the serving loop: depth-map jobs admitted under an HBM
// budget, source images cached on the device, and continuous batching of
// the jobs' phases into shared persistent launches.
//
// Threads.  The loop thread admits jobs, uploads them (on the upload stream)
// and issues one step after another on the compute stream, keeping at most
// two steps queued on the device.  The completion thread waits for a
// finished job's last launch, copies its maps to the caller's buffers (on
// the copy stream) and returns its memory.  `mu` guards the queues, the
// cache index and the statistics; device memory comes from one arena
// allocated at creation (the budget), so no hipMalloc/hipFree runs while
// jobs are in flight.
#include "pm_host.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <cstdlib>
#include <new>
#include <set>
#include <thread>

using namespace pmh;

namespace {

using Clock = std::chrono::steady_clock;
using CacheKey = std::pair<int32_t, int32_t>;  // (image cache_id, storage)

constexpr int kRing = 4;            // step table buffers

struct CacheEntry {
  size_t off = 0, bytes = 0;
  int32_t width = 0, height = 0;  // the image the entry holds: a key names one image
  int refs = 0;
  uint64_t last_use = 0;
};

enum JobState { kPending = 0, kActive = 1, kDone = 2 };

struct Job {
  uint64_t ticket = 0;
  pm_options opt{};
  pm_view_desc view{};
  pm_outputs out{};
  pm_problem* prob = nullptr;
  std::vector<std::vector<PhaseRec>> ticks;
  size_t next = 0;                    // next tick
  std::vector<CacheKey> keys;         // cache entries this job holds a reference on
  std::vector<const void*> src_tex;   // per source: cached quads or null
  std::vector<CacheKey> fresh;        // entries this job must upload (misses)
  size_t block_off = 0, block_bytes = 0;
  bool has_block = false;
  hipEvent_t done_ev = nullptr;
  JobState state = kPending;
  pm_status status = PM_OK;
  std::string error;
  Clock::time_point t_submit, t_admit, t_done;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct pm_server {
  pmh::EngineRef engine_ref;  // first member: released after the others
  pm_engine* engine = nullptr;
  pm_server_config cfg{};
  hipStream_t compute = nullptr, upload = nullptr, copy = nullptr;
  Arena arena;
  float* ws = nullptr;
  size_t ws_floats = 0;
  uint8_t* staging = nullptr;
  size_t staging_bytes = 0;
  void* table_dev[kRing] = {};
  size_t table_dev_bytes[kRing] = {};
  uint8_t* table_host[kRing] = {};
  size_t table_host_bytes[kRing] = {};
  hipEvent_t table_ev[kRing] = {};
  bool table_ev_used[kRing] = {};
  int ring = 0;
  hipEvent_t step_ev[2] = {};
  bool step_ev_used[2] = {};
  int step_parity = 0;

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job*> pending;
  std::vector<Job*> active;
  std::deque<Job*> completing;
  int copying = 0;  // jobs the completion thread took from `completing` and has not released yet
  std::map<uint64_t, Job*> jobs;  // submitted and not yet waited for
  std::map<CacheKey, CacheEntry> cache;
  uint64_t next_ticket = 1, use_clock = 0;
  pm_server_stats st{};
  size_t cache_bytes = 0;
  bool stop = false, loop_done = false;
  bool abandon = false;  // process exit: stop after the step in progress, leave jobs unfinished
  pm_status fatal = PM_OK;
  std::string fatal_msg;
  std::thread loop, completer;
};

namespace {

// ---- admission (under mu) ----------------------------------------------------

// Free the least recently used cache entry no job holds.
bool evict_one(pm_server* sv) {
  auto victim = sv->cache.end();
  for (auto it = sv->cache.begin(); it != sv->cache.end(); ++it) {
    if (it->second.refs != 0) continue;
    if (victim == sv->cache.end() || it->second.last_use < victim->second.last_use) victim = it;
  }
  if (victim == sv->cache.end()) return false;
  sv->arena.release(victim->second.off, victim->second.bytes);
  sv->cache_bytes -= victim->second.bytes;
  sv->cache.erase(victim);
  sv->st.cache_evictions++;
  return true;
}

bool alloc_evicting(pm_server* sv, size_t bytes, size_t* off) {
  for (;;) {
    if (sv->arena.alloc(bytes, off)) return true;
    if (!evict_one(sv)) return false;
  }
}

// A cache key names one image: a job whose source has a cached key but
// other dimensions (or that lists one key with two sizes) is an error, not
// a hit.  Empty when there is no conflict (under mu).
std::string key_conflict(const pm_server* sv, const pm_view_desc& v) {
  for (int i = 0; i < v.num_src; ++i) {
    const pm_image& a = v.src[i];
    if (a.cache_id <= 0) continue;
    for (int k = 0; k < i; ++k) {
      const pm_image& b = v.src[k];
      if (b.cache_id == a.cache_id && (b.width != a.width || b.height != a.height))
        return "sources " + std::to_string(k) + " and " + std::to_string(i) + " share cache_id " +
               std::to_string(a.cache_id) + " but differ in size";
    }
    for (const auto& kv : sv->cache) {
      if (kv.first.first != a.cache_id) continue;
      if (kv.second.width != a.width || kv.second.height != a.height)
        return "source " + std::to_string(i) + " (" + std::to_string(a.width) + "x" + std::to_string(a.height) +
               ") has cache_id " + std::to_string(a.cache_id) + ", which the server holds for a " +
               std::to_string(kv.second.width) + "x" + std::to_string(kv.second.height) + " image";
    }
  }
  return std::string();
}

// Reserve cache entries and the job's memory block.  Returns 1 when
// reserved, 0 (holding nothing) when the job does not fit now, -1 (holding
// nothing, *why set) when a source's key names another image.
int reserve(pm_server* sv, Job* j, std::string* why) {
  const int N = j->view.num_src;
  const int storage = j->opt.image_storage;
  j->keys.clear();
  j->fresh.clear();
  j->src_tex.assign(N, nullptr);
  *why = key_conflict(sv, j->view);
  if (!why->empty()) return -1;
  std::vector<bool> cached(N, false);
  // hits first, so eviction below never takes an entry this job needs
  for (int i = 0; i < N; ++i) {
    const int32_t id = j->view.src[i].cache_id;
    if (id <= 0) continue;
    const CacheKey k{id, storage};
    auto it = sv->cache.find(k);
    if (it == sv->cache.end()) continue;
    if (std::find(j->keys.begin(), j->keys.end(), k) == j->keys.end()) {
      it->second.refs++;
      j->keys.push_back(k);
    }
    j->src_tex[i] = sv->arena.base() + it->second.off;
    cached[i] = true;
  }
  bool ok = true;
  for (int i = 0; i < N && ok; ++i) {
    const int32_t id = j->view.src[i].cache_id;
    if (id <= 0 || cached[i]) continue;
    const CacheKey k{id, storage};
    auto it = sv->cache.find(k);
    if (it != sv->cache.end()) {  // a source listed twice
      j->src_tex[i] = sv->arena.base() + it->second.off;
      cached[i] = true;
      continue;
    }
    CacheEntry e;
    e.bytes = src_quad_bytes(j->view.src[i], storage);
    e.width = j->view.src[i].width;
    e.height = j->view.src[i].height;
    if (!alloc_evicting(sv, e.bytes, &e.off)) {
      ok = false;
      break;
    }
    e.refs = 1;
    sv->cache[k] = e;
    sv->cache_bytes += e.bytes;
    j->keys.push_back(k);
    j->fresh.push_back(k);
    j->src_tex[i] = sv->arena.base() + e.off;
    cached[i] = true;
  }
  if (ok) {
    std::unique_ptr<bool[]> c(new bool[N > 0 ? N : 1]);
    for (int i = 0; i < N; ++i) c[i] = cached[i];
    j->block_bytes = job_block_bytes(&j->opt, &j->view, c.get());
    ok = alloc_evicting(sv, j->block_bytes, &j->block_off);
    j->has_block = ok;
  }
  if (!ok) {
    // roll back: drop the entries created for this job, release the hits
    for (const CacheKey& k : j->keys) {
      auto it = sv->cache.find(k);
      if (it == sv->cache.end()) continue;
      it->second.refs--;
      if (std::find(j->fresh.begin(), j->fresh.end(), k) != j->fresh.end()) {
        sv->arena.release(it->second.off, it->second.bytes);
        sv->cache_bytes -= it->second.bytes;
        sv->cache.erase(it);
      }
    }
    j->keys.clear();
    j->fresh.clear();
    return 0;
  }
  const uint64_t now = ++sv->use_clock;
  for (const CacheKey& k : j->keys) sv->cache[k].last_use = now;
  sv->st.cache_misses += j->fresh.size();
  int hits = 0;
  for (int i = 0; i < N; ++i)
    if (j->view.src[i].cache_id > 0) hits++;
  sv->st.cache_hits += (uint64_t)std::max(0, hits - (int)j->fresh.size());
  return 1;
}

// Give back everything job j holds (under mu).  drop_fresh: the job's own
// uploads may be incomplete, so the cache entries it created are removed.
void release(pm_server* sv, Job* j, bool drop_fresh) {
  if (j->has_block) sv->arena.release(j->block_off, j->block_bytes);
  j->has_block = false;
  const uint64_t now = ++sv->use_clock;
  for (const CacheKey& k : j->keys) {
    auto it = sv->cache.find(k);
    if (it == sv->cache.end()) continue;
    it->second.refs--;
    it->second.last_use = now;
    if (drop_fresh && it->second.refs == 0 &&
        std::find(j->fresh.begin(), j->fresh.end(), k) != j->fresh.end()) {
      sv->arena.release(it->second.off, it->second.bytes);
      sv->cache_bytes -= it->second.bytes;
      sv->cache.erase(it);
    }
  }
  j->keys.clear();
  j->fresh.clear();
}

// Device bytes a job needs when none of its sources is cached: more than
// the arena means it can never run.
// Arena bytes a job needs with nothing cached: its block with every source
// in it (a source that goes to the cache instead takes the same rounded
// bytes as a cache entry).
size_t worst_case_bytes(const pm_options* o, const pm_view_desc* v) { return job_block_bytes(o, v, nullptr); }

// ---- loop thread ---------------------------------------------------------------

pm_status ensure_staging(pm_server* sv, size_t bytes) {
  if (bytes <= sv->staging_bytes) return PM_OK;
  PM_HIP(hipStreamSynchronize(sv->upload));
  if (sv->staging) PM_HIP(hipFree(sv->staging));
  sv->staging = nullptr;
  sv->staging_bytes = 0;
  PM_HIP(hipMalloc(&sv->staging, bytes));
  sv->staging_bytes = bytes;
  return PM_OK;
}

// Upload the job's cache misses and create its problem in its block.
pm_status upload_job(pm_server* sv, Job* j) {
  size_t need = 0;
  for (int i = 0; i < j->view.num_src; ++i) {
    const size_t n = (size_t)j->view.src[i].width * j->view.src[i].height;
    need = std::max(need, j->opt.geom_consistency ? n * sizeof(float) : n);
  }
  pm_status rs = ensure_staging(sv, need);
  if (rs != PM_OK) return rs;
  for (const CacheKey& k : j->fresh) {
    for (int i = 0; i < j->view.num_src; ++i) {
      if (j->view.src[i].cache_id != k.first) continue;
      if ((rs = upload_quads(j->view.src[i], k.second, (void*)j->src_tex[i], sv->staging, sv->upload)) != PM_OK)
        return rs;
      break;
    }
  }
  CreateEnv env;
  env.stream = sv->upload;
  env.arena = &sv->arena;
  env.block_off = j->block_off;
  env.block_bytes = j->block_bytes;
  env.src_tex = j->src_tex.data();
  env.staging = sv->staging;
  env.staging_bytes = sv->staging_bytes;
  if ((rs = create_problem(sv->engine, &j->opt, &j->view, env, &j->prob)) != PM_OK) return rs;
  // initial state of views given one (what prepare_views does per run)
  if (!j->opt.init_depth_random) {
    const size_t plane = (size_t)j->prob->W * j->prob->H;
    PM_HIP(hipMemcpyAsync(j->prob->host.depth, j->prob->init_depth, plane * sizeof(float), hipMemcpyDeviceToDevice,
                          sv->upload));
    PM_HIP(hipMemcpyAsync(j->prob->host.normal, j->prob->init_normal, 3 * plane * sizeof(float),
                          hipMemcpyDeviceToDevice, sv->upload));
    PM_HIP(hipStreamSynchronize(sv->upload));
  }
  j->ticks = view_ticks(*j->prob);
  PM_HIP(hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming));
  return PM_OK;
}

// Host buffer and device buffer of ring slot k with room for `bytes`.
pm_status ring_slot(pm_server* sv, int k, size_t bytes) {
  if (sv->table_ev_used[k]) PM_HIP(hipEventSynchronize(sv->table_ev[k]));
  if (bytes > sv->table_host_bytes[k]) {
    if (sv->table_host[k]) PM_HIP(hipHostFree(sv->table_host[k]));
    sv->table_host[k] = nullptr;
    const size_t nb = std::max(bytes * 2, (size_t)1 << 16);
    PM_HIP(hipHostMalloc((void**)&sv->table_host[k], nb, hipHostMallocDefault));
    sv->table_host_bytes[k] = nb;
  }
  if (bytes > sv->table_dev_bytes[k]) {
    PM_HIP(hipStreamSynchronize(sv->compute));  // no queued launch reads the old buffer
    if (sv->table_dev[k]) PM_HIP(hipFree(sv->table_dev[k]));
    sv->table_dev[k] = nullptr;
    const size_t nb = std::max(bytes * 2, (size_t)1 << 16);
    PM_HIP(hipMalloc(&sv->table_dev[k], nb));
    sv->table_dev_bytes[k] = nb;
  }
  return PM_OK;
}

// One step: every active job advances by one phase.  Returns the jobs whose
// last phase was issued.
pm_status issue_step(pm_server* sv, std::vector<Job*>& finished) {
  std::vector<std::pair<pm_problem*, const std::vector<PhaseRec>*>> phases;
  for (Job* j : sv->active) phases.emplace_back(j->prob, &j->ticks[j->next]);
  TickPlan plan;
  pm_status rs = plan_launches(phases, plan, nullptr);
  if (rs != PM_OK) return rs;
  if (plan.ws_floats > sv->ws_floats) return fail(PM_ERR_INTERNAL, "step needs more workspace than reserved");
  const size_t qoff = (plan.tables.size() + 255) & ~(size_t)255;
  const size_t bytes = qoff + 8 * sizeof(uint32_t) * plan.launches.size();
  const int k = sv->ring;
  sv->ring = (sv->ring + 1) % kRing;
  if ((rs = ring_slot(sv, k, bytes)) != PM_OK) return rs;
  uint8_t* hb = sv->table_host[k];
  uint8_t* db = (uint8_t*)sv->table_dev[k];
  std::memcpy(hb, plan.tables.data(), plan.tables.size());
  std::memset(hb + plan.tables.size(), 0, bytes - plan.tables.size());  // queue counters start at zero
  for (size_t l = 0; l < plan.launches.size(); ++l) {
    pm::LaunchTable* lt = (pm::LaunchTable*)(hb + plan.launches[l].offset);
    lt->ws = sv->ws;
    lt->queue = (uint32_t*)(db + qoff) + 8 * l;
  }
  PM_HIP(hipMemcpyAsync(db, hb, bytes, hipMemcpyHostToDevice, sv->compute));
  PM_HIP(hipEventRecord(sv->table_ev[k], sv->compute));
  sv->table_ev_used[k] = true;
  for (const pm_schedule::Launch& l : plan.launches)
    PM_HIP(pm::launch_phase(l.kind, l.cfg, (const pm::LaunchTable*)(db + l.offset), l.grid, sv->compute));
  for (Job* j : sv->active) {
    if (++j->next == j->ticks.size()) {
      PM_HIP(hipEventRecord(j->done_ev, sv->compute));
      finished.push_back(j);
    }
  }
  // at most two steps queued on the device
  const int par = sv->step_parity;
  PM_HIP(hipEventRecord(sv->step_ev[par], sv->compute));
  sv->step_ev_used[par] = true;
  sv->step_parity ^= 1;
  if (sv->step_ev_used[par ^ 1]) PM_HIP(hipEventSynchronize(sv->step_ev[par ^ 1]));
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->st.steps++;
  sv->st.launches += plan.launches.size();
  return PM_OK;
}

void fail_job(pm_server* sv, Job* j, pm_status s, const std::string& msg) {
  // under mu
  release(sv, j, true);
  j->status = s;
  j->error = msg;
  j->state = kDone;
  j->t_done = Clock::now();
  sv->st.failed++;
}

void loop_main(pm_server* sv) {
  (void)hipSetDevice(sv->engine->device);
  for (;;) {
    Job* admit = nullptr;
    bool quit = false;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        if (sv->abandon || (sv->stop && sv->pending.empty() && sv->active.empty())) {
          quit = true;
          break;
        }
        if (!sv->pending.empty() && (int)sv->active.size() < sv->cfg.max_active_views) {
          Job* j = sv->pending.front();
          std::string why;
          const int r = reserve(sv, j, &why);
          if (r > 0) {
            sv->pending.pop_front();
            admit = j;
            break;
          }
          if (r < 0) {
            sv->pending.pop_front();
            fail_job(sv, j, PM_ERR_INVALID_ARG, why);
            sv->cv.notify_all();
            continue;
          }
          if (sv->active.empty()) {
            if (sv->completing.empty() && sv->copying == 0) {
              // nothing runs, yet the job does not fit: the cached sources
              // it would reuse, pinned where they lie, can split the free
              // space into holes smaller than its block.  Drop every cached
              // source (none is held now) so the arena is one hole; laid out
              // from its start, the job's entries and block take exactly
              // the bytes the submit check allowed.
              while (evict_one(sv)) {
              }
              if (reserve(sv, j, &why) > 0) {
                sv->pending.pop_front();
                admit = j;
                break;
              }
              // still no room (cannot happen for a job that passed the
              // submit check): fail it rather than wait forever
              sv->pending.pop_front();
              int held = 0;
              for (const auto& c : sv->cache) held += c.second.refs != 0;
              fail_job(sv, j, PM_ERR_OVER_BUDGET,
                       "job does not fit in the budget: needs " + std::to_string(worst_case_bytes(&j->opt, &j->view)) +
                           " bytes; arena " + std::to_string(sv->arena.used()) + " of " +
                           std::to_string(sv->arena.capacity()) + " used, largest hole " +
                           std::to_string(sv->arena.largest_hole()) + ", " + std::to_string(sv->cache.size()) +
                           " cached sources (" + std::to_string(held) + " held)");
              sv->cv.notify_all();
              continue;
            }
            sv->cv.wait(lk);  // a finishing job will give memory back
            continue;
          }
        }
        if (!sv->active.empty()) break;  // issue a step
        sv->cv.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
      }
    }
    if (quit) break;
    if (admit) {
      const pm_status rs = upload_job(sv, admit);
      std::lock_guard<std::mutex> lk(sv->mu);
      if (rs != PM_OK) {
        delete admit->prob;
        admit->prob = nullptr;
        fail_job(sv, admit, rs, last_error());
        sv->cv.notify_all();
      } else {
        admit->state = kActive;
        admit->t_admit = Clock::now();
        sv->active.push_back(admit);
        sv->st.max_active_seen = std::max(sv->st.max_active_seen, (int32_t)sv->active.size());
      }
      continue;  // admit as many as fit before the next step
    }
    if (sv->active.empty()) continue;
    std::vector<Job*> finished;
    const pm_status rs = issue_step(sv, finished);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (rs != PM_OK) {
      // a device error ends the server: every job not yet done fails
      sv->fatal = rs;
      sv->fatal_msg = last_error();
      for (Job* j : sv->active) {
        delete j->prob;
        j->prob = nullptr;
        fail_job(sv, j, rs, sv->fatal_msg);
      }
      sv->active.clear();
      for (Job* j : sv->pending) fail_job(sv, j, rs, sv->fatal_msg);
      sv->pending.clear();
      sv->cv.notify_all();
      continue;
    }
    for (Job* j : finished) {
      sv->active.erase(std::find(sv->active.begin(), sv->active.end(), j));
      sv->completing.push_back(j);
    }
    if (!finished.empty()) sv->cv.notify_all();
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->loop_done = true;
  sv->cv.notify_all();
}

pm_status copy_outputs(pm_server* sv, Job* j) {
  PM_HIP(hipEventSynchronize(j->done_ev));
  const pm_problem* p = j->prob;
  const size_t plane = (size_t)p->W * p->H;
  hipStream_t s = sv->copy;
  if (j->out.depth) PM_HIP(hipMemcpyAsync(j->out.depth, p->host.depth, plane * sizeof(float), hipMemcpyDeviceToHost, s));
  if (j->out.normal)
    PM_HIP(hipMemcpyAsync(j->out.normal, p->host.normal, 3 * plane * sizeof(float), hipMemcpyDeviceToHost, s));
  if (j->out.conf) PM_HIP(hipMemcpyAsync(j->out.conf, p->host.conf, plane * sizeof(float), hipMemcpyDeviceToHost, s));
  if (j->out.cost)
    PM_HIP(hipMemcpyAsync(j->out.cost, p->host.cost, (size_t)p->N * plane * sizeof(float), hipMemcpyDeviceToHost, s));
  if (j->out.view_sel)
    PM_HIP(hipMemcpyAsync(j->out.view_sel, p->host.view_sel, plane * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (j->out.curvature && p->host.curv)
    PM_HIP(hipMemcpyAsync(j->out.curvature, p->host.curv, 2 * plane * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
  PM_HIP(hipStreamSynchronize(s));
  return PM_OK;
}

void completer_main(pm_server* sv) {
  (void)hipSetDevice(sv->engine->device);
  for (;;) {
    Job* j = nullptr;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv.wait(lk, [&] { return !sv->completing.empty() || sv->loop_done; });
      if (sv->completing.empty() || (sv->abandon && sv->loop_done)) break;
      j = sv->completing.front();
      sv->completing.pop_front();
      sv->copying++;  // still holds its memory until released below
    }
    const pm_status rs = copy_outputs(sv, j);
    const std::string msg = rs == PM_OK ? std::string() : last_error();
    std::lock_guard<std::mutex> lk(sv->mu);
    delete j->prob;  // arena-backed: frees nothing on the device
    j->prob = nullptr;
    if (rs != PM_OK) {
      fail_job(sv, j, rs, msg);
    } else {
      release(sv, j, false);
      j->status = PM_OK;
      j->state = kDone;
      j->t_done = Clock::now();
      sv->st.completed++;
    }
    sv->copying--;
    sv->cv.notify_all();
  }
}

void destroy_job(Job* j) {
  if (j->done_ev) (void)hipEventDestroy(j->done_ev);
  delete j->prob;
  delete j;
}

// Servers alive at process exit.  A caller may exit without destroying a
// server (CPython, for one, does not finalise every object still alive at
// exit).  Its loop thread would then go on issuing launches and its
// completion thread waiting on events while exit() runs the HIP runtime's
// static destructors, which tear down the streams and events those calls
// use.  The first server registers stop_live_servers with atexit after the
// runtime is initialised (pm_server_create has called into it), and atexit
// handlers run in reverse order of registration, so the handler runs before
// the runtime's teardown.  It stops each live server's threads after the
// step in progress and waits for the device work already issued; the jobs
// left are never finished, and nothing is freed (the process is ending).
// A pm_server_destroy after that makes no HIP call.
struct LiveServers {
  std::mutex mu;
  std::set<pm_server*> servers;
  bool registered = false, exiting = false;
};

LiveServers& live_servers() {
  static LiveServers* l = new LiveServers();  // never destroyed: used from atexit
  return *l;
}

void stop_live_servers() {
  LiveServers& l = live_servers();
  std::set<pm_server*> servers;
  {
    std::lock_guard<std::mutex> lk(l.mu);
    l.exiting = true;
    servers = l.servers;
  }
  for (pm_server* sv : servers) {
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->abandon = true;
      sv->stop = true;
      sv->cv.notify_all();
    }
    if (sv->loop.joinable()) sv->loop.join();
    if (sv->completer.joinable()) sv->completer.join();
    (void)hipSetDevice(sv->engine->device);
    (void)hipStreamSynchronize(sv->compute);
    (void)hipStreamSynchronize(sv->upload);
    (void)hipStreamSynchronize(sv->copy);
  }
}

}  // namespace

PM_API pm_status pm_server_create(pm_engine* e, const pm_server_config* cfg, pm_server** out) {
  if (!e || !cfg || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (cfg->max_active_views < 1) return fail(PM_ERR_INVALID_ARG, "max_active_views must be >= 1");
  PM_HIP(hipSetDevice(e->device));
  int n_cu = 0;
  PM_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, e->device));
  const size_t ws_floats = (size_t)n_cu * pm::kMaxBlocksPerCU * pm::kSweepThreads * pm::ws_floats_per_slot(PM_MAX_NUM_SRC);
  const size_t ws_bytes = Arena::round(ws_floats * sizeof(float));
  if (cfg->hbm_budget_bytes < ws_bytes + ((size_t)1 << 20))
    return fail(PM_ERR_OVER_BUDGET, "budget below the workspace reserve (" + std::to_string(ws_bytes) + " bytes)");
  std::unique_ptr<pm_server> sv(new (std::nothrow) pm_server());
  if (!sv) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  sv->engine_ref.hold(e);
  sv->engine = e;
  sv->cfg = *cfg;
  PM_HIP(hipStreamCreateWithFlags(&sv->compute, hipStreamNonBlocking));
  PM_HIP(hipStreamCreateWithFlags(&sv->upload, hipStreamNonBlocking));
  PM_HIP(hipStreamCreateWithFlags(&sv->copy, hipStreamNonBlocking));
  for (int k = 0; k < kRing; ++k) PM_HIP(hipEventCreateWithFlags(&sv->table_ev[k], hipEventDisableTiming));
  for (int k = 0; k < 2; ++k) PM_HIP(hipEventCreateWithFlags(&sv->step_ev[k], hipEventDisableTiming));
  PM_HIP(hipMalloc(&sv->ws, ws_bytes));
  sv->ws_floats = ws_floats;
  pm_status rs = sv->arena.init(cfg->hbm_budget_bytes - ws_bytes);
  if (rs != PM_OK) return rs;
  pm_server* p = sv.release();
  p->st.arena_capacity = p->arena.capacity();
  {
    LiveServers& l = live_servers();
    std::lock_guard<std::mutex> lk(l.mu);
    if (!l.registered) {
      l.registered = true;
      std::atexit(stop_live_servers);
    }
    l.servers.insert(p);
  }
  p->loop = std::thread(loop_main, p);
  p->completer = std::thread(completer_main, p);
  *out = p;
  return PM_OK;
}

PM_API pm_status pm_server_submit(pm_server* sv, const pm_options* o, const pm_view_desc* v, const pm_outputs* outp,
                                  uint64_t* ticket) {
  if (!sv || !o || !v || !outp || !ticket) return fail(PM_ERR_INVALID_ARG, "null argument");
  pm_status rs = pm_options_check(o);
  if (rs != PM_OK) return rs;
  if ((rs = check_view(o, v)) != PM_OK) return rs;
  const size_t worst = worst_case_bytes(o, v);
  if (worst > sv->arena.capacity())
    return fail(PM_ERR_OVER_BUDGET, "job needs " + std::to_string(worst) + " bytes, budget holds " +
                                        std::to_string(sv->arena.capacity()));
  Job* j = new (std::nothrow) Job();
  if (!j) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  j->opt = *o;
  j->view = *v;
  j->out = *outp;
  j->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->stop) {
    delete j;
    return fail(PM_ERR_NOT_READY, "server is shutting down");
  }
  if (sv->fatal != PM_OK) {
    delete j;
    return fail(sv->fatal, "server stopped after a device error: " + sv->fatal_msg);
  }
  const std::string conflict = key_conflict(sv, *v);
  if (!conflict.empty()) {
    delete j;
    return fail(PM_ERR_INVALID_ARG, conflict);
  }
  j->ticket = sv->next_ticket++;
  sv->jobs[j->ticket] = j;
  sv->pending.push_back(j);
  sv->st.submitted++;
  *ticket = j->ticket;
  sv->cv.notify_all();
  return PM_OK;
}

PM_API pm_status pm_server_poll(pm_server* sv, uint64_t ticket, int32_t* done) {
  if (!sv || !done) return fail(PM_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end()) return fail(PM_ERR_INVALID_ARG, "unknown ticket");
  *done = it->second->state == kDone ? 1 : 0;
  return PM_OK;
}

PM_API pm_status pm_server_wait(pm_server* sv, uint64_t ticket, pm_job_stats* stats) {
  if (!sv) return fail(PM_ERR_INVALID_ARG, "null server");
  Job* j = nullptr;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(ticket);
    if (it == sv->jobs.end()) return fail(PM_ERR_INVALID_ARG, "unknown ticket");
    j = it->second;
    sv->cv.wait(lk, [&] { return j->state == kDone; });
    sv->jobs.erase(it);
  }
  const pm_status s = j->status;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->queue_ms = j->t_admit.time_since_epoch().count() ? ms_between(j->t_submit, j->t_admit) : 0.0;
    stats->latency_ms = ms_between(j->t_submit, j->t_done);
    stats->status = s;
  }
  const std::string msg = j->error;
  destroy_job(j);
  return s == PM_OK ? PM_OK : fail(s, msg);
}

PM_API pm_status pm_server_get_stats(pm_server* sv, pm_server_stats* out) {
  if (!sv || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *out = sv->st;
  out->arena_used = sv->arena.used();
  out->cache_bytes = sv->cache_bytes;
  out->active = (int32_t)sv->active.size();
  out->pending = (int32_t)sv->pending.size();
  return PM_OK;
}

PM_API void pm_server_destroy(pm_server* sv) {
  if (!sv) return;
  {
    LiveServers& l = live_servers();
    std::lock_guard<std::mutex> lk(l.mu);
    if (l.exiting) return;  // stopped by stop_live_servers; the runtime may be gone
    l.servers.erase(sv);
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv.notify_all();
  }
  if (sv->loop.joinable()) sv->loop.join();
  if (sv->completer.joinable()) sv->completer.join();
  (void)hipSetDevice(sv->engine->device);
  (void)hipStreamSynchronize(sv->compute);
  for (auto& kv : sv->jobs) destroy_job(kv.second);
  for (int k = 0; k < kRing; ++k) {
    if (sv->table_host[k]) (void)hipHostFree(sv->table_host[k]);
    if (sv->table_dev[k]) (void)hipFree(sv->table_dev[k]);
    if (sv->table_ev[k]) (void)hipEventDestroy(sv->table_ev[k]);
  }
  for (int k = 0; k < 2; ++k)
    if (sv->step_ev[k]) (void)hipEventDestroy(sv->step_ev[k]);
  if (sv->ws) (void)hipFree(sv->ws);
  if (sv->staging) (void)hipFree(sv->staging);
  (void)hipStreamDestroy(sv->compute);
  (void)hipStreamDestroy(sv->upload);
  (void)hipStreamDestroy(sv->copy);
  delete sv;
}
