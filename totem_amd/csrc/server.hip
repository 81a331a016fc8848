// Served STEM engine: continuous batching of slice steps across jobs.
//
// Jobs (totem_stem_job, every array already in HBM) are submitted at any time.
// Each totem_server_step advances every in-flight probe by one slice.  Probes of
// all jobs with the same grid side share one rows launch and one columns launch
// (the kernels read each item's own job, window and slice = step - start_step),
// and every grid side runs on a stream of its own so that launches of different
// sides overlap on the device.  A job's probes are admitted into wave slots as
// slots free up, between steps, so they join launches already in progress
// ("cohorts": the probes of one job admitted at one step advance together).
//
// Wave slots (wave [N][N] complex64 + FP64 detector partials) live in one arena
// split into pages of the largest slot size; a page serves one grid side at a
// time and returns to the shared pool when its last slot is released.  Admission
// is first come, first served over all grid sides; a job holds at most half the
// pages while jobs submitted after it started wait (a long job does not stall
// later arrivals; jobs queued together still run in order); when a waiting job
// with nothing in flight cannot get a single slot, the jobs behind it stop taking
// new work until pages drain to it, so no job starves.
//
// Results do not depend on what else is in flight: every probe's arithmetic and
// its detector reduction are the same as in a run of its job alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "host_common.h"
#include "kernels.h"
#include "launch.h"

namespace totem {
namespace srv {

#ifndef TOTEM_SERVER_AHEAD
#define TOTEM_SERVER_AHEAD 4
#endif
#ifndef TOTEM_SERVER_SHARE
#define TOTEM_SERVER_SHARE 2  // a job holds at most 1/TOTEM_SERVER_SHARE of the pages while others wait
#endif
constexpr int kMaxJobs = 1024;     // job-table entries (jobs in flight or waiting)
constexpr int kMaxSlices = 4096;   // slices per job
constexpr int kMaxSide = 2048;     // largest grid side (sets the page size)

struct Cohort {
  long long job;
  int start;                 // group step of the cohort's slice 0
  long first;                // first item index (m * D + d order) of the job
  int count;
  int offset;                // position in the group's device item list
  std::vector<int> slots;
};

struct Job {
  totem_stem_job j;
  long long id;
  int table;
  long total = 0, next = 0, done = 0;
  long long started = 0;     // next job id when admission first considered the job (0: not yet)
  std::vector<int> mos;
  int band_ng = 0;           // column groups of the job's band (the columns pass runs these)
  hipEvent_t ev = nullptr;   // recorded after the job's last reduction
  bool closed = false;
};

struct Group {
  int n = 0;
  int plan = 0;                       // pass variant of this side, fixed when the side is created
  hipStream_t st = nullptr;
  int step = 0;
  size_t slot_bytes = 0, wave_bytes = 0;
  int per_page = 0;
  std::vector<int> free_slots;        // page * per_page + k
  std::map<int, int> page_used;       // owned page -> slots in use
  std::vector<Cohort> cohorts;
  std::vector<ItemDev> host_items;
  DevBuf items_dev;
  static constexpr int kAhead = TOTEM_SERVER_AHEAD;  // slice steps a side may run ahead of its device work
  hipEvent_t ahead_ev[kAhead] = {};
  bool ahead_set[kAhead] = {};
  static constexpr int kRing = 3;     // pinned staging buffers of the item list
  ItemDev* stage[kRing] = {nullptr, nullptr, nullptr};
  size_t stage_cap[kRing] = {0, 0, 0};
  hipEvent_t stage_ev[kRing] = {nullptr, nullptr, nullptr};
  bool stage_used[kRing] = {false, false, false};
  int ring = 0;
  bool dirty = false;
};

}  // namespace srv
}  // namespace totem

// The opaque handle of the C ABI.
struct totem_server {
  int dev = 0;
  int max_aper = 0, max_mid = 0;
  size_t page_bytes = 0;
  int n_pages = 0;
  totem::DevBuf arena, jobs_dev, meta_dev, col_flags;
  std::vector<int> free_pages;
  std::vector<hipEvent_t> page_ev;    // recorded on the releasing stream when a page returns to the pool
  std::vector<bool> page_ev_set;
  std::vector<int> free_tables;
  std::map<long long, totem::srv::Job> jobs;
  std::deque<long long> order;        // submission order of jobs with items not yet admitted
  std::map<int, totem::srv::Group> groups;
  long long next_id = 1;

  size_t slot_bytes(int n, size_t* wave) const {
    const size_t w = totem::align_up((size_t)n * n * sizeof(float2));
    if (wave) *wave = w;
    return w + totem::align_up((size_t)(max_mid + 1) * max_aper * totem::detector_groups_max(n) * sizeof(double));
  }
  int meta_ints() const { return totem::srv::kMaxSlices + std::max(max_mid, 1); }
};

namespace totem {
namespace srv {

using Server = ::totem_server;

Group& group_of(Server& s, int n) {
  auto it = s.groups.find(n);
  if (it != s.groups.end()) return it->second;
  Group& g = s.groups[n];
  g.n = n;
  g.plan = current_plan(n);
  g.slot_bytes = s.slot_bytes(n, &g.wave_bytes);
  g.per_page = (int)(s.page_bytes / g.slot_bytes);
  HIP_OK(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking));
  for (int r = 0; r < Group::kRing; ++r) HIP_OK(hipEventCreateWithFlags(&g.stage_ev[r], hipEventDisableTiming));
  for (int r = 0; r < Group::kAhead; ++r) HIP_OK(hipEventCreateWithFlags(&g.ahead_ev[r], hipEventDisableTiming));
  tw_table(n);
  return g;
}

void take_page(Server& s, Group& g) {
  const int page = s.free_pages.back();
  s.free_pages.pop_back();
  if (s.page_ev_set[page]) HIP_OK(hipStreamWaitEvent(g.st, s.page_ev[page], 0));  // last user's work first
  g.page_used[page] = 0;
  for (int k = g.per_page - 1; k >= 0; --k) g.free_slots.push_back(page * g.per_page + k);
}

void release_slot(Server& s, Group& g, int slot) {
  const int page = slot / g.per_page;
  g.free_slots.push_back(slot);
  if (--g.page_used[page] > 0) return;
  // the page is empty: hand it back to the pool, ordered after this stream's work
  g.page_used.erase(page);
  g.free_slots.erase(std::remove_if(g.free_slots.begin(), g.free_slots.end(),
                                    [&](int x) { return x / g.per_page == page; }),
                     g.free_slots.end());
  HIP_OK(hipEventRecord(s.page_ev[page], g.st));
  s.page_ev_set[page] = true;
  s.free_pages.push_back(page);
}

// Admits waiting jobs' probes, oldest job first.  While other jobs wait, a job holds
// at most half the arena's pages (in slots of its side), so one large job cannot
// occupy the whole arena for its full run; a job that reached that share lets the
// jobs behind it in.  A job with no probe in flight that gets no slot stops the
// pass: the jobs behind it wait until pages drain to it, so no job starves.
void admit(Server& s) {
  for (auto it = s.order.begin(); it != s.order.end();) {
    Job& J = s.jobs.at(*it);
    Group& g = group_of(s, J.j.n);
    const long held = J.next - J.done;  // probes of J in flight
    long want = J.total - J.next;
    if (J.started == 0) J.started = s.next_id;
    // jobs submitted after J started (ids only grow, so the last one waiting tells)
    // hold J to its share of the pages; jobs that were queued before J started wait
    // their turn, first come first served
    if (TOTEM_SERVER_SHARE > 1 && s.order.size() > 1 && s.order.back() >= J.started) {
      const long share = (long)std::max(1, s.n_pages / TOTEM_SERVER_SHARE) * g.per_page;
      want = std::min(want, share - held);
    }
    if (want <= 0) {  // at its share: the jobs behind it go next
      ++it;
      continue;
    }
    while ((long)g.free_slots.size() < want && !s.free_pages.empty()) take_page(s, g);
    const int k = (int)std::min<long>(want, (long)g.free_slots.size());
    if (k == 0) {
      if (held == 0) return;  // starved: pages drain to it before anyone behind it is admitted
      ++it;
      continue;
    }
    Cohort c;
    c.job = J.id;
    c.start = g.step;
    c.first = J.next;
    c.count = k;
    c.offset = 0;
    c.slots.reserve(k);
    for (int q = 0; q < k; ++q) {
      const int slot = g.free_slots.back();
      g.free_slots.pop_back();
      ++g.page_used[slot / g.per_page];
      c.slots.push_back(slot);
    }
    J.next += k;
    g.cohorts.push_back(std::move(c));
    g.dirty = true;
    if (J.next == J.total) it = s.order.erase(it);
    else ++it;
  }
}

void upload_items(Server& s, Group& g) {
  g.host_items.clear();
  const char* base = static_cast<const char*>(s.arena.p);
  for (Cohort& c : g.cohorts) {
    c.offset = (int)g.host_items.size();
    const Job& J = s.jobs.at(c.job);
    const totem_stem_job& j = J.j;
    const float2* probes = static_cast<const float2*>(j.probes);
    for (int q = 0; q < c.count; ++q) {
      const long o = c.first + q;
      const int m = (int)(o / j.depth), d = (int)(o % j.depth);
      const int slot = c.slots[q];
      char* mem = const_cast<char*>(base) + (size_t)(slot / g.per_page) * s.page_bytes +
                  (size_t)(slot % g.per_page) * g.slot_bytes;
      ItemDev it;
      it.probe = probes + (size_t)m * j.n * j.n;
      it.wave = reinterpret_cast<float2*>(mem);
      it.part = reinterpret_cast<double*>(mem + g.wave_bytes);
      it.job = J.table;
      window_origin(j, d, it.r0, it.c0);
      it.start_step = c.start;
      it.out = (int)o;
      g.host_items.push_back(it);
    }
  }
  const size_t cnt = g.host_items.size();
  if (cnt == 0) return;
  if (g.items_dev.bytes < cnt * sizeof(ItemDev)) {
    HIP_OK(hipStreamSynchronize(g.st));
    g.items_dev.alloc(std::max(cnt, (size_t)256) * 2 * sizeof(ItemDev));
  }
  const int r = g.ring;
  g.ring = (g.ring + 1) % Group::kRing;
  if (g.stage_used[r]) HIP_OK(hipEventSynchronize(g.stage_ev[r]));
  if (g.stage_cap[r] < cnt) {
    if (g.stage[r]) HIP_OK(hipHostFree(g.stage[r]));
    g.stage[r] = nullptr;
    g.stage_cap[r] = std::max(cnt, (size_t)256) * 2;
    HIP_OK(hipHostMalloc((void**)&g.stage[r], g.stage_cap[r] * sizeof(ItemDev), hipHostMallocDefault));
  }
  std::copy(g.host_items.begin(), g.host_items.end(), g.stage[r]);
  HIP_OK(hipMemcpyAsync(g.items_dev.p, g.stage[r], cnt * sizeof(ItemDev), hipMemcpyHostToDevice, g.st));
  HIP_OK(hipEventRecord(g.stage_ev[r], g.st));
  g.stage_used[r] = true;
}

void step_group(Server& s, Group& g) {
  if (g.cohorts.empty()) return;
  // at most kAhead steps of this side queued on the device: a job submitted now joins
  // within a few slice steps instead of behind an unbounded queue
  const int ring = g.step % Group::kAhead;
  if (g.ahead_set[ring]) HIP_OK(hipEventSynchronize(g.ahead_ev[ring]));
  if (g.dirty) {
    upload_items(s, g);
    g.dirty = false;
  }
  const int cnt = (int)g.host_items.size();
  const ItemDev* items = g.items_dev.as<ItemDev>();
  const JobDev* jobs = s.jobs_dev.as<JobDev>();
  pass_rows(g.n, g.plan, g.st, items, cnt, jobs, g.step);
  // the columns pass runs the largest band of the jobs in flight (each job skips the
  // groups beyond its own)
  int groups = 0;
  for (const Cohort& c : g.cohorts) groups = std::max(groups, s.jobs.at(c.job).band_ng);
  if (groups > 0) pass_cols(g.n, g.plan, g.st, items, cnt, jobs, g.step, groups);
  bool removed = false;
  for (Cohort& c : g.cohorts) {
    Job& J = s.jobs.at(c.job);
    if (g.step - c.start != J.j.slices - 1) continue;
    reduce_det(g.st, items + c.offset, nullptr, c.count, jobs, detector_groups(g.n, g.plan), s.max_mid + 1,
               s.max_aper);
    for (int slot : c.slots) release_slot(s, g, slot);
    J.done += c.count;
    if (J.done == J.total) {
      HIP_OK(hipEventRecord(J.ev, g.st));
      J.closed = true;
    }
    c.count = 0;
    removed = true;
  }
  if (removed) {
    g.cohorts.erase(std::remove_if(g.cohorts.begin(), g.cohorts.end(), [](const Cohort& c) { return c.count == 0; }),
                    g.cohorts.end());
    g.dirty = true;
  }
  HIP_OK(hipEventRecord(g.ahead_ev[ring], g.st));
  g.ahead_set[ring] = true;
  ++g.step;
}

void submit_impl(Server& s, const totem_stem_job& j, long long* id) {
  HIP_OK(hipSetDevice(s.dev));
  validate_job(j);
  require(j.n <= kMaxSide, "grid side above " + std::to_string(kMaxSide), TOTEM_ERR_UNSUPPORTED);
  require(j.n_aper <= s.max_aper, "job has more detectors than the server's max_aper");
  require(j.n_mid <= s.max_mid, "job has more mid slices than the server's max_mid");
  require(j.slices <= kMaxSlices, "job has more than " + std::to_string(kMaxSlices) + " slices");
  require(!s.free_tables.empty(), "server job table is full", TOTEM_ERR_MEMORY);
  Job J;
  J.j = j;
  J.j.mid = nullptr;  // the host list is read here only
  J.id = s.next_id++;
  J.total = (long)j.n_probes * j.depth;
  std::vector<int> hit;
  mid_walk(j, J.mos, hit);
  J.table = s.free_tables.back();
  int* meta = s.meta_dev.as<int>() + (size_t)J.table * s.meta_ints();
  JobDev jd;
  jd.t = static_cast<const float2*>(j.t);
  jd.prop = static_cast<const float2*>(j.prop);
  jd.aper = j.aper;
  jd.aper2 = j.aper2;
  jd.mid_of_slice = j.n_mid > 0 ? meta : nullptr;
  jd.layer_hit = meta + kMaxSlices;
  jd.det = j.det;
  jd.exit = static_cast<float2*>(j.exit_wave);
  jd.np = j.np;
  jd.S = j.slices;
  jd.A = j.n_aper;
  jd.L = j.n_mid;
  jd.n_out = (int)J.total;
  jd.wrap = j.wrap;
  jd.prop_cols = 0;  // the caller's row-major propagator
  // the job's band, as a run alone finds it (the same groups, so the same images); the
  // null stream does not wait for the server's (non-blocking) streams
  Group& g = group_of(s, j.n);
  const int C = pass_cols_width(j.n, g.plan);
  job_band(nullptr, jd, jd.prop, s.col_flags.as<int>(), j.n, C);
  J.band_ng = std::min(jd.band_ng, j.n / C);
  if (jd.band_len < j.n && j.exit_wave) {  // exit-wave columns outside the band stay zero
    fill(nullptr, jd.exit, J.total * j.n * j.n, 0.f, 0.f);
    HIP_OK(hipStreamSynchronize(nullptr));
  }
  // the entry is free (its previous job completed), so a blocking copy cannot race a kernel
  HIP_OK(hipMemcpy(meta, J.mos.data(), J.mos.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(meta + kMaxSlices, hit.data(), hit.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s.jobs_dev.as<JobDev>() + J.table, &jd, sizeof(jd), hipMemcpyHostToDevice));
  HIP_OK(hipEventCreateWithFlags(&J.ev, hipEventDisableTiming));
  s.free_tables.pop_back();
  s.order.push_back(J.id);
  *id = J.id;
  s.jobs.emplace(J.id, std::move(J));
}

}  // namespace srv
}  // namespace totem

using totem::guarded;
using totem::require;

extern "C" {

int totem_server_create(size_t arena_bytes, int max_aper, int max_mid, totem_server** out) {
  return guarded([&] {
    require(out != nullptr, "out is null");
    *out = nullptr;
    require(max_aper > 0 && max_aper <= 64 && max_mid >= 0 && max_mid <= 256, "max_aper in 1..64, max_mid in 0..256");
    auto s = std::make_unique<totem_server>();
    HIP_OK(hipGetDevice(&s->dev));
    s->max_aper = max_aper;
    s->max_mid = max_mid;
    s->page_bytes = s->slot_bytes(totem::srv::kMaxSide, nullptr);
    s->n_pages = (int)(arena_bytes / s->page_bytes);
    require(s->n_pages > 0, "arena smaller than one page (" + std::to_string(s->page_bytes) + " bytes)");
    s->arena.alloc((size_t)s->n_pages * s->page_bytes);
    s->jobs_dev.alloc(totem::srv::kMaxJobs * sizeof(totem::JobDev));
    s->meta_dev.alloc((size_t)totem::srv::kMaxJobs * s->meta_ints() * sizeof(int));
    s->col_flags.alloc(totem::srv::kMaxSide * sizeof(int));
    for (int p = s->n_pages - 1; p >= 0; --p) s->free_pages.push_back(p);
    s->page_ev.resize(s->n_pages);
    s->page_ev_set.assign(s->n_pages, false);
    for (auto& e : s->page_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (int t = totem::srv::kMaxJobs - 1; t >= 0; --t) s->free_tables.push_back(t);
    *out = s.release();
  });
}

int totem_server_submit(totem_server* srv, const totem_stem_job* job, long long* id) {
  return guarded([&] {
    require(srv && job && id, "null argument");
    totem::srv::submit_impl(*srv, *job, id);
  });
}

int totem_server_step(totem_server* srv, int* in_flight) {
  return guarded([&] {
    require(srv != nullptr, "null server");
    HIP_OK(hipSetDevice(srv->dev));
    totem::srv::admit(*srv);
    long n = 0;
    for (auto& kv : srv->groups) {
      totem::srv::step_group(*srv, kv.second);
      for (const auto& c : kv.second.cohorts) n += c.count;
    }
    HIP_OK(hipGetLastError());
    if (in_flight) *in_flight = (int)n;
  });
}

int totem_server_poll(totem_server* srv, long long* ids, int cap, int* n_done) {
  return guarded([&] {
    require(srv && n_done && (cap == 0 || ids), "null argument");
    int n = 0;
    for (auto it = srv->jobs.begin(); it != srv->jobs.end() && n < cap;) {
      totem::srv::Job& J = it->second;
      if (J.closed && hipEventQuery(J.ev) == hipSuccess) {
        ids[n++] = J.id;
        HIP_OK(hipEventDestroy(J.ev));
        srv->free_tables.push_back(J.table);
        it = srv->jobs.erase(it);
      } else {
        ++it;
      }
    }
    *n_done = n;
  });
}

int totem_server_sync(totem_server* srv) {
  return guarded([&] {
    require(srv != nullptr, "null server");
    for (auto& kv : srv->groups) HIP_OK(hipStreamSynchronize(kv.second.st));
  });
}

int totem_server_destroy(totem_server* srv) {
  return guarded([&] {
    if (!srv) return;
    std::unique_ptr<totem_server> s(srv);
    HIP_OK(hipSetDevice(s->dev));
    for (auto& kv : s->groups) {
      totem::srv::Group& g = kv.second;
      HIP_OK(hipStreamSynchronize(g.st));
      for (int r = 0; r < totem::srv::Group::kRing; ++r) {
        if (g.stage[r]) HIP_OK(hipHostFree(g.stage[r]));
        HIP_OK(hipEventDestroy(g.stage_ev[r]));
      }
      for (int r = 0; r < totem::srv::Group::kAhead; ++r) HIP_OK(hipEventDestroy(g.ahead_ev[r]));
      HIP_OK(hipStreamDestroy(g.st));
    }
    for (auto& e : s->page_ev) HIP_OK(hipEventDestroy(e));
    for (auto& kv : s->jobs) HIP_OK(hipEventDestroy(kv.second.ev));
  });
}

}  // extern "C"
