// Batched execution and the serving engine (no counterpart in the reference,
// whose cudaMain runs one simulation per call: src/cudaMain.cu:377-1667).
//
// Batch: members of uneven size are spread over several HIP streams by a
// longest-processing-time-first schedule (each stream with its own work
// buffers), their constant tables are made resident in one upload, and each
// distinct grid shape is autotuned once.  A batch plan captures that whole
// multi-stream schedule in a hipGraph once and replays it.
//
// Engine: requests are admitted FIFO by HBM footprint against a budget; each
// of `slots` host threads drives its own stream, so requests that arrive while
// others run are picked up as soon as a slot and memory free up and their work
// overlaps on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "host_math.hpp"
#include "kernels.hpp"
#include "pipeline.hpp"

namespace cyr {

namespace {

// Relative cost of one simulation, for scheduling: HBM bytes the executor moves.
double sim_cost(const cyr_sim_desc *d, const SimGeometry &g) {
  const double V = (double)g.V, M = g.M, units = (double)g.nE * g.nK;
  const int pipe = resolved_pipeline(d, g);
  if (pipe == CYR_PIPE_MBASIS)  // basis passes per k, then 2D work per energy
    return g.nK * V * (16.0 * M + 7.0 * M * 8.0 * 6.0) + units * g.P * (16.0 * 7.0 * M + 8.0 * g.nA);
  if (pipe == CYR_PIPE_EBASIS) return units * V * (16.0 * M + 80.0);
  return units * g.nA * V * (16.0 * M + 24.0 + 96.0);
}

// Tunes every distinct E-basis grid shape of the batch that has no tuning yet.
cyr_status tune_shapes(uint32_t count, const cyr_sim_desc *descs, const std::vector<SimGeometry> &geo,
                       WorkBuffers &wb, hipStream_t st) {
  std::set<std::tuple<uint32_t, uint32_t, uint32_t, int>> seen;
  for (uint32_t i = 0; i < count; i++) {
    const cyr_sim_desc *d = &descs[i];
    const SimGeometry &g = geo[i];
    if (resolved_pipeline(d, g) != CYR_PIPE_EBASIS || g.nz < 2) continue;
    const auto key = std::make_tuple(g.nx, g.ny, g.nz, d->precision);
    if (!seen.insert(key).second) continue;
    const YzTuning t = ebasis_yz_tuning(g.nx, g.ny, g.nz, d->precision);
    if (t.columns || t.target_blocks) continue;
    YzTuning chosen{};
    CYR_TRY(autotune_shape(d, g, wb, st, &chosen));
  }
  return CYR_OK;
}

// A batch laid out on `nl` streams: member order per stream, resident tables.
struct Schedule {
  std::vector<SimGeometry> geo;
  std::vector<std::vector<uint32_t>> per_lane;
  std::vector<size_t> table_off;
  DeviceBuffer tables;
};

cyr_status plan_schedule(uint32_t count, const cyr_sim_desc *descs, const double *const *mats, uint32_t nl,
                         hipStream_t st, Schedule &s) {
  s.geo.resize(count);
  for (uint32_t i = 0; i < count; i++) CYR_TRY(build_geometry(&descs[i], s.geo[i]));
  // longest processing time first onto the least-loaded stream
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; i++) order[i] = i;
  std::vector<double> cost(count);
  for (uint32_t i = 0; i < count; i++) cost[i] = sim_cost(&descs[i], s.geo[i]);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  s.per_lane.assign(nl, {});
  std::vector<double> load(nl, 0.0);
  for (uint32_t i : order) {
    const uint32_t l = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
    s.per_lane[l].push_back(i);
    load[l] += cost[i];
  }
  // every member's constant tables in one device buffer, one upload
  s.table_off.resize(count);
  size_t total = 0;
  for (uint32_t i = 0; i < count; i++) {
    s.table_off[i] = total;
    total += (table_bytes(&descs[i], s.geo[i]) + 255) & ~(size_t)255;
  }
  CYR_TRY(ensure_buffer(s.tables, total));
  std::vector<unsigned char> host(total);
  for (uint32_t i = 0; i < count; i++) fill_tables(&descs[i], s.geo[i], mats[i], host.data() + s.table_off[i]);
  CYR_HIP(hipMemcpyAsync(s.tables.ptr, host.data(), total, hipMemcpyHostToDevice, st));
  CYR_HIP(hipStreamSynchronize(st));
  return CYR_OK;
}

struct Lanes {
  std::vector<hipStream_t> st;
  std::vector<std::unique_ptr<WorkBuffers>> wb;
  std::vector<hipEvent_t> ev;
  hipEvent_t fork = nullptr;
  cyr_status create(uint32_t n) {
    CYR_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (uint32_t i = 0; i < n; i++) {
      hipStream_t s;
      CYR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
      st.push_back(s);
      hipEvent_t e;
      CYR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev.push_back(e);
      wb.emplace_back(new WorkBuffers());
      wb.back()->max_lanes = 1;
    }
    return CYR_OK;
  }
  void destroy() {
    for (auto s : st) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    for (auto &w : wb) w->release();
    for (auto e : ev) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    st.clear();
    wb.clear();
    ev.clear();
    fork = nullptr;
  }
};

// Queues the whole schedule: fork from `main`, each stream runs its members with
// resident tables, join back into `main`.
cyr_status queue_schedule(const cyr_sim_desc *descs, const void *const *d_voxels, const double *const *mats,
                          void *const *d_outs, const Schedule &s, Lanes &L, hipStream_t main) {
  CYR_HIP(hipEventRecord(L.fork, main));
  for (size_t l = 0; l < L.st.size(); l++) {
    CYR_HIP(hipStreamWaitEvent(L.st[l], L.fork, 0));
    for (uint32_t i : s.per_lane[l]) {
      const void *tab = reinterpret_cast<const unsigned char *>(s.tables.ptr) + s.table_off[i];
      CYR_TRY(run_simulation(&descs[i], s.geo[i], d_voxels[i], mats[i], d_outs[i], *L.wb[l], L.st[l], kRunResident,
                             tab));
    }
    CYR_HIP(hipEventRecord(L.ev[l], L.st[l]));
    CYR_HIP(hipStreamWaitEvent(main, L.ev[l], 0));
  }
  return CYR_OK;
}

uint32_t default_lanes(uint32_t count) { return count < 4 ? (count ? count : 1) : 4; }

// Streams and work buffers of cyr_simulate_batch, kept per device across calls
// (grown on demand) so a call does not re-allocate them.
struct BatchState {
  std::mutex mu;
  Lanes lanes;
};
std::mutex g_batch_mu;
std::map<int, BatchState *> g_batch;

cyr_status batch_state(uint32_t nl, BatchState **out) {
  int dev = 0;
  CYR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_batch_mu);
  BatchState *&b = g_batch[dev];
  if (!b) b = new BatchState();
  *out = b;
  (void)nl;
  return CYR_OK;
}

}  // namespace

void release_batch_buffers() {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_batch) {
    std::lock_guard<std::mutex> l2(kv.second->mu);
    (void)hipSetDevice(kv.first);
    kv.second->lanes.destroy();
  }
  (void)hipSetDevice(cur);
}

// ---- batch plan --------------------------------------------------------------

struct BatchPlan {
  int device = 0;
  std::vector<cyr_sim_desc> descs;
  std::vector<std::vector<double>> energies, kvecs, mats;
  std::vector<const void *> vox;
  std::vector<const double *> matp;
  std::vector<void *> outs;
  Schedule sched;
  Lanes lanes;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // also what a failed cyr_batch_create leaves behind
  ~BatchPlan() {
    if (cap) (void)hipStreamSynchronize(cap);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    lanes.destroy();
    free_buffer(sched.tables);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

}  // namespace cyr

using namespace cyr;

struct cyr_batch_s {
  BatchPlan p;
};

extern "C" {

cyr_status cyr_simulate_batch(uint32_t count, const cyr_sim_desc *descs, const void *const *d_voxels,
                              const double *const *mats, void *const *d_outs, void *stream) {
  if (count == 0) return CYR_OK;
  if (!descs || !d_voxels || !mats || !d_outs) return fail(CYR_E_INVALID, "null batch arrays");
  for (uint32_t i = 0; i < count; i++)
    if (!d_voxels[i] || !mats[i] || !d_outs[i]) return fail(CYR_E_INVALID, "null batch member buffer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto ws = workspace_use();
  BatchState *bs;
  CYR_TRY(batch_state(default_lanes(count), &bs));
  std::lock_guard<std::mutex> lk(bs->mu);
  Lanes &L = bs->lanes;
  cyr_status rc = CYR_OK;
  if (L.st.size() < default_lanes(count)) {
    L.destroy();
    rc = L.create(default_lanes(count));
  }
  Schedule s;
  if (rc == CYR_OK) rc = plan_schedule(count, descs, mats, (uint32_t)L.st.size(), st, s);
  if (rc == CYR_OK) rc = tune_shapes(count, descs, s.geo, *L.wb[0], L.st[0]);
  if (rc == CYR_OK) {
    for (auto ls : L.st) (void)hipStreamSynchronize(ls);
    rc = queue_schedule(descs, d_voxels, mats, d_outs, s, L, st);
  }
  const hipError_t se = hipStreamSynchronize(st);
  if (rc == CYR_OK && se != hipSuccess) rc = hip_fail(se, "hipStreamSynchronize", __FILE__, __LINE__);
  free_buffer(s.tables);
  return rc;
}

cyr_status cyr_batch_create(uint32_t count, const cyr_sim_desc *descs, const void *const *d_voxels,
                            const double *const *mats, void *const *d_outs, uint32_t lanes, cyr_batch *out) {
  if (!out) return fail(CYR_E_INVALID, "null output handle");
  *out = nullptr;
  if (count == 0 || !descs || !d_voxels || !mats || !d_outs) return fail(CYR_E_INVALID, "empty or null batch");
  auto ws = workspace_use();
  std::unique_ptr<cyr_batch_s> b(new cyr_batch_s());
  BatchPlan &p = b->p;
  CYR_HIP(hipGetDevice(&p.device));
  // own copies of the descriptors and the arrays they point to
  p.descs.assign(descs, descs + count);
  p.energies.resize(count);
  p.kvecs.resize(count);
  p.mats.resize(count);
  for (uint32_t i = 0; i < count; i++) {
    cyr_sim_desc &d = p.descs[i];
    if (!d_voxels[i] || !mats[i] || !d_outs[i] || !d.energies) return fail(CYR_E_INVALID, "null batch member");
    p.energies[i].assign(d.energies, d.energies + d.num_energies);
    d.energies = p.energies[i].data();
    if (d.num_kvectors && d.kvectors) {
      p.kvecs[i].assign(d.kvectors, d.kvectors + 3 * (size_t)d.num_kvectors);
      d.kvectors = p.kvecs[i].data();
    }
    p.mats[i].assign(mats[i], mats[i] + 4 * (size_t)d.num_energies * d.num_materials);
    p.vox.push_back(d_voxels[i]);
    p.matp.push_back(p.mats[i].data());
    p.outs.push_back(d_outs[i]);
  }
  CYR_HIP(hipStreamCreateWithFlags(&p.cap, hipStreamNonBlocking));
  CYR_TRY(p.lanes.create(lanes ? std::min(lanes, 8u) : default_lanes(count)));
  CYR_TRY(plan_schedule(count, p.descs.data(), p.matp.data(), (uint32_t)p.lanes.st.size(), p.cap, p.sched));
  CYR_TRY(tune_shapes(count, p.descs.data(), p.sched.geo, *p.lanes.wb[0], p.lanes.st[0]));
  // eager pass: sizes every buffer and creates every twiddle table, so the
  // captured graph holds kernels and event edges only
  CYR_TRY(queue_schedule(p.descs.data(), p.vox.data(), p.matp.data(), p.outs.data(), p.sched, p.lanes, p.cap));
  CYR_HIP(hipStreamSynchronize(p.cap));
  CYR_HIP(hipStreamBeginCapture(p.cap, hipStreamCaptureModeRelaxed));
  const cyr_status qs = queue_schedule(p.descs.data(), p.vox.data(), p.matp.data(), p.outs.data(), p.sched, p.lanes,
                                       p.cap);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(p.cap, &g);
  if (qs != CYR_OK) {
    if (g) (void)hipGraphDestroy(g);
    return qs;
  }
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
  p.graph = g;
  CYR_HIP(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
  live_handles_add(1);  // the graph holds twiddle and work-buffer pointers
  *out = b.release();
  return CYR_OK;
}

cyr_status cyr_batch_launch(cyr_batch b, void *stream) {
  if (!b || !b->p.exec) return fail(CYR_E_INVALID, "invalid batch plan");
  CYR_HIP(hipGraphLaunch(b->p.exec, reinterpret_cast<hipStream_t>(stream)));
  return CYR_OK;
}

cyr_status cyr_batch_destroy(cyr_batch b) {
  if (!b) return CYR_OK;
  delete b;  // ~BatchPlan drains and frees
  live_handles_add(-1);
  return CYR_OK;
}

}  // extern "C"

// ---- serving engine ------------------------------------------------------------

namespace cyr {
namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kCancelled = 3, kFailed = 4 };

struct Request {
  uint64_t id = 0;
  cyr_sim_desc desc{};
  std::vector<double> energies, kvecs, mats;
  const void *vox = nullptr;
  void *out = nullptr;
  uint64_t footprint = 0;
  int state = kQueued;
  cyr_status status = CYR_OK;
  Clock::time_point t_submit, t_start, t_done;
};

struct Slot {
  hipStream_t st = nullptr;
  WorkBuffers wb;
  DeviceBuffer vox, out;
  std::thread th;
  // device bytes the slot's buffers hold between requests, and the bytes
  // counted against the budget for it (held, plus a running request's footprint)
  uint64_t held = 0, reserved = 0;
  uint64_t bytes_now() const { return vox.bytes + out.bytes + wb.bytes(); }
  void release_buffers() {
    wb.release();
    free_buffer(vox);
    free_buffer(out);
  }
};

}  // namespace

struct Engine {
  int device = 0;
  uint64_t budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::map<uint64_t, std::shared_ptr<Request>> reqs;
  std::vector<std::unique_ptr<Slot>> slots;
  uint64_t next_id = 1, resident = 0;
  uint64_t submitted = 0, completed = 0, rejected = 0, cancelled = 0;
  uint32_t in_flight = 0;
  bool stop = false;

  void slot_loop(Slot *s);
  cyr_status run_request(Slot *s, Request &r);
};

cyr_status Engine::run_request(Slot *s, Request &r) {
  SimGeometry g;
  CYR_TRY(build_geometry(&r.desc, g));
  const uint64_t vb = morphology_bytes(&r.desc), ob = output_bytes(&r.desc, g);
  CYR_TRY(ensure_buffer(s->vox, vb));
  CYR_TRY(ensure_buffer(s->out, ob));
  std::vector<unsigned char> conv;
  if (r.desc.storage == CYR_STORE_NATIVE) {
    CYR_TRY(upload_host(s->vox.ptr, r.vox, vb, s->st));
  } else {
    convert_storage(&r.desc, r.vox, conv);
    CYR_TRY(upload_host(s->vox.ptr, conv.data(), vb, s->st));
  }
  CYR_TRY(run_simulation(&r.desc, g, s->vox.ptr, r.mats.data(), s->out.ptr, s->wb, s->st, kRunAsync));
  CYR_HIP(hipMemcpyAsync(r.out, s->out.ptr, ob, hipMemcpyDeviceToHost, s->st));
  CYR_HIP(hipStreamSynchronize(s->st));
  return CYR_OK;
}

// Admission (FIFO): `resident` sums every slot's reservation, so buffers an idle
// slot keeps from an earlier request count against the budget too.  A slot takes
// the queue head when its footprint fits next to the other slots' reservations
// and what this slot holds (buffers only grow, so the run may hold both); when it
// fits only without this slot's buffers, the slot frees them first; when it does
// not fit at all, an idle slot that holds buffers frees them, since they may be
// what keeps the head out.  With every slot idle and empty the head always fits
// (submit rejects footprints above the budget), so the queue keeps draining.
void Engine::slot_loop(Slot *s) {
  (void)hipSetDevice(device);
  for (;;) {
    std::shared_ptr<Request> r;
    bool shrink = false, admit = false;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (stop) return;
        if (!queue.empty()) {
          const auto &h = reqs[queue.front()];
          const uint64_t other = resident - s->reserved;
          if (other + s->held + h->footprint <= budget) {
            admit = true;
          } else if (other + h->footprint <= budget) {
            admit = shrink = true;
          } else if (s->held > 0) {
            shrink = true;
          }
        }
        if (admit || shrink) break;
        cv_work.wait(lk);
      }
      if (admit) {
        r = reqs[queue.front()];
        queue.pop_front();
        r->state = kRunning;
        r->t_start = Clock::now();
        in_flight++;
        const uint64_t res = (shrink ? 0 : s->held) + r->footprint;
        resident = resident - s->reserved + res;
        s->reserved = res;
      }
    }
    if (shrink) {  // this slot's stream is idle: nothing queued on it uses the buffers
      s->release_buffers();
      std::lock_guard<std::mutex> lk(mu);
      s->held = 0;
      if (!admit) {
        resident -= s->reserved;
        s->reserved = 0;
      }
    }
    if (!admit) {
      cv_work.notify_all();
      continue;
    }
    const cyr_status st = run_request(s, *r);
    {
      std::lock_guard<std::mutex> lk(mu);
      r->status = st;
      r->state = st == CYR_OK ? kDone : kFailed;
      r->t_done = Clock::now();
      s->held = s->bytes_now();
      resident = resident - s->reserved + s->held;
      s->reserved = s->held;
      in_flight--;
      completed++;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

}  // namespace cyr

struct cyr_engine_s {
  Engine e;
};

extern "C" {

cyr_status cyr_engine_create(int device, uint64_t budget_bytes, uint32_t slots, cyr_engine *out) {
  if (!out) return fail(CYR_E_INVALID, "null output handle");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    return fail(CYR_E_NODEVICE, "no HIP device found");
  }
  if (device < 0 || device >= n) return fail(CYR_E_INVALID, "device out of range");
  if (slots < 1 || slots > 16) return fail(CYR_E_INVALID, "slots must be 1..16");
  CYR_HIP(hipSetDevice(device));
  if (budget_bytes == 0) {  // default: 90% of the device memory
    size_t fr = 0, tot = 0;
    CYR_HIP(hipMemGetInfo(&fr, &tot));
    budget_bytes = (uint64_t)(0.9 * (double)tot);
  }
  std::unique_ptr<cyr_engine_s> h(new cyr_engine_s());
  Engine &e = h->e;
  e.device = device;
  e.budget = budget_bytes;
  for (uint32_t i = 0; i < slots; i++) {
    std::unique_ptr<Slot> s(new Slot());
    CYR_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));  // default priority
    s->wb.max_lanes = 1;
    e.slots.push_back(std::move(s));
  }
  for (auto &s : e.slots) {
    Slot *sp = s.get();
    sp->th = std::thread([&e, sp] { e.slot_loop(sp); });
  }
  live_handles_add(1);  // slot threads use the shared device caches while it lives
  *out = h.release();
  return CYR_OK;
}

cyr_status cyr_engine_submit(cyr_engine eng, const cyr_sim_desc *desc, const void *voxels, const double *mats,
                             void *out, uint64_t *request_id) {
  if (!eng || !desc || !voxels || !mats || !out) return fail(CYR_E_INVALID, "null argument");
  Engine &e = eng->e;
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  auto r = std::make_shared<Request>();
  r->desc = *desc;
  r->energies.assign(desc->energies, desc->energies + desc->num_energies);
  r->desc.energies = r->energies.data();
  if (desc->num_kvectors && desc->kvectors) {
    r->kvecs.assign(desc->kvectors, desc->kvectors + 3 * (size_t)desc->num_kvectors);
    r->desc.kvectors = r->kvecs.data();
  }
  r->mats.assign(mats, mats + 4 * (size_t)desc->num_energies * desc->num_materials);
  r->vox = voxels;
  r->out = out;
  r->footprint = cyr_footprint_bytes(&r->desc);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(e.mu);
  if (e.stop) return fail(CYR_E_INVALID, "engine is shutting down");
  if (r->footprint > e.budget) {
    e.rejected++;
    return fail(CYR_E_ADMISSION, "request footprint exceeds the engine's HBM budget");
  }
  r->id = e.next_id++;
  e.reqs[r->id] = r;
  e.queue.push_back(r->id);
  e.submitted++;
  if (request_id) *request_id = r->id;
  e.cv_work.notify_all();
  return CYR_OK;
}

cyr_status cyr_engine_wait(cyr_engine eng, uint64_t request_id, double timeout_s, cyr_status *request_status) {
  if (!eng) return fail(CYR_E_INVALID, "null engine");
  Engine &e = eng->e;
  std::unique_lock<std::mutex> lk(e.mu);
  auto it = e.reqs.find(request_id);
  if (it == e.reqs.end()) return fail(CYR_E_NOTFOUND, "unknown request id");
  std::shared_ptr<Request> r = it->second;
  auto done = [&] { return r->state >= kDone; };
  if (timeout_s < 0) {
    e.cv_done.wait(lk, done);
  } else if (!e.cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
    return fail(CYR_E_TIMEOUT, "request still in flight");
  }
  if (request_status) *request_status = r->state == kCancelled ? CYR_E_CANCELLED : r->status;
  return CYR_OK;
}

cyr_status cyr_engine_poll(cyr_engine eng, uint64_t request_id, int *done, cyr_status *request_status) {
  if (!eng || !done) return fail(CYR_E_INVALID, "null argument");
  Engine &e = eng->e;
  std::lock_guard<std::mutex> lk(e.mu);
  auto it = e.reqs.find(request_id);
  if (it == e.reqs.end()) return fail(CYR_E_NOTFOUND, "unknown request id");
  const Request &r = *it->second;
  *done = r.state >= kDone ? 1 : 0;
  if (request_status) *request_status = r.state == kCancelled ? CYR_E_CANCELLED : r.status;
  return CYR_OK;
}

cyr_status cyr_engine_request_times(cyr_engine eng, uint64_t request_id, double *queued_s, double *run_s,
                                    double *latency_s) {
  if (!eng) return fail(CYR_E_INVALID, "null engine");
  Engine &e = eng->e;
  std::lock_guard<std::mutex> lk(e.mu);
  auto it = e.reqs.find(request_id);
  if (it == e.reqs.end()) return fail(CYR_E_NOTFOUND, "unknown request id");
  const Request &r = *it->second;
  if (r.state != kDone && r.state != kFailed) return fail(CYR_E_INVALID, "request has not run to completion");
  auto sec = [](Clock::duration d) { return std::chrono::duration<double>(d).count(); };
  if (queued_s) *queued_s = sec(r.t_start - r.t_submit);
  if (run_s) *run_s = sec(r.t_done - r.t_start);
  if (latency_s) *latency_s = sec(r.t_done - r.t_submit);
  return CYR_OK;
}

cyr_status cyr_engine_release(cyr_engine eng, uint64_t request_id) {
  if (!eng) return fail(CYR_E_INVALID, "null engine");
  Engine &e = eng->e;
  std::lock_guard<std::mutex> lk(e.mu);
  auto it = e.reqs.find(request_id);
  if (it == e.reqs.end()) return fail(CYR_E_NOTFOUND, "unknown request id");
  if (it->second->state < kDone) return fail(CYR_E_INVALID, "request still in flight");
  e.reqs.erase(it);
  return CYR_OK;
}

cyr_status cyr_engine_cancel(cyr_engine eng, uint64_t request_id) {
  if (!eng) return fail(CYR_E_INVALID, "null engine");
  Engine &e = eng->e;
  {
    std::lock_guard<std::mutex> lk(e.mu);
    auto it = e.reqs.find(request_id);
    if (it == e.reqs.end()) return fail(CYR_E_NOTFOUND, "unknown request id");
    Request &r = *it->second;
    if (r.state != kQueued) return fail(CYR_E_INVALID, "request is no longer queued");
    e.queue.erase(std::find(e.queue.begin(), e.queue.end(), request_id));
    r.state = kCancelled;
    r.status = CYR_E_CANCELLED;
    e.cancelled++;
  }
  e.cv_done.notify_all();
  e.cv_work.notify_all();
  return CYR_OK;
}

cyr_status cyr_engine_get_stats(cyr_engine eng, cyr_engine_stats *stats) {
  if (!eng || !stats) return fail(CYR_E_INVALID, "null argument");
  Engine &e = eng->e;
  std::lock_guard<std::mutex> lk(e.mu);
  stats->submitted = e.submitted;
  stats->completed = e.completed;
  stats->rejected = e.rejected;
  stats->cancelled = e.cancelled;
  stats->resident_bytes = e.resident;
  stats->budget_bytes = e.budget;
  stats->in_flight = e.in_flight;
  stats->queued = (uint32_t)e.queue.size();
  return CYR_OK;
}

cyr_status cyr_engine_destroy(cyr_engine eng) {
  if (!eng) return CYR_OK;
  Engine &e = eng->e;
  {
    std::lock_guard<std::mutex> lk(e.mu);
    e.stop = true;
    for (uint64_t id : e.queue) {
      e.reqs[id]->state = kCancelled;
      e.reqs[id]->status = CYR_E_CANCELLED;
      e.cancelled++;
    }
    e.queue.clear();
  }
  e.cv_work.notify_all();
  e.cv_done.notify_all();
  for (auto &s : e.slots) {
    if (s->th.joinable()) s->th.join();  // a running request finishes first
  }
  (void)hipSetDevice(e.device);
  for (auto &s : e.slots) {
    (void)hipStreamSynchronize(s->st);
    s->release_buffers();
    (void)hipStreamDestroy(s->st);
  }
  delete eng;
  live_handles_add(-1);
  return CYR_OK;
}

}  // extern "C"
