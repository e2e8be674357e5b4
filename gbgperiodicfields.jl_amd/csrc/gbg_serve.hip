// Serving loop: continuous batching of transform / operator / product jobs.
//
// Jobs are submitted from any host thread and queued in arrival order.  A
// dispatcher thread admits the oldest job as soon as one of `slots` streams
// is free and the job's plan (tables + work area, per grid, precision and
// slot) fits the HBM budget -- evicting idle plans least recently used first
// -- so new work starts while earlier jobs are still on the device and their
// kernels overlap across streams.  Completion is signalled by a host function
// enqueued behind the job on its stream; latency = completion - submission.
// Every job runs on a plan of its own slot, so its result does not depend on
// what else is in flight.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "gbg_internal.hpp"

namespace gbg {
using Clock = std::chrono::steady_clock;

using PlanKey = std::tuple<int, int, int, double, double, double, int, int>;

struct ServeJob {
  gbg_job job;
  unsigned long long ticket;
  Clock::time_point t_submit, t_done;
  int slot;
  PlanKey key;
  bool done;        // kernels finished (or the job failed)
  bool dispatched;  // the dispatcher no longer touches the job
  gbg_status status;
};

struct PlanEntry {
  gbg_plan* plan;
  int busy;
  unsigned long long last_use;
};

}  // namespace gbg

struct gbg_server {
  size_t budget;
  int slots;
  std::vector<hipStream_t> streams;
  std::vector<int> slot_busy;
  std::mutex mu;
  std::condition_variable cv;       // dispatcher wake-ups
  std::condition_variable cv_done;  // waiters
  std::deque<gbg::ServeJob*> pending;
  std::map<unsigned long long, gbg::ServeJob*> jobs;
  std::map<gbg::PlanKey, gbg::PlanEntry> plans;
  size_t resident;  // bytes held by cached plans
  unsigned long long next_ticket, use_clock;
  bool stop;
  std::thread worker;
  std::vector<double> latencies;
  gbg::Clock::time_point first_submit, last_done;
  bool any;
};

namespace gbg {

struct DoneArg {
  gbg_server* srv;
  ServeJob* job;
};

// Runs on the runtime's callback thread once the job's kernels are done: no
// HIP calls here, only bookkeeping under the server mutex (which the
// dispatcher never holds across a HIP call).
static void on_done(void* p) {
  DoneArg* a = static_cast<DoneArg*>(p);
  gbg_server* s = a->srv;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    a->job->t_done = Clock::now();
    a->job->done = true;
    s->slot_busy[a->job->slot] = 0;
    auto it = s->plans.find(a->job->key);
    if (it != s->plans.end()) it->second.busy = 0;
    s->latencies.push_back(
        std::chrono::duration<double, std::milli>(a->job->t_done - a->job->t_submit).count());
    s->last_done = a->job->t_done;
  }
  s->cv.notify_all();
  s->cv_done.notify_all();
  delete a;
}

// device bytes a job needs in its plan's work area
static size_t work_need(const gbg_job& j) {
  const size_t M = (size_t)(j.n[0] / 2 + 1) * j.n[1] * j.n[2];
  const size_t P = (size_t)j.n[0] * j.n[1] * j.n[2];
  const size_t cs = 2 * (size_t)j.elem_bytes;
  switch (j.kind) {
    case GBG_JOB_TO_PHYSICAL: return (size_t)j.ncomp * M * cs;
    case GBG_JOB_DEALIASED_PRODUCT:
    case GBG_JOB_UXW: {
      // the fused paths' compact intermediates depend on the cutoff
      gbg_plan geo;
      std::memset(&geo, 0, sizeof(geo));
      for (int a = 0; a < 3; ++a) {
        geo.n[a] = j.n[a];
        geo.L[a] = j.L[a];
      }
      geo.elem = j.elem_bytes;
      geo.nh = j.n[0] / 2 + 1;
      geo.M = M;
      geo.P = P;
      const char* env = std::getenv("GBG_GENERIC_FFT");
      geo.generic_only = (env && env[0] == '1') ? 1 : 0;
      return j.kind == GBG_JOB_UXW ? uxw_work_bytes(&geo, j.kc, 0)
                                   : dealiased_product_work_bytes(&geo, j.kc);
    }
    default: return 0;
  }
}

// Upper bound of a new plan's own device bytes before any work area, as
// gbg_plan_create lays them out: per axis the two FP64 wavenumber tables and
// the twiddle table, the reduction partials and results; plus one tile list
// of the fused pipelines' last pass (at most two ints per output tile of at
// least 8 pencils, one per (ky, kx) pencil bounds it).
static size_t plan_fixed_bytes(const gbg_job& j) {
  size_t b = sizeof(double) * kRedBlocks * 4 + sizeof(double) * 4;
  for (int a = 0; a < 3; ++a) {
    const size_t stored = a == 0 ? (size_t)(j.n[0] / 2 + 1) : (size_t)j.n[a];
    b += 2 * sizeof(double) * stored + 2 * (size_t)j.elem_bytes * (size_t)j.n[a];
  }
  if (j.kind == GBG_JOB_UXW || j.kind == GBG_JOB_DEALIASED_PRODUCT)
    b += (size_t)(j.n[0] / 2 + 1) * (size_t)j.n[1];
  return b;
}

static size_t plan_bytes(const gbg_plan* p) { return p->device_bytes; }

// Choose idle plans (least recently used first, never `keep`) whose release
// makes room for `need` more bytes; they leave the cache at once and are
// destroyed by the caller outside the lock.  False if even evicting every
// idle plan would not make room (the caller then waits for a completion).
static bool make_room(gbg_server* s, size_t need, const PlanKey& keep,
                      std::vector<gbg_plan*>& victims) {
  size_t freeable = 0;
  for (auto& kv : s->plans)
    if (!kv.second.busy && kv.first != keep) freeable += plan_bytes(kv.second.plan);
  if (s->resident + need > s->budget + freeable) return false;
  while (s->resident + need > s->budget) {
    auto victim = s->plans.end();
    for (auto it = s->plans.begin(); it != s->plans.end(); ++it)
      if (!it->second.busy && it->first != keep &&
          (victim == s->plans.end() || it->second.last_use < victim->second.last_use))
        victim = it;
    if (victim == s->plans.end()) return false;
    s->resident -= plan_bytes(victim->second.plan);
    victims.push_back(victim->second.plan);
    s->plans.erase(victim);
  }
  return true;
}

static bool buffers_overlap(const void* a, size_t na, const void* b, size_t nb) {
  const char *x = (const char*)a, *y = (const char*)b;
  return x < y + nb && y < x + na;
}

static gbg_status run_job(gbg_plan* p, const gbg_job& j, hipStream_t st) {
  switch (j.kind) {
    case GBG_JOB_TO_SPECTRAL: return fwd_transform(p, j.out, j.in0, j.ncomp, st);
    case GBG_JOB_TO_PHYSICAL: {
      gbg_status e = ensure_work(p, (size_t)j.ncomp * p->M * 2 * p->elem);
      if (e) return e;
      return inv_transform(p, j.out, (void*)j.in0, p->work, j.ncomp, st);
    }
    case GBG_JOB_CURL: return op_curl(p, j.out, j.in0, 0, st);
    case GBG_JOB_INVERSE_LAPLACIAN: return op_laplacian(p, j.out, j.in0, j.ncomp, 1, st);
    case GBG_JOB_DEALIASED_PRODUCT: return pipe_dealiased_product(p, j.out, j.in0, j.in1, j.kc, st);
    case GBG_JOB_UXW: return pipe_uxw(p, j.in0, j.out, j.kc, 0, st);
  }
  return fail(GBG_E_ARG, "unknown job kind");
}

// The dispatcher decides under the mutex and releases it for every HIP call
// (plan creation and destruction may synchronise the device, which waits on
// completion callbacks that take the mutex).
static void dispatcher(gbg_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto free_slot = [&]() {
    for (int i = 0; i < s->slots; ++i)
      if (!s->slot_busy[i]) return i;
    return -1;
  };
  while (true) {
    s->cv.wait(lk, [&] { return s->stop || (!s->pending.empty() && free_slot() >= 0); });
    if (s->pending.empty()) {
      if (s->stop) break;
      continue;
    }
    const int slot = free_slot();
    ServeJob* j = s->pending.front();
    const gbg_job jb = j->job;
    const PlanKey key =
        std::make_tuple(jb.n[0], jb.n[1], jb.n[2], jb.L[0], jb.L[1], jb.L[2], jb.elem_bytes, slot);
    auto it = s->plans.find(key);
    const bool fresh = it == s->plans.end();
    const size_t need = work_need(jb);
    const size_t have = fresh ? 0 : it->second.plan->work_bytes;
    const size_t grow = fresh ? need + plan_fixed_bytes(jb) : (have < need ? need - have : 0);
    std::vector<gbg_plan*> victims;
    if (grow && !make_room(s, grow, key, victims)) {
      bool inflight = false;
      for (int i = 0; i < s->slots; ++i) inflight = inflight || s->slot_busy[i];
      if (inflight) {
        s->cv.wait(lk);  // admitted once a completion frees a plan
        continue;
      }
      // nothing runs and even every idle plan evicted leaves too little room
      // (the job's own cached plan holds more than the budget allows beside
      // its work area): no completion will ever make room, so the job ends
      // with GBG_E_OOM instead of waiting
      s->pending.pop_front();
      j->status = GBG_E_OOM;
      j->done = true;
      j->dispatched = true;
      j->t_done = Clock::now();
      s->cv_done.notify_all();
      continue;
    }
    s->pending.pop_front();
    s->slot_busy[slot] = 1;
    j->slot = slot;
    j->key = key;
    s->resident += grow;  // reservation, settled below
    gbg_plan* p = fresh ? nullptr : it->second.plan;
    if (!fresh) {
      it->second.busy = 1;
      it->second.last_use = ++s->use_clock;
    }
    const long long before = fresh ? 0 : (long long)plan_bytes(p);
    lk.unlock();
    for (gbg_plan* v : victims) gbg_plan_destroy(v);
    gbg_status st = GBG_OK;
    if (fresh) st = gbg_plan_create(jb.n, jb.L, jb.elem_bytes, &p);
    if (!st) st = run_job(p, jb, s->streams[slot]);
    if (!st) {
      DoneArg* a = new DoneArg{s, j};
      st = check_hip(hipLaunchHostFunc(s->streams[slot], on_done, a), "completion callback");
      if (st) delete a;
    }
    lk.lock();
    const long long after = p ? (long long)plan_bytes(p) : 0;
    s->resident = (size_t)((long long)s->resident - (long long)grow + (after - before));
    if (fresh && p) s->plans[key] = PlanEntry{p, (st || j->done) ? 0 : 1, ++s->use_clock};
    if (st) {
      auto e = s->plans.find(key);
      if (e != s->plans.end()) e->second.busy = 0;
      s->slot_busy[slot] = 0;
      j->status = st;
      j->done = true;
      j->t_done = Clock::now();
    }
    j->dispatched = true;
    s->cv_done.notify_all();
  }
}

}  // namespace gbg

using namespace gbg;

extern "C" {

gbg_status gbg_server_create(size_t hbm_budget, int slots, gbg_server** out) {
  try {
    if (!out) return fail(GBG_E_ARG, "null pointer");
    *out = nullptr;
    if (slots < 1 || slots > 16) return fail(GBG_E_ARG, "slots must be in 1..16");
    if (hbm_budget == 0) return fail(GBG_E_ARG, "budget must be positive");
    gbg_server* s = new gbg_server();
    s->budget = hbm_budget;
    s->slots = slots;
    s->slot_busy.assign(slots, 0);
    s->resident = 0;
    s->next_ticket = 1;
    s->use_clock = 0;
    s->stop = false;
    s->any = false;
    for (int i = 0; i < slots; ++i) {
      hipStream_t st = nullptr;
      gbg_status e = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "server stream");
      if (e) {
        for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
        delete s;
        return e;
      }
      s->streams.push_back(st);
    }
    s->worker = std::thread(dispatcher, s);
    *out = s;
    return GBG_OK;
  } catch (...) {
    return fail(GBG_E_OOM, "server creation failed");
  }
}

gbg_status gbg_server_submit(gbg_server* s, const gbg_job* job, unsigned long long* ticket) {
  try {
    if (!s || !job || !ticket) return fail(GBG_E_ARG, "null argument");
    const gbg_job& j = *job;
    if (j.kind < GBG_JOB_TO_SPECTRAL || j.kind > GBG_JOB_UXW) return fail(GBG_E_ARG, "unknown job kind");
    if (j.elem_bytes != 4 && j.elem_bytes != 8) return fail(GBG_E_ARG, "elem_bytes must be 4 or 8");
    for (int a = 0; a < 3; ++a)
      if (j.n[a] < 1 || !(j.L[a] > 0)) return fail(GBG_E_ARG, "bad grid");
    if (!j.in0 || !j.out || (j.kind == GBG_JOB_DEALIASED_PRODUCT && !j.in1))
      return fail(GBG_E_ARG, "null field buffer");
    const int nc = j.ncomp;
    if ((j.kind == GBG_JOB_CURL || j.kind == GBG_JOB_UXW) && nc != 3) return fail(GBG_E_DIM, "ncomp must be 3");
    if (j.kind == GBG_JOB_DEALIASED_PRODUCT && nc != 1) return fail(GBG_E_DIM, "ncomp must be 1");
    if (nc != 1 && nc != 3 && nc != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
    if ((j.kind == GBG_JOB_UXW || j.kind == GBG_JOB_DEALIASED_PRODUCT) && (!std::isfinite(j.kc) || j.kc < 0))
      return fail(GBG_E_ARG, "cutoff must be finite and >= 0");
    // the buffers a job reads and writes (elements of one component)
    const size_t M = (size_t)(j.n[0] / 2 + 1) * j.n[1] * j.n[2];
    const size_t P = (size_t)j.n[0] * j.n[1] * j.n[2];
    const size_t spec = (size_t)nc * M * 2 * j.elem_bytes, phys = (size_t)nc * P * j.elem_bytes;
    const size_t in_b = j.kind == GBG_JOB_TO_SPECTRAL ? phys : spec;
    const size_t out_b = j.kind == GBG_JOB_TO_PHYSICAL ? phys : spec;
    const bool alias_ok = (j.kind == GBG_JOB_CURL || j.kind == GBG_JOB_INVERSE_LAPLACIAN) &&
                          j.in0 == j.out;
    if (!alias_ok && buffers_overlap(j.out, out_b, j.in0, in_b))
      return fail(GBG_E_ALIAS, "the job's output overlaps its input");
    if (j.kind == GBG_JOB_DEALIASED_PRODUCT && buffers_overlap(j.out, out_b, j.in1, in_b))
      return fail(GBG_E_ALIAS, "the job's output overlaps its second input");
    if (work_need(j) + plan_fixed_bytes(j) > s->budget)
      return fail(GBG_E_OOM, "job needs more device memory than the server's budget");
    ServeJob* sj = new ServeJob();
    sj->job = j;
    sj->done = false;
    sj->dispatched = false;
    sj->status = GBG_OK;
    sj->slot = -1;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      sj->ticket = s->next_ticket++;
      sj->t_submit = Clock::now();
      if (!s->any) {
        s->first_submit = sj->t_submit;
        s->any = true;
      }
      s->jobs[sj->ticket] = sj;
      s->pending.push_back(sj);
      *ticket = sj->ticket;
    }
    s->cv.notify_all();
    return GBG_OK;
  } catch (...) {
    return fail(GBG_E_OOM, "host allocation");
  }
}

gbg_status gbg_server_wait(gbg_server* s, unsigned long long ticket, double* latency_ms) {
  if (!s) return fail(GBG_E_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return fail(GBG_E_ARG, "unknown ticket");
  ServeJob* j = it->second;
  s->cv_done.wait(lk, [&] { return j->done && j->dispatched; });
  const gbg_status st = j->status;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
  s->jobs.erase(it);
  delete j;
  if (st) return fail(st, "job failed");
  return GBG_OK;
}

gbg_status gbg_server_poll(gbg_server* s, unsigned long long ticket, int* done,
                           double* latency_ms) {
  if (!s) return fail(GBG_E_ARG, "null server");
  if (!done) return fail(GBG_E_ARG, "null done flag");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return fail(GBG_E_ARG, "unknown ticket");
  ServeJob* j = it->second;
  if (!(j->done && j->dispatched)) {
    *done = 0;
    return GBG_OK;
  }
  *done = 1;
  const gbg_status st = j->status;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
  s->jobs.erase(it);
  delete j;
  if (st) return fail(st, "job failed");
  return GBG_OK;
}

gbg_status gbg_server_stats(gbg_server* s, double* p50_ms, double* p99_ms, double* jobs_per_s,
                            unsigned long long* completed) {
  if (!s) return fail(GBG_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto q = [&](double f) {
    if (l.empty()) return 0.0;
    size_t i = (size_t)std::ceil(f * l.size());
    return l[i == 0 ? 0 : std::min(i, l.size()) - 1];
  };
  if (p50_ms) *p50_ms = q(0.50);
  if (p99_ms) *p99_ms = q(0.99);
  if (completed) *completed = l.size();
  if (jobs_per_s) {
    const double span = s->any ? std::chrono::duration<double>(s->last_done - s->first_submit).count() : 0;
    *jobs_per_s = span > 0 ? l.size() / span : 0.0;
  }
  return GBG_OK;
}

gbg_status gbg_server_resident(gbg_server* s, size_t* bytes, int* plans) {
  if (!s) return fail(GBG_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (bytes) *bytes = s->resident;
  if (plans) *plans = (int)s->plans.size();
  return GBG_OK;
}

gbg_status gbg_server_destroy(gbg_server* s) {
  if (!s) return GBG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->jobs) delete kv.second;
    s->jobs.clear();
  }
  for (auto& kv : s->plans) gbg_plan_destroy(kv.second.plan);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  delete s;
  return GBG_OK;
}

}  // extern "C"
