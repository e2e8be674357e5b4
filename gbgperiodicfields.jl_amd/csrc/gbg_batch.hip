// Transform autotuner and the batched, graph-captured form.
//
// gbg_plan_autotune times, per axis, the candidate pass configurations --
// the register-resident pass (power-of-two lengths), and the mixed-radix LDS
// pass under several radix factorisations/orderings and pencils-per-workgroup
// widths -- on a scratch vector field and keeps the fastest (coordinate
// descent over the three axes, forward + inverse transform as the score);
// then, on grids the fused pipelines cover, the x pass's tile shape (row-pair
// units per workgroup), scored by one u x omega term.
//
// gbg_batch holds many independent items (a field on its own grid and its
// operation), spreads them over a few streams (largest first, to the least
// loaded stream), runs them once eagerly so every work area is allocated,
// then stream-captures all of them -- fork/join over the streams -- into one
// hipGraph that a single gbg_batch_launch replays.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "fft_core.hpp"

namespace gbg {

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
struct AxisChoice {
  AxisPlan ap;
  int use_fast;
  int bsel;
};

static bool is_fast_len(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}

// factorisations of n into the implemented radices, greedy with a preferred
// largest power-of-two radix (8 or 4)
static bool factor_with(int n, int big, std::vector<int>& out) {
  out.clear();
  int m = n;
  for (int r : {big, 4, 2, 3, 5, 7, 11, 13}) {
    if (r > big && (r == 8 || r == 4)) continue;
    while (m % r == 0) {
      out.push_back(r);
      m /= r;
    }
  }
  return m == 1 && (int)out.size() <= kMaxStages;
}

static std::vector<AxisChoice> candidates(gbg_plan* p, int axis) {
  std::vector<AxisChoice> c;
  const int n = p->n[axis];
  AxisChoice base;
  base.ap = p->ax[axis];
  base.use_fast = 1;
  base.bsel = 0;
  if (is_fast_len(n)) c.push_back(base);  // register-resident pass
  // a chirp-z axis: the orders of its convolution length's radices
  const int len = pencil_len(base.ap);
  std::vector<std::vector<int>> orders;
  std::vector<int> f;
  for (int big : {8, 4}) {
    if (!factor_with(len, big, f)) continue;
    orders.push_back(f);
    std::vector<int> rev(f.rbegin(), f.rend());
    orders.push_back(rev);
    std::vector<int> oddfirst = f;
    std::stable_partition(oddfirst.begin(), oddfirst.end(), [](int r) { return r % 2 == 1; });
    orders.push_back(oddfirst);
  }
  std::sort(orders.begin(), orders.end());
  orders.erase(std::unique(orders.begin(), orders.end()), orders.end());
  const int cap = 65536 / (2 * p->elem);
  const int b0 = pick_b(len, cap);
  for (const auto& o : orders) {
    for (int b : {b0, b0 / 2, b0 / 4}) {
      if (b < 1) continue;
      AxisChoice ch;
      ch.ap.n = n;
      ch.ap.m = base.ap.m;
      ch.ap.nst = (int)o.size();
      for (int i = 0; i < ch.ap.nst; ++i) ch.ap.radix[i] = o[i];
      ch.use_fast = 0;
      ch.bsel = b;
      c.push_back(ch);
    }
  }
  return c;
}

static void apply(gbg_plan* p, int axis, const AxisChoice& ch) {
  p->ax[axis] = ch.ap;
  p->use_fast[axis] = ch.use_fast;
  p->bsel[axis] = ch.bsel;
}

static gbg_status time_roundtrip(gbg_plan* p, void* phys, void* spec, hipStream_t s, hipEvent_t e0,
                                 hipEvent_t e1, int reps, double* ms) {
  gbg_status st;
  for (int i = 0; i < 1; ++i) {  // warm-up
    if ((st = fwd_transform(p, spec, phys, 3, s)) || (st = inv_transform(p, phys, spec, spec, 3, s)))
      return st;
  }
  if ((st = check_hip(hipEventRecord(e0, s), "tune event"))) return st;
  for (int i = 0; i < reps; ++i) {
    if ((st = fwd_transform(p, spec, phys, 3, s)) || (st = inv_transform(p, phys, spec, spec, 3, s)))
      return st;
  }
  if ((st = check_hip(hipEventRecord(e1, s), "tune event"))) return st;
  if ((st = check_hip(hipEventSynchronize(e1), "tune sync"))) return st;
  float t = 0.f;
  if ((st = check_hip(hipEventElapsedTime(&t, e0, e1), "tune elapsed"))) return st;
  *ms = (double)t / reps;
  return GBG_OK;
}

gbg_status autotune(gbg_plan* p, hipStream_t s, double* best_ms) {
  const size_t pb = 3 * p->P * p->elem, sb = 3 * p->M * 2 * p->elem;
  void* phys = nullptr;
  void* spec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  gbg_status st = check_hip(hipMalloc(&phys, pb), "tune buffer");
  if (!st) st = check_hip(hipMalloc(&spec, sb), "tune buffer");
  if (!st) st = check_hip(hipMemsetAsync(phys, 0, pb, s), "tune init");
  if (!st) st = check_hip(hipEventCreate(&e0), "tune event");
  if (!st) st = check_hip(hipEventCreate(&e1), "tune event");
  const int reps = p->P >= (1u << 24) ? 2 : 5;
  double best = 1e30;
  if (!st) st = time_roundtrip(p, phys, spec, s, e0, e1, reps, &best);
  for (int axis = 0; axis < 3 && !st; ++axis) {
    const AxisChoice keep{p->ax[axis], p->use_fast[axis], p->bsel[axis]};
    AxisChoice win = keep;
    for (const AxisChoice& ch : candidates(p, axis)) {
      apply(p, axis, ch);
      double ms = 0;
      if ((st = time_roundtrip(p, phys, spec, s, e0, e1, reps, &ms))) break;
      if (ms < best) {
        best = ms;
        win = ch;
      }
    }
    apply(p, axis, st ? keep : win);
  }
  // the fused pipelines' x-pass shape: one u x omega term (2/3 cutoff of the
  // x axis) on a zero field in each shape, the faster kept
  if (!st && is_fast_len(p->n[0]) && is_fast_len(p->n[1]) && is_fast_len(p->n[2]) &&
      !p->generic_only) {
    void* nh = nullptr;
    st = check_hip(hipMalloc(&nh, sb), "tune buffer");
    const double kc = ((2.0 / 3.0) * (p->n[0] / 2.0)) * (2.0 * M_PI / p->L[0]);
    const int keep = p->x_var;
    int win = keep;
    double wbest = 1e30;
    for (int v = 0; v < 2 && !st; ++v) {
      p->x_var = v;
      if ((st = pipe_uxw(p, spec, nh, kc, 0, s))) break;   // warm-up (work area)
      if ((st = check_hip(hipEventRecord(e0, s), "tune event"))) break;
      for (int i = 0; i < reps && !st; ++i) st = pipe_uxw(p, spec, nh, kc, 0, s);
      if (st || (st = check_hip(hipEventRecord(e1, s), "tune event")) ||
          (st = check_hip(hipEventSynchronize(e1), "tune sync")))
        break;
      float t = 0.f;
      if ((st = check_hip(hipEventElapsedTime(&t, e0, e1), "tune elapsed"))) break;
      if (t < wbest) {
        wbest = t;
        win = v;
      }
    }
    p->x_var = st ? keep : win;
    if (nh) (void)hipFree(nh);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (phys) (void)hipFree(phys);
  if (spec) (void)hipFree(spec);
  if (best_ms) *best_ms = best;
  return st;
}

}  // namespace gbg

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------
struct gbg_batch_item {
  int kind;
  int n[3];
  double L[3];
  int elem;
  const void* in;
  void* out0;
  void* out1;
  double kc;
  int stream;
  gbg_plan* plan;
  double load;
};

struct gbg_batch {
  int nstreams;
  int tune;
  std::vector<hipStream_t> streams;
  std::vector<gbg_batch_item> items;
  std::map<std::tuple<int, int, int, double, double, double, int, int>, gbg_plan*> plans;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipEvent_t> events;
};

namespace gbg {

static gbg_status run_item(const gbg_batch_item& it, hipStream_t s) {
  gbg_plan* p = it.plan;
  gbg_status st;
  if (it.kind == GBG_ITEM_SCALAR_GRAD_PRODUCT) {
    if ((st = op_gradient(p, it.out0, it.in, 0, s))) return st;
    return pipe_dealiased_product(p, it.out1, it.in, it.in, it.kc, s);
  }
  if ((st = op_gradient(p, it.out0, it.in, 1, s))) return st;
  return pipe_uxw(p, it.in, it.out1, it.kc, 0, s);
}

static void destroy_batch(gbg_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (auto& kv : b->plans) gbg_plan_destroy(kv.second);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  delete b;
}

static gbg_status instantiate(gbg_batch* b, hipStream_t user) {
  gbg_status st;
  if (b->exec) return fail(GBG_E_ARG, "batch already instantiated");
  if (b->items.empty()) return fail(GBG_E_ARG, "empty batch");
  // streams: one capture origin plus nstreams workers, default priority
  for (int i = 0; i <= b->nstreams; ++i) {
    hipStream_t s = nullptr;
    if ((st = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch stream"))) return st;
    b->streams.push_back(s);
  }
  // largest items first onto the least loaded worker
  std::vector<int> order(b->items.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return b->items[x].load > b->items[y].load; });
  std::vector<double> acc(b->nstreams, 0.0);
  for (int i : order) {
    const int w = (int)(std::min_element(acc.begin(), acc.end()) - acc.begin());
    acc[w] += b->items[i].load;
    b->items[i].stream = w;
  }
  // plans per (grid, precision, worker); tuned once per (grid, precision)
  std::map<std::tuple<int, int, int, double, double, double, int>, gbg_plan*> tuned;
  for (auto& it : b->items) {
    auto key = std::make_tuple(it.n[0], it.n[1], it.n[2], it.L[0], it.L[1], it.L[2], it.elem, it.stream);
    auto f = b->plans.find(key);
    if (f == b->plans.end()) {
      gbg_plan* p = nullptr;
      if ((st = gbg_plan_create(it.n, it.L, it.elem, &p))) return st;
      b->plans[key] = p;
      auto tkey = std::make_tuple(it.n[0], it.n[1], it.n[2], it.L[0], it.L[1], it.L[2], it.elem);
      auto t = tuned.find(tkey);
      if (b->tune) {
        if (t == tuned.end()) {
          if ((st = autotune(p, b->streams[0], nullptr))) return st;
          tuned[tkey] = p;
        } else {
          for (int a = 0; a < 3; ++a) {
            p->ax[a] = t->second->ax[a];
            p->use_fast[a] = t->second->use_fast[a];
            p->bsel[a] = t->second->bsel[a];
          }
          p->x_var = t->second->x_var;
        }
      }
      it.plan = p;
    } else {
      it.plan = f->second;
    }
  }
  // eager pass: allocates every work area outside the capture
  for (auto& it : b->items)
    if ((st = run_item(it, b->streams[1 + it.stream]))) return st;
  for (hipStream_t s : b->streams)
    if ((st = check_hip(hipStreamSynchronize(s), "batch eager sync"))) return st;
  // capture: fork from the origin stream to the workers, join back
  for (int i = 0; i <= b->nstreams; ++i) {
    hipEvent_t e = nullptr;
    if ((st = check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event"))) return st;
    b->events.push_back(e);
  }
  hipStream_t origin = b->streams[0];
  if ((st = check_hip(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal), "begin capture")))
    return st;
  gbg_status cst = check_hip(hipEventRecord(b->events[0], origin), "fork");
  for (int w = 0; w < b->nstreams && !cst; ++w)
    cst = check_hip(hipStreamWaitEvent(b->streams[1 + w], b->events[0], 0), "fork wait");
  for (auto& it : b->items) {
    if (cst) break;
    cst = run_item(it, b->streams[1 + it.stream]);
  }
  for (int w = 0; w < b->nstreams && !cst; ++w) {
    cst = check_hip(hipEventRecord(b->events[1 + w], b->streams[1 + w]), "join");
    if (!cst) cst = check_hip(hipStreamWaitEvent(origin, b->events[1 + w], 0), "join wait");
  }
  hipGraph_t g = nullptr;
  st = check_hip(hipStreamEndCapture(origin, &g), "end capture");
  if (cst) {
    if (g) (void)hipGraphDestroy(g);
    return cst;
  }
  if (st) return st;
  b->graph = g;
  (void)user;
  return check_hip(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
}

}  // namespace gbg

using namespace gbg;

extern "C" {

gbg_status gbg_plan_autotune(gbg_plan* plan, void* stream, double* best_ms) {
  try {
    if (!plan) return fail(GBG_E_ARG, "null plan");
    return autotune(plan, as_stream(stream), best_ms);
  } catch (...) {
    return fail(GBG_E_DEVICE, "exception in autotune");
  }
}

gbg_status gbg_batch_create(int nstreams, int autotune_plans, gbg_batch** out) {
  try {
    if (!out) return fail(GBG_E_ARG, "null pointer");
    *out = nullptr;
    if (nstreams < 1 || nstreams > 16) return fail(GBG_E_ARG, "nstreams must be in 1..16");
    gbg_batch* b = new gbg_batch();
    b->nstreams = nstreams;
    b->tune = autotune_plans ? 1 : 0;
    *out = b;
    return GBG_OK;
  } catch (...) {
    return fail(GBG_E_OOM, "host allocation");
  }
}

gbg_status gbg_batch_add(gbg_batch* b, int kind, const int n[3], const double L[3], int elem_bytes,
                         const void* in, void* out0, void* out1, double kc) {
  try {
    if (!b || !n || !L || !in || !out0 || !out1) return fail(GBG_E_ARG, "null argument");
    if (b->exec) return fail(GBG_E_ARG, "batch already instantiated");
    if (kind != GBG_ITEM_SCALAR_GRAD_PRODUCT && kind != GBG_ITEM_VECTOR_GRAD_UXW)
      return fail(GBG_E_ARG, "unknown batch item kind");
    if (elem_bytes != 4 && elem_bytes != 8) return fail(GBG_E_ARG, "elem_bytes must be 4 or 8");
    if (!std::isfinite(kc) || kc < 0) return fail(GBG_E_ARG, "cutoff must be finite and >= 0");
    gbg_batch_item it{};
    it.kind = kind;
    for (int a = 0; a < 3; ++a) {
      if (n[a] < 1 || !(L[a] > 0)) return fail(GBG_E_ARG, "bad grid");
      it.n[a] = n[a];
      it.L[a] = L[a];
    }
    {
      // the refusals of gbg_gradient, gbg_uxw and gbg_dealiased_product: no
      // output may overlap the input or the other output
      const size_t cb = (size_t)(n[0] / 2 + 1) * n[1] * n[2] * 2 * elem_bytes;
      const size_t nin = kind == GBG_ITEM_VECTOR_GRAD_UXW ? 3 : 1;
      const char *i0 = (const char*)in, *o0 = (const char*)out0, *o1 = (const char*)out1;
      auto ov = [](const char* a, size_t na, const char* c, size_t nc) {
        return a < c + nc && c < a + na;
      };
      if (ov(o0, 3 * nin * cb, i0, nin * cb) || ov(o1, nin * cb, i0, nin * cb))
        return fail(GBG_E_ALIAS, "a batch item's output overlaps its input");
      if (ov(o0, 3 * nin * cb, o1, nin * cb))
        return fail(GBG_E_ALIAS, "a batch item's two outputs overlap");
    }
    it.elem = elem_bytes;
    it.in = in;
    it.out0 = out0;
    it.out1 = out1;
    it.kc = kc;
    it.load = (double)(n[0] / 2 + 1) * n[1] * n[2] * (kind == GBG_ITEM_VECTOR_GRAD_UXW ? 54.0 : 19.0);
    b->items.push_back(it);
    return GBG_OK;
  } catch (...) {
    return fail(GBG_E_OOM, "host allocation");
  }
}

gbg_status gbg_batch_instantiate(gbg_batch* b, void* stream) {
  try {
    if (!b) return fail(GBG_E_ARG, "null batch");
    return instantiate(b, as_stream(stream));
  } catch (...) {
    return fail(GBG_E_DEVICE, "exception in batch instantiate");
  }
}

gbg_status gbg_batch_launch(gbg_batch* b, void* stream) {
  if (!b) return fail(GBG_E_ARG, "null batch");
  if (!b->exec) return fail(GBG_E_ARG, "batch not instantiated");
  return check_hip(hipGraphLaunch(b->exec, as_stream(stream)), "graph launch");
}

gbg_status gbg_batch_destroy(gbg_batch* b) {
  destroy_batch(b);
  return GBG_OK;
}

}  // extern "C"
