// Ensembles of independent instances of mixed grid sizes (C ABI s2t_batch_*).
//
// Members of the same size form one class: a Solver whose `batch` slots share buffers, so each
// fused kernel covers every member of the class in one launch (blockIdx.y = slot). Each class has
// its own stream; a step of all classes is captured once per buffer phase as a hipGraph whose
// class branches fork from, and join back to, the batch's main stream, so classes run
// concurrently and a replay needs no host work. dt, stepIndex and the forcing phases live in the
// per-member device scalars, so a replayed graph needs no parameter updates.
// Grouped mode (the default): the classes of sizes 256, 512 and 1024 are stepped together by
// grouped launches (one launch per row or column pass covering every instance of those classes,
// on the main stream), the other classes on their branches as before. On the 64-member mix of
// 256^2, 512^2 and 1024^2 members it runs 32,100-32,200 Mpts-steps/s against 31,100-31,200 for
// the per-class branches (400-step runs, alternated, on three boxes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/s2t.h"
#include "core.hpp"

namespace s2t {

extern thread_local std::string g_err;
int api_bad(const std::string& m);

struct Batch {
  std::vector<Solver*> cls;                 // one per distinct grid size
  std::vector<std::pair<int, int>> member;  // member -> (class, slot)
  hipStream_t main = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  bool use_graphs = true;
  bool grouped = true;
  bool captured = false;
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
};

static void drop_graphs(Batch* b) {
  for (int p = 0; p < 3; ++p) {
    if (b->exec[p]) (void)hipGraphExecDestroy(b->exec[p]);
    if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
    b->exec[p] = nullptr;
    b->graph[p] = nullptr;
  }
  b->captured = false;
}

static void batch_free(Batch* b) {
  if (!b) return;
  if (b->main) (void)hipStreamSynchronize(b->main);
  drop_graphs(b);
  for (Solver* s : b->cls) solver_free(s);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->main) (void)hipStreamDestroy(b->main);
  delete b;
}

// all queued work of the batch (graphs on main, eager work on the class streams) is done
static void batch_quiesce(Batch* b) {
  hipchk(hipStreamSynchronize(b->main), "hipStreamSynchronize");
  for (Solver* s : b->cls) sync(s);
}

// the classes stepped by grouped launches (grouped mode: the grouped sizes), and the others
static void split_classes(const Batch* b, std::vector<Solver*>& grp, std::vector<Solver*>& rest) {
  grp.clear();
  rest.clear();
  for (Solver* s : b->cls) {
    if (b->grouped && group_size(s->n) && grp.size() < 4)
      grp.push_back(s);
    else
      rest.push_back(s);
  }
  // the largest size first in the work list: its longer workgroups start first and the short
  // ones of the small sizes fill the tail of the launch
  std::sort(grp.begin(), grp.end(), [](const Solver* a, const Solver* b) { return a->n > b->n; });
}

// one step of every class, each on its own stream (grouped classes: on main, after their own
// streams have drained, and finished before returning)
static void step_classes(Batch* b) {
  std::vector<Solver*> grp, rest;
  split_classes(b, grp, rest);
  for (Solver* s : rest) step_once(s, false);
  if (!grp.empty()) {
    for (Solver* s : grp) {
      ensure_primed(s);
      sync(s);
    }
    group_step(grp.data(), (int)grp.size(), b->main);
    hipchk(hipStreamSynchronize(b->main), "hipStreamSynchronize");
  }
}

static void capture(Batch* b) {
  drop_graphs(b);
  for (Solver* s : b->cls) {
    ensure_primed(s);
    sync(s);
  }
  const int phase0 = b->cls[0]->phase;
  std::vector<Solver*> grp, rest;
  split_classes(b, grp, rest);
  for (int k = 0; k < 3; ++k) {
    const int p = (phase0 + k) % 3;
    for (Solver* s : b->cls) s->phase = p;
    hipchk(hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      hipchk(hipEventRecord(b->fork, b->main), "hipEventRecord");
      for (size_t c = 0; c < rest.size(); ++c) {
        Solver* s = rest[c];
        hipchk(hipStreamWaitEvent(s->stream, b->fork, 0), "hipStreamWaitEvent");
        step_once(s, false);
        hipchk(hipEventRecord(b->join[c], s->stream), "hipEventRecord");
      }
      if (!grp.empty()) group_step(grp.data(), (int)grp.size(), b->main);
      for (size_t c = 0; c < rest.size(); ++c)
        hipchk(hipStreamWaitEvent(b->main, b->join[c], 0), "hipStreamWaitEvent");
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->main, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hipchk(hipStreamEndCapture(b->main, &b->graph[p]), "hipStreamEndCapture");
    hipchk(hipGraphInstantiate(&b->exec[p], b->graph[p], nullptr, nullptr, 0),
           "hipGraphInstantiate");
  }
  // capturing advanced the host phase counters three times: back where they started
  for (Solver* s : b->cls) s->phase = phase0;
  b->captured = true;
}

static void batch_step(Batch* b, int nsteps) {
  bool fresh = false;
  for (Solver* s : b->cls)
    for (char pr : s->primed_slot) fresh |= !pr;
  if (!b->use_graphs) {
    for (int i = 0; i < nsteps; ++i) step_classes(b);
    return;
  }
  if (fresh) {
    // prime the replaced members eagerly, then make main wait for it
    for (Solver* s : b->cls) {
      ensure_primed(s);
      sync(s);
    }
  }
  if (!b->captured) capture(b);
  for (int i = 0; i < nsteps; ++i) {
    const int p = b->cls[0]->phase;
    hipchk(hipGraphLaunch(b->exec[p], b->main), "hipGraphLaunch");
    for (Solver* s : b->cls) s->phase = (p + 1) % 3;
  }
}

}  // namespace s2t

using namespace s2t;

#define S2T_API_BEGIN try {
#define S2T_API_END                 \
  return S2T_OK;                    \
  }                                 \
  catch (const Fail& f) {           \
    g_err = f.what();               \
    return f.code;                  \
  }                                 \
  catch (const std::exception& e) { \
    g_err = e.what();               \
    return S2T_EHIP;                \
  }

#define BATCH_MEMBER(h, i)                                                          \
  Batch* b = reinterpret_cast<Batch*>(h);                                           \
  if (!b) return api_bad(std::string(__func__) + ": null batch");                   \
  if ((i) < 0 || (i) >= (int)b->member.size()) return api_bad("member out of range"); \
  Solver* s = b->cls[b->member[i].first];                                           \
  const int slot = b->member[i].second

extern "C" {

int s2t_batch_create(const int* sizes, int count, s2t_batch** out) {
  if (!out) return api_bad("s2t_batch_create: out is NULL");
  *out = nullptr;
  if (!sizes || count <= 0) return api_bad("s2t_batch_create: need at least one member");
  for (int i = 0; i < count; ++i) {
    const std::string e = size_error(sizes[i]);
    if (!e.empty()) return api_bad(e);
  }
  Batch* b = new Batch();
  S2T_API_BEGIN
  std::map<int, int> per_size;  // size -> members
  for (int i = 0; i < count; ++i) per_size[sizes[i]] += 1;
  std::map<int, int> cls_of, next_slot;
  for (auto& kv : per_size) {
    cls_of[kv.first] = (int)b->cls.size();
    b->cls.push_back(solver_new(kv.first, 1, kv.second, false));
    default_class_width(b->cls.back(), kv.second);
    next_slot[kv.first] = 0;
  }
  for (int i = 0; i < count; ++i) b->member.push_back({cls_of[sizes[i]], next_slot[sizes[i]]++});
  hipchk(hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking), "hipStreamCreate");
  hipchk(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
  b->join.resize(b->cls.size(), nullptr);
  for (auto& e : b->join) hipchk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  *out = reinterpret_cast<s2t_batch*>(b);
  return S2T_OK;
  }
  catch (const Fail& f) {
    g_err = f.what();
    batch_free(b);
    return f.code;
  }
  catch (const std::exception& e) {
    g_err = e.what();
    batch_free(b);
    return S2T_EHIP;
  }
}

int s2t_batch_destroy(s2t_batch* h) {
  batch_free(reinterpret_cast<Batch*>(h));
  return S2T_OK;
}

int s2t_batch_size(s2t_batch* h, int member, int* n) {
  BATCH_MEMBER(h, member);
  (void)slot;
  if (!n) return api_bad("n is NULL");
  *n = s->n;
  return S2T_OK;
}

int s2t_batch_configure(s2t_batch* h, int member, const s2t_params* p) {
  BATCH_MEMBER(h, member);
  if (!p) return api_bad("params is NULL");
  const std::string e = check_params(*p);
  if (!e.empty()) return api_bad(e);
  S2T_API_BEGIN
  batch_quiesce(b);
  configure_slot(s, slot, *p);
  S2T_API_END
}

int s2t_batch_seed_spectrum(s2t_batch* h, int member, double k0, double amplitude, uint32_t seed) {
  BATCH_MEMBER(h, member);
  if (!(k0 > 0)) return api_bad("k0 must be positive");
  S2T_API_BEGIN
  batch_quiesce(b);
  seed_spectrum_slot(s, slot, k0, amplitude, seed);
  S2T_API_END
}

int s2t_batch_seed_physical(s2t_batch* h, int member, const float* field) {
  BATCH_MEMBER(h, member);
  if (!field) return api_bad("field is NULL");
  S2T_API_BEGIN
  batch_quiesce(b);
  seed_physical_slot(s, slot, field);
  S2T_API_END
}

int s2t_batch_write_state(s2t_batch* h, int member, const float* spectrum) {
  BATCH_MEMBER(h, member);
  if (!spectrum) return api_bad("spectrum is NULL");
  S2T_API_BEGIN
  batch_quiesce(b);
  write_state_slot(s, slot, spectrum);
  S2T_API_END
}

int s2t_batch_read_state(s2t_batch* h, int member, float* spectrum) {
  BATCH_MEMBER(h, member);
  if (!spectrum) return api_bad("spectrum is NULL");
  S2T_API_BEGIN
  batch_quiesce(b);
  read_state_slot(s, slot, spectrum);
  S2T_API_END
}

int s2t_batch_diagnostics(s2t_batch* h, int member, s2t_diag* out, double* spectrum,
                          double* zonal) {
  BATCH_MEMBER(h, member);
  S2T_API_BEGIN
  batch_quiesce(b);
  diag_slot(s, slot, out, spectrum, zonal);
  S2T_API_END
}

int s2t_batch_use_graphs(s2t_batch* h, int enable) {
  Batch* b = reinterpret_cast<Batch*>(h);
  if (!b) return api_bad("s2t_batch_use_graphs: null batch");
  S2T_API_BEGIN
  batch_quiesce(b);
  b->use_graphs = enable != 0;
  if (!b->use_graphs) drop_graphs(b);
  S2T_API_END
}

int s2t_batch_use_grouped(s2t_batch* h, int enable) {
  Batch* b = reinterpret_cast<Batch*>(h);
  if (!b) return api_bad("s2t_batch_use_grouped: null batch");
  S2T_API_BEGIN
  batch_quiesce(b);
  b->grouped = enable != 0;
  drop_graphs(b);
  S2T_API_END
}

int s2t_batch_autotune(s2t_batch* h, int steps) {
  Batch* b = reinterpret_cast<Batch*>(h);
  if (!b) return api_bad("s2t_batch_autotune: null batch");
  if (steps < 1) return api_bad("s2t_batch_autotune: steps must be >= 1");
  S2T_API_BEGIN
  batch_quiesce(b);
  // each class with its own member count; captured steps hold the old widths
  for (Solver* s : b->cls) autotune_cols(s, steps, nullptr, nullptr, nullptr);
  drop_graphs(b);
  S2T_API_END
}

int s2t_batch_column_width(s2t_batch* h, int member, int* width) {
  BATCH_MEMBER(h, member);
  if (!width) return api_bad("s2t_batch_column_width: null width");
  S2T_API_BEGIN
  int w[MAX_COL_WIDTHS];
  s->ops->col_widths(w);
  *width = s->cols_c != 0 ? s->cols_c : w[0];
  S2T_API_END
}

int s2t_batch_step(s2t_batch* h, int nsteps) {
  Batch* b = reinterpret_cast<Batch*>(h);
  if (!b) return api_bad("s2t_batch_step: null batch");
  if (nsteps < 0) return api_bad("nsteps must be >= 0");
  S2T_API_BEGIN
  batch_step(b, nsteps);
  S2T_API_END
}

int s2t_batch_sync(s2t_batch* h) {
  Batch* b = reinterpret_cast<Batch*>(h);
  if (!b) return api_bad("s2t_batch_sync: null batch");
  S2T_API_BEGIN
  batch_quiesce(b);
  S2T_API_END
}

}  // extern "C"
