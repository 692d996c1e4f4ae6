// Continuous-batching serving loop (C ABI s2t_server_*).
//
// Requests (grid size, step count, scenario, parameters) arrive from any thread. A native worker
// thread admits them against a device-memory budget: the instances of one grid size live in
// chunks (a Solver of a fixed number of slots: one launch per fused kernel covers every slot of a
// chunk, idle slots exit at once); a request takes a free slot of a chunk of its size, or the
// worker allocates a new chunk if the budget has room, freeing idle chunks of other sizes when it
// does not. A request that fits nowhere yet waits; later requests that fit are admitted past it.
// Admission is enqueued on the chunk's stream with no host wait (the scenario field built on the
// device, the slot's parameters and budget set by one launch), behind the steps already queued.
// Chunks run on four streams of the server (lanes), by size: 2048 and larger, 1024, 512, and the
// smaller sizes; the device maps a process's streams onto a few hardware queues, and with a
// stream per chunk a large chunk's steps could wait in a shared queue behind a small chunk's
// chain of short launches. Each slot retires itself on the
// device after exactly its budgeted steps (the step epilogue), so a result does not depend on
// when the host polls or on which other requests share the chunk: it is bit-identical to the same
// request run alone. Completions record their latency and FP64 diagnostics (and, on request, the
// final spectrum); the server reports p50/p99 latency, aggregate throughput and memory use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/s2t.h"
#include "aux.hpp"
#include "core.hpp"

namespace s2t {

extern thread_local std::string g_err;
int api_bad(const std::string& m);

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------------------------
// server state
// ---------------------------------------------------------------------------------------------
struct Request {
  uint64_t id;
  s2t_request req;
  double t_submit;
};

struct Record {
  int status = S2T_REQ_QUEUED;
  bool cancel = false;  // asked to stop while running; the worker frees the slot
  int n = 0, steps = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  s2t_diag diag{};
  std::vector<float> state;
  bool keep_state = false;
};

// A chunk: one Solver with `slots` instances of one grid size.
struct Chunk {
  Solver* s = nullptr;
  std::vector<int64_t> slot_req;  // request id per slot, -1 when free
  std::vector<int> slot_phase;    // chunk buffer phase when the slot's request was admitted
  std::vector<char> slot_fresh;   // admitted after this poll's clock copy was enqueued
  int active = 0;
  // running estimate of the time per step, from events on the chunk's lane stream: chunks of one
  // lane interleave their steps there, so each estimate includes the others' steps, and a lane's
  // work per poll stays near target_ms however many chunks share it
  double ms_per_step = 1.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  size_t bytes = 0;
  Inst* host_inst = nullptr;      // pinned copy of the slots' parameters and clocks, per poll
};

struct Server {
  std::map<int, std::vector<Chunk*>> cls;  // grid size -> chunks (worker thread only)
  std::map<int, int> chunk_slots;          // slots per chunk of a configured size
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> pending;
  std::unordered_map<uint64_t, Record> rec;
  std::thread worker;
  std::atomic<bool> stop{false};
  uint64_t next_id = 1;
  double target_ms = 4.0;         // GPU time per lane per poll (Chunk::ms_per_step)
  size_t budget = 0, used = 0, peak = 0;  // device bytes: limit (0 = none), in chunks now, most
  // statistics
  std::vector<double> latencies;
  double work_pts_steps = 0;      // sum n^2 * steps of completed requests
  double t_first = 0, t_last = 0;
  uint64_t submitted = 0, completed = 0, rejected = 0, cancelled = 0;
  std::string fatal;
  // failure injection (s2t_server_inject_failure = k: the worker fails at the admission after k
  // admissions), so that callers can test their handling of a failed server
  std::atomic<int> fail_after{-1};
  int admitted = 0;
  s2t_server_callback callback = nullptr;
  void* callback_user = nullptr;
  hipStream_t lanes[4] = {nullptr, nullptr, nullptr, nullptr};
};

// the lane of a chunk of size n
static int lane_of(int n) { return n >= 2048 ? 0 : n == 1024 ? 1 : n == 512 ? 2 : 3; }

// Completion callback (called without sv->mu held).
static void notify(Server* sv, uint64_t id, int status) {
  s2t_server_callback cb;
  void* user;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    cb = sv->callback;
    user = sv->callback_user;
  }
  if (cb) cb(id, status, user);
}

static size_t chunk_bytes(int n, int slots) {
  // device bytes of a chunk of `slots` instances of size n (solver_new's allocations, plus the
  // full-field scratch that seeding from a physical field allocates)
  const int K = n / 3, KC = K + 1, NJ = 2 * K + 1;
  const size_t pitch = (NJ + 63) / 64 * 64;
  const size_t kp = (KC + 15) / 16 * 16;
  const size_t state = (size_t)KC * pitch * 8, plane = (size_t)n * kp * 8;
  // per slot: 3 states, 5 mixed planes; per chunk: 3 full fields and one state of scratch
  return (size_t)slots * (3 * state + 5 * plane) + (size_t)n * n * 3 * 8 + state;
}

// Slots per chunk of a size without a configured class: enough instances that one launch covers
// about 2^22 points (a chunk of small grids fills the GPU), at most 32, at least 1.
static int default_chunk_slots(int n) {
  const long want = (1L << 22) / ((long)n * n);
  return (int)std::max(1L, std::min(32L, want));
}

static int slots_for(const Server* sv, int n) {
  auto it = sv->chunk_slots.find(n);
  return it != sv->chunk_slots.end() ? it->second : default_chunk_slots(n);
}

static Chunk* new_chunk(int n, int slots, hipStream_t lane) {
  Chunk* c = new Chunk();
  try {
    c->s = solver_new(n, 1, slots, false, lane);
    c->slot_req.assign(slots, -1);
    c->slot_phase.assign(slots, 0);
    c->slot_fresh.assign(slots, 0);
    c->bytes = chunk_bytes(n, slots);
    hipchk(hipEventCreate(&c->e0), "hipEventCreate");
    hipchk(hipEventCreate(&c->e1), "hipEventCreate");
    hipchk(hipHostMalloc((void**)&c->host_inst, sizeof(Inst) * slots, hipHostMallocDefault),
           "hipHostMalloc");
    // a new chunk starts with every slot idle
    for (int k = 0; k < slots; ++k) {
      const int zero = 0;
      hipchk(hipMemcpy(&c->s->inst[k].st.active, &zero, sizeof(int), hipMemcpyHostToDevice),
             "hipMemcpy");
    }
  } catch (...) {
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->host_inst) (void)hipHostFree(c->host_inst);
    solver_free(c->s);
    delete c;
    throw;
  }
  return c;
}

static void chunk_width(Chunk& c) { c.s->cols_c = class_width(c.s, c.active); }

static void free_chunk(Chunk* c) {
  if (!c) return;
  if (c->e0) (void)hipEventDestroy(c->e0);
  if (c->e1) (void)hipEventDestroy(c->e1);
  solver_free(c->s);  // waits for the chunk's stream, so no copy into host_inst is in flight
  if (c->host_inst) (void)hipHostFree(c->host_inst);
  delete c;
}

// Seed slot `slot` of chunk c for request r, which the caller has taken off the queue and marked
// running.
static void admit(Chunk& c, int slot, const Request& r) {
  Solver* s = c.s;
  const s2t_request& q = r.req;
  // enqueued behind the chunk's queued steps, with no host wait: the worker goes on admitting
  if (q.scenario == S2T_SCENARIO_TURBULENCE)
    seed_spectrum_slot_async(s, slot, q.k0 > 0 ? q.k0 : 14.0, q.amplitude > 0 ? q.amplitude : 6.0,
                             q.seed);
  else
    seed_scenario_slot_async(s, slot, q.scenario, q.scenario_param);
  admit_slot_async(s, slot, q.params, q.steps);
  c.slot_req[slot] = (int64_t)r.id;
  c.slot_phase[slot] = s->phase;
  c.active += 1;
}

// A chunk of size n with a free slot for a new request: an existing one, else a new one if the
// budget has room (freeing idle chunks of any size first when it does not); null when neither.
// Worker thread only.
static Chunk* place(Server* sv, int n) {
  for (Chunk* c : sv->cls[n])
    if (c->active < (int)c->slot_req.size()) return c;
  const int want = slots_for(sv, n);
  auto fits = [&](size_t bytes) { return sv->budget == 0 || sv->used + bytes <= sv->budget; };
  size_t bytes = chunk_bytes(n, want);
  int slots = want;
  if (!fits(bytes)) {
    // free idle chunks (no active slot) until the new chunk fits, largest first
    std::vector<std::pair<size_t, std::pair<int, Chunk*>>> idle;
    for (auto& kv : sv->cls)
      for (Chunk* c : kv.second)
        if (c->active == 0) idle.push_back({c->bytes, {kv.first, c}});
    std::sort(idle.begin(), idle.end(),
              [](const auto& a, const auto& b) { return a.first > b.first; });
    for (auto& it : idle) {
      if (fits(bytes)) break;
      auto& v = sv->cls[it.second.first];
      v.erase(std::find(v.begin(), v.end(), it.second.second));
      free_chunk(it.second.second);
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->used -= it.first;
    }
    if (!fits(bytes) && fits(chunk_bytes(n, 1))) {
      slots = 1;  // a one-slot chunk where the full one does not fit
      bytes = chunk_bytes(n, 1);
    }
    if (!fits(bytes)) return nullptr;
  }
  Chunk* c = new_chunk(n, slots, sv->lanes[lane_of(n)]);
  sv->cls[n].push_back(c);
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->used += c->bytes;
  sv->peak = std::max(sv->peak, sv->used);
  return c;
}

static void complete(Server* sv, Chunk& c, int slot) {
  const uint64_t id = (uint64_t)c.slot_req[slot];
  bool keep;
  int steps;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    keep = sv->rec[id].keep_state;
    steps = sv->rec[id].steps;
  }
  // the chunk kept rotating its buffers after this slot retired: its final state is where its
  // own last step left it
  const int phase = (c.slot_phase[slot] + steps) % 3;
  s2t_diag d{};
  diag_slot(c.s, slot, &d, nullptr, nullptr, phase);
  std::vector<float> st;
  if (keep) {
    st.resize((size_t)2 * c.s->n * c.s->n);
    read_state_slot(c.s, slot, st.data(), phase);
  }
  const double t = now_s();
  c.slot_req[slot] = -1;
  c.active -= 1;
  std::lock_guard<std::mutex> lk(sv->mu);
  Record& rc = sv->rec[id];
  rc.status = S2T_REQ_DONE;
  rc.t_done = t;
  rc.diag = d;
  rc.state.swap(st);
  sv->latencies.push_back(t - rc.t_submit);
  sv->work_pts_steps += (double)rc.n * rc.n * rc.steps;
  sv->completed += 1;
  sv->t_last = t;
}

// Stop a running request whose cancellation was asked for: its slot goes idle on the device and is
// freed. The flag is cleared on the chunk's stream, so it lands after the slot's admission (which
// may still be queued there) and before any later step or admission of the slot.
static void cancel_slot(Server* sv, Chunk& c, int slot) {
  const uint64_t id = (uint64_t)c.slot_req[slot];
  fill_u32(c.s->stream, &c.s->inst[slot].st.active, 1, 0u);
  c.slot_req[slot] = -1;
  c.active -= 1;
  std::lock_guard<std::mutex> lk(sv->mu);
  Record& rc = sv->rec[id];
  rc.status = S2T_REQ_CANCELLED;
  rc.t_done = now_s();
  sv->cancelled += 1;
}

static void worker_loop(Server* sv) {
  try {
    while (true) {
      // 1. the queue, or a wait for work
      std::vector<Request> queue;
      std::vector<uint64_t> dropped;  // queued requests cancelled by the shutdown
      {
        std::unique_lock<std::mutex> lk(sv->mu);
        if (sv->stop) {
          // shutdown: nothing more is admitted, and running requests stop between steps
          for (const Request& r : sv->pending) {
            Record& rc = sv->rec[r.id];
            rc.status = S2T_REQ_CANCELLED;
            rc.t_done = now_s();
            sv->cancelled += 1;
            dropped.push_back(r.id);
          }
          sv->pending.clear();
          for (auto& kv : sv->cls)
            for (Chunk* c : kv.second)
              for (int64_t id : c->slot_req)
                if (id >= 0) sv->rec[(uint64_t)id].cancel = true;
        }
        bool any_active = false;
        for (auto& kv : sv->cls)
          for (Chunk* c : kv.second) any_active |= c->active > 0;
        if (!any_active && sv->pending.empty()) {
          if (sv->stop) {
            lk.unlock();
            for (uint64_t id : dropped) notify(sv, id, S2T_REQ_CANCELLED);
            break;
          }
          sv->cv.wait_for(lk, std::chrono::milliseconds(2));
          continue;
        }
        queue.assign(sv->pending.begin(), sv->pending.end());
      }
      for (uint64_t id : dropped) notify(sv, id, S2T_REQ_CANCELLED);
      // cancellations of running requests, between steps (every chunk stream is idle here)
      std::vector<uint64_t> stopped;
      for (auto& kv : sv->cls)
        for (Chunk* c : kv.second)
          for (int slot = 0; slot < (int)c->slot_req.size(); ++slot) {
            if (c->slot_req[slot] < 0) continue;
            bool want;
            {
              std::lock_guard<std::mutex> lk(sv->mu);
              want = sv->rec[(uint64_t)c->slot_req[slot]].cancel;
            }
            if (want) {
              stopped.push_back((uint64_t)c->slot_req[slot]);
              cancel_slot(sv, *c, slot);
            }
          }
      for (uint64_t id : stopped) notify(sv, id, S2T_REQ_CANCELLED);
      // 2. step every chunk with active slots, each on its own stream (enqueued only). The steps
      // are enqueued round-robin over the chunks, one step of each in turn: streams share the
      // device's few hardware queues, and chunk after chunk a large chunk's steps waited in a
      // shared queue behind a small chunk's whole chain of short, latency-bound launches.
      std::vector<std::pair<Chunk*, int>> ran;
      int kmax = 0;
      for (auto& kv : sv->cls)
        for (Chunk* c : kv.second) {
          if (c->active == 0) continue;
          // the column-kernel width for the chunk's busy slots (idle slots exit at once): every
          // width gives the same state bit for bit
          chunk_width(*c);
          const int k = std::max(1, std::min(64, (int)std::lround(sv->target_ms / c->ms_per_step)));
          ensure_primed(c->s);
          hipchk(hipEventRecord(c->e0, c->s->stream), "hipEventRecord");
          ran.push_back({c, k});
          kmax = std::max(kmax, k);
        }
      for (int i = 0; i < kmax; ++i)
        for (auto& pr : ran)
          if (i < pr.second) step_once(pr.first->s, false);
      for (auto& pr : ran) {
        Chunk* c = pr.first;
        hipchk(hipEventRecord(c->e1, c->s->stream), "hipEventRecord");
        // every slot's clock in one copy, behind the steps
        hipchk(hipMemcpyAsync(c->host_inst, c->s->inst, sizeof(Inst) * c->s->batch,
                              hipMemcpyDeviceToHost, c->s->stream),
               "hipMemcpyAsync");
      }
      // 3. admission, FIFO, while the steps just enqueued run: each pending request takes a slot
      // wherever `place` finds one (its seeding is ordered behind the chunk's steps on the
      // chunk's stream); one that fits nowhere yet keeps its place in the queue
      std::vector<uint64_t> taken;
      for (const Request& r : queue) {
        {
          // cancelled while queued: s2t_server_cancel removed it from `pending`
          std::lock_guard<std::mutex> lk(sv->mu);
          if (sv->rec[r.id].status != S2T_REQ_QUEUED) continue;
        }
        Chunk* c = place(sv, r.req.n);
        if (!c) continue;
        const int fa = sv->fail_after.load();
        if (fa >= 0 && sv->admitted >= fa)
          throw Fail(S2T_EHIP, "worker failure injected by s2t_server_inject_failure");
        int slot = 0;
        while (c->slot_req[slot] >= 0) ++slot;
        {
          // taken off the queue and marked running together, so that a cancellation sees either
          // a queued request (and removes it) or a running one (and flags it for the worker)
          std::lock_guard<std::mutex> lk(sv->mu);
          Record& rc = sv->rec[r.id];
          if (rc.status != S2T_REQ_QUEUED) continue;  // cancelled since the check above
          for (auto p = sv->pending.begin(); p != sv->pending.end(); ++p)
            if (p->id == r.id) {
              sv->pending.erase(p);
              break;
            }
          rc.status = S2T_REQ_RUNNING;
          rc.t_start = now_s();
        }
        admit(*c, slot, r);
        c->slot_fresh[slot] = 1;
        sv->admitted += 1;
      }
      // 4. poll: per-slot device flags tell which requests retired
      for (auto& pr : ran) {
        Chunk& c = *pr.first;
        sync(c.s);
        float ms = 0.f;
        hipchk(hipEventElapsedTime(&ms, c.e0, c.e1), "hipEventElapsedTime");
        c.ms_per_step = 0.5 * c.ms_per_step + 0.5 * std::max(1e-3, (double)ms / pr.second);
        for (int slot = 0; slot < c.s->batch; ++slot) {
          if (c.slot_req[slot] < 0 || c.slot_fresh[slot]) continue;  // copied before its admission
          const DevState& d = c.host_inst[slot].st;
          if (!d.active) {
            const uint64_t id = (uint64_t)c.slot_req[slot];
            complete(sv, c, slot);
            notify(sv, id, S2T_REQ_DONE);
          }
        }
      }
      // the next poll's copies are enqueued after this iteration's admissions
      for (auto& kv : sv->cls)
        for (Chunk* c : kv.second) std::fill(c->slot_fresh.begin(), c->slot_fresh.end(), 0);
    }
  } catch (const std::exception& e) {
    // the worker stops: every request still queued, being admitted or running fails, and its
    // callback says so, so that no caller waits for it
    std::vector<uint64_t> failed;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->fatal = e.what();
      for (auto& kv : sv->rec)
        if (kv.second.status == S2T_REQ_QUEUED || kv.second.status == S2T_REQ_RUNNING) {
          kv.second.status = S2T_REQ_FAILED;
          kv.second.t_done = now_s();
          failed.push_back(kv.first);
        }
      sv->pending.clear();
    }
    std::sort(failed.begin(), failed.end());
    for (uint64_t id : failed) notify(sv, id, S2T_REQ_FAILED);
  }
}

static void server_free(Server* sv) {
  if (!sv) return;
  sv->stop = true;
  sv->cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (auto& kv : sv->cls)
    for (Chunk* c : kv.second) free_chunk(c);
  for (hipStream_t l : sv->lanes)
    if (l) (void)hipStreamDestroy(l);
  delete sv;
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

extern "C" {

int s2t_server_create(const int* sizes, const int* slots, int nclasses, double budget_bytes,
                      s2t_server** out) {
  if (!out) return api_bad("s2t_server_create: out is NULL");
  *out = nullptr;
  if (nclasses < 0 || (nclasses > 0 && (!sizes || !slots)))
    return api_bad("s2t_server_create: bad class list");
  if (!(budget_bytes >= 0)) return api_bad("s2t_server_create: budget must be >= 0");
  size_t need = 0;
  for (int i = 0; i < nclasses; ++i) {
    const std::string e = size_error(sizes[i]);
    if (!e.empty()) return api_bad(e);
    if (slots[i] <= 0) return api_bad("slot counts must be positive");
    for (int j = 0; j < i; ++j)
      if (sizes[j] == sizes[i]) return api_bad("duplicate class size");
    need += chunk_bytes(sizes[i], slots[i]);
  }
  if (budget_bytes > 0 && (double)need > budget_bytes)
    return api_bad("the first chunk of each class needs " + std::to_string(need) +
                   " bytes, over the budget of " + std::to_string((size_t)budget_bytes));
  Server* sv = new Server();
  S2T_API_BEGIN
  sv->budget = (size_t)budget_bytes;
  for (hipStream_t& l : sv->lanes)
    hipchk(hipStreamCreateWithFlags(&l, hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i < nclasses; ++i) {
    sv->chunk_slots[sizes[i]] = slots[i];
    Chunk* c = new_chunk(sizes[i], slots[i], sv->lanes[lane_of(sizes[i])]);
    sv->cls[sizes[i]].push_back(c);
    sv->used += c->bytes;
  }
  sv->peak = sv->used;
  sv->worker = std::thread(worker_loop, sv);
  *out = reinterpret_cast<s2t_server*>(sv);
  return S2T_OK;
  }
  catch (const Fail& f) {
    g_err = f.what();
    server_free(sv);
    return f.code;
  }
  catch (const std::exception& e) {
    g_err = e.what();
    server_free(sv);
    return S2T_EHIP;
  }
}

int s2t_server_destroy(s2t_server* h) {
  server_free(reinterpret_cast<Server*>(h));
  return S2T_OK;
}

int s2t_server_submit(s2t_server* h, const s2t_request* r, uint64_t* id) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv || !r || !id) return api_bad("s2t_server_submit: NULL argument");
  if (r->steps <= 0) return api_bad("steps must be positive");
  if (r->scenario < 0 || r->scenario > S2T_SCENARIO_MERGER) return api_bad("unknown scenario");
  const std::string e = check_params(r->params);
  if (!e.empty()) return api_bad(e);
  std::unique_lock<std::mutex> lk(sv->mu);
  if (!sv->fatal.empty()) return api_bad("server stopped: " + sv->fatal);
  const uint64_t rid = sv->next_id++;
  Record& rc = sv->rec[rid];
  rc.n = r->n;
  rc.steps = r->steps;
  rc.t_submit = now_s();
  rc.keep_state = r->keep_state != 0;
  sv->submitted += 1;
  if (sv->t_first == 0) sv->t_first = rc.t_submit;
  // any supported size is admitted against the budget; a size none of whose instances fits in
  // the budget at all is rejected
  const bool reject = !size_error(r->n).empty() ||
                      (sv->budget > 0 && chunk_bytes(r->n, 1) > sv->budget);
  if (reject) {
    rc.status = S2T_REQ_REJECTED;
    sv->rejected += 1;
  } else {
    sv->pending.push_back({rid, *r, rc.t_submit});
    sv->cv.notify_all();
  }
  *id = rid;
  if (reject) {
    lk.unlock();
    notify(sv, rid, S2T_REQ_REJECTED);
  }
  return S2T_OK;
}

int s2t_server_cancel(s2t_server* h, uint64_t id) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv) return api_bad("s2t_server_cancel: null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->rec.find(id);
  if (it == sv->rec.end()) return api_bad("unknown request id");
  Record& rc = it->second;
  if (rc.status == S2T_REQ_QUEUED) {
    for (auto p = sv->pending.begin(); p != sv->pending.end(); ++p)
      if (p->id == id) {
        sv->pending.erase(p);
        break;
      }
    rc.status = S2T_REQ_CANCELLED;
    rc.t_done = now_s();
    sv->cancelled += 1;
    lk.unlock();
    notify(sv, id, S2T_REQ_CANCELLED);
    return S2T_OK;
  }
  if (rc.status == S2T_REQ_RUNNING) {
    rc.cancel = true;  // the worker stops it between steps and calls back
    return S2T_OK;
  }
  g_err = "request already finished";
  return S2T_ESTATE;
}

int s2t_server_inject_failure(s2t_server* h, int after_admissions) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv) return api_bad("s2t_server_inject_failure: null server");
  sv->fail_after = after_admissions < 0 ? -1 : after_admissions;
  return S2T_OK;
}

int s2t_server_set_callback(s2t_server* h, s2t_server_callback cb, void* user) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv) return api_bad("s2t_server_set_callback: null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->callback = cb;
  sv->callback_user = user;
  return S2T_OK;
}

int s2t_server_poll(s2t_server* h, uint64_t id, int* status, s2t_diag* diag, double* latency_s) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv || !status) return api_bad("s2t_server_poll: NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (!sv->fatal.empty()) return api_bad("server stopped: " + sv->fatal);
  auto it = sv->rec.find(id);
  if (it == sv->rec.end()) return api_bad("unknown request id");
  *status = it->second.status;
  if (diag) *diag = it->second.diag;
  if (latency_s)
    *latency_s = it->second.status == S2T_REQ_DONE ? it->second.t_done - it->second.t_submit : -1.0;
  return S2T_OK;
}

int s2t_server_read_state(s2t_server* h, uint64_t id, float* spectrum) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv || !spectrum) return api_bad("s2t_server_read_state: NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->rec.find(id);
  if (it == sv->rec.end()) return api_bad("unknown request id");
  if (it->second.status != S2T_REQ_DONE || it->second.state.empty())
    return api_bad("request has no stored state (not done, or keep_state was 0)");
  memcpy(spectrum, it->second.state.data(), it->second.state.size() * sizeof(float));
  return S2T_OK;
}

int s2t_server_drain(s2t_server* h, double timeout_s) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv) return api_bad("s2t_server_drain: null server");
  const double t0 = now_s();
  while (true) {
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (!sv->fatal.empty()) return api_bad("server stopped: " + sv->fatal);
      if (sv->completed + sv->rejected + sv->cancelled == sv->submitted) return S2T_OK;
    }
    if (timeout_s > 0 && now_s() - t0 > timeout_s) {
      g_err = "drain timed out";
      return S2T_ESTATE;
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  }
}

int s2t_server_get_stats(s2t_server* h, s2t_server_stats* out) {
  Server* sv = reinterpret_cast<Server*>(h);
  if (!sv || !out) return api_bad("s2t_server_get_stats: NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  std::vector<double> l = sv->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    const double pos = p * (l.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - i;
    return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
  };
  out->submitted = sv->submitted;
  out->completed = sv->completed;
  out->rejected = sv->rejected;
  out->cancelled = sv->cancelled;
  out->p50_latency_s = pct(0.5);
  out->p99_latency_s = pct(0.99);
  out->mean_latency_s = l.empty() ? 0.0 : [&] { double t = 0; for (double x : l) t += x; return t / l.size(); }();
  const double span = sv->t_last - sv->t_first;
  out->mpts_steps_per_s = span > 0 ? sv->work_pts_steps / span / 1e6 : 0.0;
  out->budget_bytes = (double)sv->budget;
  out->used_bytes = (double)sv->used;
  out->peak_bytes = (double)sv->peak;
  return S2T_OK;
}

}  // extern "C"
