// serve.cpp -- request loop with continuous batching for elasticity solves.
//
// Threads:
//  * the scheduler admits arrived requests (first come, first served, or the
//    smallest estimate first with an aging bound) while a slot is free and the
//    request's device-memory estimate fits in the budget,
//    and interleaves the CG chunks of all running requests: whenever the result
//    copy of a request's earlier chunk has landed it either retires the request
//    (device state says done, or the iteration budget is queued) or queues the
//    next chunk on the request's slot stream, keeping two chunks in flight
//    (plain stream launches: the scheduler is the only thread that launches
//    CG kernels, and no chunk graph exists for a worker to create or destroy
//    while it launches);
//  * setup workers turn an admitted request into device state on its slot's
//    context (structured operator built and assembled on the device, Jacobi,
//    initial residual) and retire finished requests (result,
//    true residual, solution to the host, device block returned).
// Each slot owns a femhip context with its own non-blocking stream, so a
// request's kernels, reductions and launch shapes are those of a solve alone.
// Device memory: the server reserves its HBM budget once, at creation, as one
// arena; admission hands a request a contiguous block of its estimate (first
// fit), its buffers are carved from that block while it is set up, and the
// block goes back to the arena when the request retires.  Each slot also owns
// the pinned result slots and events its requests' solvers use.  So no
// hipMalloc, hipFree, hipHostMalloc or hipHostFree runs while requests
// iterate; they run only at server creation and destruction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "femhip_internal.h"

using namespace femhip;

namespace {

enum ReqState { QUEUED, SETUP, RUNNING, RETIRING, DONE };

struct Req {
  int64_t ticket = 0;
  femhip_serve_request_t rq{};
  double arrival = 0.0, estimate = 0.0;
  std::atomic<int> state{QUEUED};
  int slot = -1;
  femhip_space_t *space = nullptr;
  femhip_operator_t *op = nullptr;
  femhip_solver_t *solver = nullptr;
  int cur = 0;
  int64_t queued_its = 0;
  double t_admit = 0.0, t_run = 0.0, t_done = 0.0;
  femhip_serve_result_t res{};
  std::vector<double> x;
  std::string error;   // message of a failed setup or retire
  size_t blk_off = 0, blk_size = 0;   // the request's block of the server's arena (size 0: none)
  DevArena arena;                     // bump allocator over that block
};
using ReqP = std::shared_ptr<Req>;

}  // namespace

struct femhip_server {
  femhip_ctx *ctx = nullptr;
  femhip_serve_params_t prm{};
  std::chrono::steady_clock::time_point epoch;
  std::vector<femhip_ctx *> slot_ctx;
  std::vector<int> free_slots;
  std::mutex mu;
  std::condition_variable cv_sched, cv_work, cv_done;
  std::map<int64_t, ReqP> reqs;              // every request not yet released
  std::multimap<std::pair<double, int64_t>, ReqP> pending;   // by (arrival, ticket)
  std::deque<ReqP> setup_q, ready_q, retire_q;
  int64_t next_ticket = 1, submitted = 0, completed = 0, failed = 0;
  int active = 0, peak_active = 0, running_peak = 0;
  double hbm_in_use = 0.0, hbm_peak = 0.0;
  bool stop = false;
  std::thread sched;
  std::vector<std::thread> workers;
  // device arena (the HBM budget, reserved once) and its free ranges by offset
  char *arena = nullptr;
  size_t arena_size = 0;
  std::map<size_t, size_t> free_blocks;
  int64_t arena_fallbacks = 0;                // allocations that did not fit a request's block
  std::vector<SolverHostSlots> slot_host;     // per slot: pinned result slots and events

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch).count();
  }
};

namespace {

// Device memory bound for a request on the structured stencil operator: 81
// FP64 + FP32 values per row, RHS, per-slice data, FP64 and FP32 Krylov
// vectors, block partials and the residual history (rows = 3 x the free
// vertices padded to a multiple of 64).
double estimate_bytes(int32_t n, int32_t maxit) {
  const double nn = n, nodes = nn * (nn + 1) * (nn + 1);
  const double rows = 3.0 * std::ceil(nodes / 64.0) * 64.0;
  const double values = 81.0 * rows * 12.0;
  const double slices = rows / 64.0;
  const double vecs = 8.0 * rows + 48.0 * rows + 20.0 * rows + 8.0 * std::max(1, maxit);
  return values + 24.0 * slices + vecs + (4 << 20);
}

constexpr size_t kBlockAlign = 256;

// first-fit block of `bytes` from the arena's free ranges (false: none fits)
bool arena_take(femhip_server &S, size_t bytes, size_t *off) {
  for (auto it = S.free_blocks.begin(); it != S.free_blocks.end(); ++it) {
    if (it->second < bytes) continue;
    *off = it->first;
    const size_t rest = it->second - bytes;
    const size_t at = it->first + bytes;
    S.free_blocks.erase(it);
    if (rest > 0) S.free_blocks[at] = rest;
    return true;
  }
  return false;
}

// give a block back, merged with its free neighbours
void arena_give(femhip_server &S, size_t off, size_t bytes) {
  auto next = S.free_blocks.lower_bound(off);
  if (next != S.free_blocks.end() && off + bytes == next->first) {
    bytes += next->second;
    next = S.free_blocks.erase(next);
  }
  if (next != S.free_blocks.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      prev->second += bytes;
      return;
    }
  }
  S.free_blocks[off] = bytes;
}

void free_request_objects(Req &r) {
  if (r.solver) femhip_solver_destroy(r.solver);
  if (r.op) femhip_operator_destroy(r.op);
  if (r.space) femhip_space_destroy(r.space);
  r.solver = nullptr;
  r.op = nullptr;
  r.space = nullptr;
}

void check_rc(int rc) {
  if (rc != FEMHIP_OK) throw Error{rc, femhip_last_error()};
}

// operator (built on the device), solver format and solver start on the
// request's slot context
void setup_request(femhip_server &S, Req &r) {
  femhip_ctx *c = S.slot_ctx[(size_t)r.slot];
  FH_HIP(hipSetDevice(c->device));
  femhip_physics_t ph{};
  ph.kind = FEMHIP_PHYS_ELASTICITY;
  ph.E = r.rq.E;
  ph.nu = r.rq.nu;
  for (int k = 0; k < 3; ++k) ph.body[k] = r.rq.body[k];
  {
    ArenaScope scope(r.blk_size ? &r.arena : nullptr);   // the request's buffers come from its block
    r.op = create_structured_elasticity(c, r.rq.n, ph);
    check_rc(femhip_solver_create(c, r.op, &S.prm.cg, &r.solver));
    solver_stream_launches(*r.solver);
    solver_use_host_slots(*r.solver, S.slot_host[(size_t)r.slot]);
    solver_start(*r.solver);
  }
  if (r.arena.misses) {
    std::lock_guard<std::mutex> lk(S.mu);
    S.arena_fallbacks += (int64_t)r.arena.misses;
  }
  FH_HIP(hipStreamSynchronize(c->stream));
  r.res.hbm_bytes = (double)(operator_device_bytes(*r.op) + solver_device_bytes(*r.solver));
  r.res.num_dofs = 3 * r.op->nodes;
}

// result, true residual and solution copy (node-interleaved: 3 node + c), then
// the device memory is freed
void retire_request(femhip_server &S, Req &r) {
  (void)S;
  femhip_cg_result_t cg{};
  solver_finish(*r.solver, &cg);
  std::vector<double> xi((size_t)r.op->n);
  solver_copy_solution(*r.solver, xi.data());
  const int32_t blk = r.op->blk, nodes = r.op->nodes;
  r.x.resize((size_t)r.res.num_dofs);
  for (int c = 0; c < 3; ++c)
    for (int32_t v = 0; v < nodes; ++v) r.x[(size_t)3 * v + c] = xi[(size_t)c * blk + v];
  r.res.status = cg.status;
  r.res.num_iterations = cg.num_iterations;
  r.res.outer_iterations = cg.outer_iterations;
  r.res.did_converge = cg.did_converge;
  r.res.true_residual = cg.true_residual;
  r.res.rhs_norm = cg.rhs_norm;
}

void finish_locked(femhip_server &S, Req &r, int status_if_failed) {
  r.t_done = S.now();
  if (status_if_failed != FEMHIP_OK) {
    r.res.status = status_if_failed;
    ++S.failed;
  }
  r.res.queue_ms = (r.t_admit - r.arrival) * 1e3;
  r.res.setup_ms = (r.t_run > 0 ? r.t_run - r.t_admit : 0.0) * 1e3;
  r.res.solve_ms = (r.t_run > 0 ? r.t_done - r.t_run : 0.0) * 1e3;
  r.res.latency_ms = (r.t_done - r.arrival) * 1e3;
  r.res.hbm_estimate = r.estimate;
  r.state = DONE;
  ++S.completed;
  S.free_slots.push_back(r.slot);
  if (r.blk_size) arena_give(S, r.blk_off, r.blk_size);   // the slot's stream is idle (worker_loop)
  r.blk_size = 0;
  S.hbm_in_use -= r.estimate;
  --S.active;
  S.cv_done.notify_all();
  S.cv_sched.notify_all();
}

void worker_loop(femhip_server &S) {
  (void)hipSetDevice(S.ctx->device);
  for (;;) {
    ReqP r;
    bool retire = false;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      // exit only once nothing admitted is left to set up or retire
      S.cv_work.wait(lk, [&] { return !S.retire_q.empty() || !S.setup_q.empty() || (S.stop && S.active == 0); });
      if (!S.retire_q.empty()) {
        r = S.retire_q.front();
        S.retire_q.pop_front();
        retire = true;
      } else if (!S.setup_q.empty()) {
        r = S.setup_q.front();
        S.setup_q.pop_front();
      } else {
        return;
      }
    }
    int status = FEMHIP_OK;
    std::string msg;
    {
      std::lock_guard<std::mutex> lk(S.mu);
      if (retire && !r->error.empty()) {   // the scheduler's launch or poll failed
        status = r->res.status;
        msg = r->error;
      }
    }
    try {
      if (status == FEMHIP_OK) {
        if (retire) retire_request(S, *r);
        else setup_request(S, *r);
      }
    } catch (const Error &e) {
      status = e.code;
      msg = e.msg;
    } catch (...) {
      status = FEMHIP_ERR_OOM;
      msg = "host allocation failed";
    }
    if (retire || status != FEMHIP_OK) {
      // nothing of the request may still run on its slot before its block is reused
      (void)hipStreamSynchronize(S.slot_ctx[(size_t)r->slot]->stream);
      free_request_objects(*r);
      std::lock_guard<std::mutex> lk(S.mu);
      if (status != FEMHIP_OK && r->error.empty()) r->error = std::string(retire ? "retire: " : "setup: ") + msg;
      finish_locked(S, *r, status);
    } else {
      std::lock_guard<std::mutex> lk(S.mu);
      S.ready_q.push_back(r);
      S.cv_sched.notify_all();
    }
  }
}

// The pending request to admit next among those arrived by t (end() if none):
// the earliest arrival (FEMHIP_ADMIT_FCFS), or the smallest memory estimate,
// ties by arrival (FEMHIP_ADMIT_SMALLEST; the estimate grows with the mesh,
// so small, short solves pass large ones), except that a request waiting
// aging_s or longer goes first so large requests are never starved.
std::multimap<std::pair<double, int64_t>, ReqP>::iterator next_admission(femhip_server &S, double t) {
  auto first = S.pending.begin();
  if (first == S.pending.end() || first->second->arrival > t) return S.pending.end();
  if (S.prm.admission != FEMHIP_ADMIT_SMALLEST) return first;
  if (S.prm.aging_s > 0 && t - first->second->arrival >= S.prm.aging_s) return first;
  auto best = first;
  for (auto it = first; it != S.pending.end() && it->second->arrival <= t; ++it)
    if (it->second->estimate < best->second->estimate) best = it;
  return best;
}

void scheduler_loop(femhip_server &S) {
  (void)hipSetDevice(S.ctx->device);
  const int chunk = solver_chunk();
  const int64_t maxit = S.prm.cg.max_num_iterations;
  std::vector<ReqP> running;
  for (;;) {
    std::vector<ReqP> started;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      if (S.stop && running.empty() && S.active == 0) return;
      // admissions: arrived requests by the policy while a slot and memory are free
      const double t = S.now();
      while (!S.stop && !S.pending.empty()) {
        auto it = next_admission(S, t);
        if (it == S.pending.end()) break;
        ReqP r = it->second;
        if (S.free_slots.empty() || S.hbm_in_use + r->estimate > S.prm.hbm_budget_bytes) break;
        if (S.arena) {   // a contiguous block of the estimate, else wait as for the budget
          const size_t need = ((size_t)r->estimate + kBlockAlign - 1) / kBlockAlign * kBlockAlign;
          size_t off = 0;
          if (!arena_take(S, need, &off)) break;
          r->blk_off = off;
          r->blk_size = need;
          r->arena = DevArena{S.arena + off, need, 0, 0};
        }
        S.pending.erase(it);
        r->slot = S.free_slots.back();
        S.free_slots.pop_back();
        r->state = SETUP;
        r->t_admit = t;
        S.hbm_in_use += r->estimate;
        S.hbm_peak = std::max(S.hbm_peak, S.hbm_in_use);
        ++S.active;
        S.peak_active = std::max(S.peak_active, S.active);
        S.setup_q.push_back(r);
        S.cv_work.notify_one();
      }
      while (!S.ready_q.empty()) {
        started.push_back(S.ready_q.front());
        S.ready_q.pop_front();
      }
      if (running.empty() && started.empty()) {
        // nothing on the device: sleep until work arrives or the next arrival time
        double wait_s = 0.05;
        if (!S.pending.empty()) wait_s = std::min(wait_s, std::max(0.0, S.pending.begin()->second->arrival - t));
        S.cv_sched.wait_for(lk, std::chrono::duration<double>(wait_s));
        continue;
      }
    }
    for (auto &r : started) {
      try {
        solver_push(*r->solver, 0);
        r->queued_its = chunk;
        if (r->queued_its < maxit) {
          solver_push(*r->solver, 1);
          r->queued_its += chunk;
        }
        r->cur = 0;
        r->state = RUNNING;
        r->t_run = S.now();
        running.push_back(r);
      } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(S.mu);
        r->res.status = e.code;
        r->error = "launch: " + e.msg;
        r->state = RETIRING;
        S.retire_q.push_back(r);
        S.cv_work.notify_one();
      }
    }
    {
      std::lock_guard<std::mutex> lk(S.mu);
      S.running_peak = std::max(S.running_peak, (int)running.size());
    }
    bool progress = !started.empty();
    for (size_t i = 0; i < running.size();) {
      Req &r = *running[i];
      int st = 0;
      try {
        st = solver_poll(*r.solver, r.cur);
      } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(S.mu);
        r.res.status = e.code;
        r.error = "poll: " + e.msg;
        st = 2;
      }
      if (st == 0) {
        ++i;
        continue;
      }
      progress = true;
      if (st == 2 || r.queued_its >= maxit) {
        std::lock_guard<std::mutex> lk(S.mu);
        r.state = RETIRING;
        S.retire_q.push_back(running[i]);
        S.cv_work.notify_one();
        running[i] = running.back();
        running.pop_back();
        continue;
      }
      try {
        solver_push(*r.solver, r.cur);   // the slot's previous result has been read
        r.queued_its += chunk;
        r.cur ^= 1;
      } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(S.mu);
        r.res.status = e.code;
        r.error = "launch: " + e.msg;
        r.state = RETIRING;
        S.retire_q.push_back(running[i]);
        S.cv_work.notify_one();
        running[i] = running.back();
        running.pop_back();
        continue;
      }
      ++i;
    }
    if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace

extern "C" {

double femhip_serve_estimate_bytes(int32_t n, int32_t max_num_iterations) {
  return estimate_bytes(n, max_num_iterations);
}

int femhip_serve_create(femhip_ctx_t *ctx, const femhip_serve_params_t *params, femhip_server_t **server) {
  try {
    FH_CHECK(ctx && params && server, FEMHIP_ERR_INVALID, "null argument");
    const femhip_serve_params_t &p = *params;
    FH_CHECK(p.max_active >= 1 && p.max_active <= 64, FEMHIP_ERR_INVALID, "max_active must be in [1, 64]");
    FH_CHECK(p.setup_threads >= 1 && p.setup_threads <= 64, FEMHIP_ERR_INVALID, "setup_threads must be in [1, 64]");
    FH_CHECK(p.hbm_budget_bytes > 0, FEMHIP_ERR_INVALID, "hbm_budget_bytes must be positive");
    FH_CHECK(p.cg.precision == FEMHIP_PRECISION_FP32_REFINED, FEMHIP_ERR_INVALID, "the server runs FP32_REFINED");
    FH_CHECK(p.cg.max_num_iterations >= 1 && p.cg.rtol >= 0 && p.cg.atol >= 0, FEMHIP_ERR_INVALID,
             "bad solver parameters");
    FH_CHECK(p.cg.stopping_criterion == FEMHIP_RES_RES || p.cg.stopping_criterion == FEMHIP_RES_RHS,
             FEMHIP_ERR_INVALID, "FP32_REFINED supports the RES_RES and RES_RHS criteria");
    FH_CHECK(p.cg.preconditioner == FEMHIP_PREC_IDENTITY || p.cg.preconditioner == FEMHIP_PREC_JACOBI,
             FEMHIP_ERR_INVALID, "unknown preconditioner");
    FH_CHECK(p.cg.inner_rtol >= 0 && p.cg.inner_rtol < 1, FEMHIP_ERR_INVALID, "reliable-update factor in [0,1)");
    FH_CHECK(p.admission == FEMHIP_ADMIT_FCFS || p.admission == FEMHIP_ADMIT_SMALLEST, FEMHIP_ERR_INVALID,
             "unknown admission policy");
    FH_CHECK(p.aging_s >= 0, FEMHIP_ERR_INVALID, "aging_s must be >= 0");
    FH_HIP(hipSetDevice(ctx->device));
    auto *S = new femhip_server();
    S->ctx = ctx_retain(ctx);
    S->prm = p;
    S->epoch = std::chrono::steady_clock::now();
    try {
      // the arena: the budget, or 90 % of the free device memory if that is less
      // (the budget is then capped at the arena)
      size_t free_b = 0, total_b = 0;
      FH_HIP(hipMemGetInfo(&free_b, &total_b));
      const size_t gran = 2u << 20, cap = (size_t)(0.9 * (double)free_b) / gran * gran;
      size_t want = (size_t)std::ceil(std::min(p.hbm_budget_bytes, (double)cap) / (double)gran) * gran;
      want = std::min(want, cap);
      if (want >= gran && hipMalloc((void **)&S->arena, want) == hipSuccess) {
        S->arena_size = want;
        S->free_blocks[0] = want;
        S->prm.hbm_budget_bytes = std::min(S->prm.hbm_budget_bytes, (double)want);
      } else {
        S->arena = nullptr;   // requests then allocate their buffers one by one
        (void)hipGetLastError();
      }
      for (int i = 0; i < p.max_active; ++i) {
        femhip_ctx_t *c = nullptr;
        check_rc(femhip_ctx_create(ctx->device, &c));
        S->slot_ctx.push_back(c);
        S->free_slots.push_back(p.max_active - 1 - i);
        SolverHostSlots hs;
        for (int k = 0; k < 2; ++k) {
          FH_HIP(hipHostMalloc(&hs.h_res[k], solver_result_bytes(), hipHostMallocMapped | hipHostMallocCoherent));
          FH_HIP(hipEventCreateWithFlags(&hs.ev[k], hipEventDisableTiming));
        }
        FH_HIP(hipHostMalloc((void **)&hs.h_flag, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
        S->slot_host.push_back(hs);
      }
      S->sched = std::thread(scheduler_loop, std::ref(*S));
      for (int i = 0; i < p.setup_threads; ++i) S->workers.emplace_back(worker_loop, std::ref(*S));
    } catch (...) {
      femhip_serve_destroy(S);
      throw;
    }
    *server = S;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return FEMHIP_ERR_OOM;
  } catch (const std::exception &e) {   // e.g. std::system_error when a thread cannot start
    set_error(std::string("server start failed: ") + e.what());
    return FEMHIP_ERR_STATE;
  } catch (...) {
    set_error("server start failed");
    return FEMHIP_ERR_STATE;
  }
}

int femhip_serve_submit(femhip_server_t *S, const femhip_serve_request_t *rq, int64_t *ticket) {
  try {
    FH_CHECK(S && rq && ticket, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(structured_elasticity_fits(rq->n), FEMHIP_ERR_INVALID,
             "request needs 1 <= n <= 206 (the operator's 32-bit entry index)");
    FH_CHECK(rq->E > 0 && rq->nu > -1.0 && rq->nu < 0.5, FEMHIP_ERR_INVALID, "invalid E or nu");
    const double est = estimate_bytes(rq->n, S->prm.cg.max_num_iterations);
    FH_CHECK(est <= S->prm.hbm_budget_bytes, FEMHIP_ERR_ADMISSION,
             "request needs an estimated " + std::to_string((long long)est) +
                 " device bytes, more than the whole HBM budget");
    auto r = std::make_shared<Req>();
    r->rq = *rq;
    r->estimate = est;
    std::lock_guard<std::mutex> lk(S->mu);
    FH_CHECK(!S->stop, FEMHIP_ERR_STATE, "server is shutting down");
    r->ticket = S->next_ticket++;
    r->arrival = rq->arrival_s >= 0 ? rq->arrival_s : S->now();
    S->reqs[r->ticket] = r;
    S->pending.emplace(std::make_pair(r->arrival, r->ticket), r);
    ++S->submitted;
    *ticket = r->ticket;
    S->cv_sched.notify_all();
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_serve_wait(femhip_server_t *S, int64_t ticket, double timeout_s, femhip_serve_result_t *result) {
  try {
    FH_CHECK(S && result, FEMHIP_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    FH_CHECK(it != S->reqs.end(), FEMHIP_ERR_INVALID, "unknown or released ticket");
    ReqP r = it->second;
    auto done = [&] { return r->state == DONE; };
    if (timeout_s < 0) S->cv_done.wait(lk, done);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
      throw Error{FEMHIP_ERR_STATE, "request not finished within the timeout"};
    *result = r->res;
    set_error(r->error);   // femhip_last_error() names a failed request's cause
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_serve_get_solution(femhip_server_t *S, int64_t ticket, double *x) {
  try {
    FH_CHECK(S && x, FEMHIP_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    FH_CHECK(it != S->reqs.end(), FEMHIP_ERR_INVALID, "unknown or released ticket");
    FH_CHECK(it->second->state == DONE, FEMHIP_ERR_STATE, "request not finished");
    FH_CHECK(!it->second->x.empty(), FEMHIP_ERR_STATE, "request failed, no solution: " + it->second->error);
    std::memcpy(x, it->second->x.data(), sizeof(double) * it->second->x.size());
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_serve_release(femhip_server_t *S, int64_t ticket) {
  try {
    FH_CHECK(S, FEMHIP_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    FH_CHECK(it != S->reqs.end(), FEMHIP_ERR_INVALID, "unknown or released ticket");
    FH_CHECK(it->second->state == DONE, FEMHIP_ERR_STATE, "request not finished");
    S->reqs.erase(it);
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_serve_stats(femhip_server_t *S, femhip_serve_stats_t *st) {
  try {
    FH_CHECK(S && st, FEMHIP_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    st->submitted = S->submitted;
    st->completed = S->completed;
    st->failed = S->failed;
    st->queued = (int32_t)S->pending.size();
    st->active = S->active;
    st->peak_active = S->peak_active;
    st->running_peak = S->running_peak;
    st->hbm_in_use = S->hbm_in_use;
    st->hbm_peak = S->hbm_peak;
    st->arena_bytes = (double)S->arena_size;
    st->arena_fallbacks = S->arena_fallbacks;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

double femhip_serve_now(femhip_server_t *S) { return S ? S->now() : 0.0; }

int femhip_serve_destroy(femhip_server_t *S) {
  if (!S) return FEMHIP_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    for (auto &kv : S->pending) {   // never admitted: dropped
      kv.second->state = DONE;
      kv.second->res.status = FEMHIP_ERR_STATE;
    }
    S->pending.clear();
    S->cv_sched.notify_all();
    S->cv_done.notify_all();
  }
  if (S->sched.joinable()) S->sched.join();   // returns once every admitted request has retired
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->cv_work.notify_all();
  }
  for (auto &w : S->workers)
    if (w.joinable()) w.join();
  for (auto *c : S->slot_ctx) femhip_ctx_destroy(c);
  for (auto &hs : S->slot_host) {   // every request retired, every thread joined
    for (int k = 0; k < 2; ++k) {
      if (hs.h_res[k]) (void)hipHostFree(hs.h_res[k]);
      if (hs.ev[k]) (void)hipEventDestroy(hs.ev[k]);
    }
    if (hs.h_flag) (void)hipHostFree(hs.h_flag);
  }
  if (S->arena) (void)hipFree(S->arena);
  ctx_release(S->ctx);
  delete S;
  return FEMHIP_OK;
}

}  // extern "C"
