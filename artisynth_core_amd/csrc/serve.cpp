// Serving loop: step-solve requests from many concurrent simulation instances
// (the kkt_serve_* entry points of include/kkt_mi355x.h), built on the
// solver's own C ABI.
//
// - An instance is one simulation's KKT pattern.  It gets its own solver handle
//   (own HIP stream) on its first request; the analysis and the factor stay
//   resident in HBM between its requests.
// - Admission by an HBM budget: after the analysis (which holds only the
//   structure) the handle reports what it will hold once factored; that much is
//   reserved before the factor allocates its fronts, evicting idle resident
//   instances least recently used first: an idle instance's factor is released
//   first (its analysis stays, so its next request refactors without a new
//   analysis), and analysis-only instances are destroyed when that is not
//   enough.  An instance larger than the whole
//   budget fails its requests; one that fits only after busy instances finish
//   waits for them.  An FP64 refactor after stalled FP32 refinement is charged
//   when it happens.
// - Continuous batching: requests enter one FIFO queue; `nworkers` threads take
//   the oldest request whose instance is idle, so arrivals start as soon as a
//   worker and memory are free, and different instances' device work overlaps
//   on their streams.  An instance's requests run in submission order, one at a
//   time, on its own handle, so its results do not depend on the mix (an
//   evicted instance refactors on its next request).
// - A request is one KKT step with the iteration-limit part of
//   KKTSolver.factorAndSolve's policy (KKTSolver.java:930-1014): BiCGStab with
//   the resident (stale) factor while it converges within max_iter, else a
//   factor and a refined direct solve.  The policy's other part, refactoring
//   early when estimateOptimalCount (KKTSolver.java:889-922) says so, reads one
//   simulation's own step times; in a shared server a request's time includes
//   the other instances' work (a group solve is shared by all its requests), so
//   the loop does not apply it.
// - Group solves (group_runners > 0): requests of resident, factored, idle
//   instances are taken together, up to group_max at a time, by a group runner
//   thread and solved in one kkt_group_iterative_solve_device call for the
//   device-array requests and one kkt_group_iterative_solve call for the
//   host-array ones (every BiCGStab step one launch for all of them); a request whose
//   group solve does not converge goes to a worker for the factor and direct
//   solve, its instance still held.  Each instance's result is bitwise the one
//   its own iterative solve gives, so the mix still does not change it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kkt_mi355x.h"

namespace {

// Forward row groups per row block of every instance's solve plan.  A group
// solve launches all its instances' tree levels at once, tens of thousands of
// row blocks per launch, so fewer, larger blocks amortise each block's gather:
// on one stale-factor step of the 64-instance mix, 16.2 ms with 4 groups per
// block against 17.1 ms for the one-handle default, 16.2 ms with 2 and 17.2 ms
// with 8.  The solution is bitwise the same for any choice.
constexpr int kServeRowGroups = 4;

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Instance {
  int iid = 0;
  int n = 0, nnz = 0, mtype = KKT_MT_SYMMETRIC;
  std::vector<int> ro, ci;
  std::vector<double> vals0;
  kkt_solver* h = nullptr;
  long long charged = 0;   // bytes reserved against the budget
  bool busy = false;
  bool factored = false;
  bool too_large = false;
  double last_used = 0;
  long long requests = 0;
};

struct Request {
  long long id = 0;
  bool group_failed = false;    // its group solve did not converge: direct next
  int iid = 0;
  bool device = false;
  std::vector<double> vals, b;  // host submissions: copies
  const double* d_vals = nullptr;
  const double* d_b = nullptr;
  double* x = nullptr;          // caller's output (host or device)
  kkt_serve_result res{};
  bool done = false;
};

}  // namespace

struct kkt_server {
  kkt_serve_config cfg{};
  std::mutex mu;
  std::condition_variable cv;       // work available / instance became idle
  std::condition_variable done_cv;  // a request completed
  std::map<int, std::unique_ptr<Instance>> inst;
  std::deque<Request*> queue;
  std::deque<Request*> fallback;    // group solves that did not converge (instance held)
  long long inflight = 0;           // taken, not yet done
  int group_live = 0;               // group runners holding a group object
  int freeing = 0;                  // instances whose device memory is being freed
  std::map<long long, std::unique_ptr<Request>> reqs;
  std::vector<std::thread> workers;
  long long next_id = 1;
  long long used = 0, peak = 0;
  bool stop = false;
  double t0 = now_ms();
  kkt_serve_stats st{};
  double service_ms = 0;
  std::string err;
  bool has_err = false;

  int fail(int code, const std::string& m) {
    std::lock_guard<std::mutex> lk(mu);
    err = m;
    has_err = true;
    return code;
  }

  // a request the group runners take: its instance resident and factored
  bool group_eligible(const Request*, const Instance* in) const {
    return group_live > 0 && in->h && in->factored && !in->too_large;
  }

  // for a worker: a failed group solve's direct step first, else the oldest
  // queued request whose instance is idle and that no group runner takes
  // (keeps per-instance order: an instance's later requests wait behind its
  // first queued one)
  Request* next_locked() {
    if (!fallback.empty()) {
      Request* r = fallback.front();
      fallback.pop_front();
      return r;
    }
    std::vector<int> blocked;
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Request* r = *it;
      if (std::find(blocked.begin(), blocked.end(), r->iid) != blocked.end()) continue;
      Instance* in = inst[r->iid].get();
      if (in->busy || group_eligible(r, in)) {
        blocked.push_back(r->iid);
        continue;
      }
      queue.erase(it);
      inflight++;
      return r;
    }
    return nullptr;
  }

  // for a group runner: the first queued request of every idle instance, in
  // queue order, while group-eligible, at most group_max
  void next_group_locked(std::vector<Request*>& out) {
    out.clear();
    std::vector<int> blocked;
    for (auto it = queue.begin(); it != queue.end() && (int)out.size() < cfg.group_max;) {
      Request* r = *it;
      if (std::find(blocked.begin(), blocked.end(), r->iid) != blocked.end()) {
        ++it;
        continue;
      }
      blocked.push_back(r->iid);
      Instance* in = inst[r->iid].get();
      if (in->busy || !group_eligible(r, in)) {
        ++it;
        continue;
      }
      in->busy = true;
      out.push_back(r);
      it = queue.erase(it);
      inflight++;
    }
  }

  // Frees the device memory of one idle instance other than `self`, least
  // recently used first: an instance's factor goes first (kkt_release_factor:
  // its analysis stays resident, so coming back costs a factor, not a new
  // analysis); when no idle instance holds a factor, an analysis-only instance
  // is destroyed.  The victim is held busy and the lock released while the
  // device memory is freed (stream synchronisation and hipFree), so submissions
  // and other workers do not wait on it; while one free is in flight, other
  // callers wait for its result instead of freeing more.  Returns false when
  // no idle instance holds device memory.
  bool free_one(const Instance* self, std::unique_lock<std::mutex>& lk) {
    if (freeing > 0) {
      cv.wait_for(lk, std::chrono::milliseconds(2));
      return true;
    }
    Instance* victim = nullptr;
    for (auto& kv : inst) {
      Instance* in = kv.second.get();
      if (in == self || !in->h || in->busy || !in->factored) continue;
      if (!victim || in->last_used < victim->last_used) victim = in;
    }
    bool whole = victim == nullptr;
    if (whole) {
      for (auto& kv : inst) {
        Instance* in = kv.second.get();
        if (in == self || !in->h || in->busy) continue;
        if (!victim || in->last_used < victim->last_used) victim = in;
      }
    }
    if (!victim) return false;
    victim->busy = true;
    freeing++;
    kkt_solver* h = victim->h;
    lk.unlock();
    long long keep = 0;
    if (!whole) {
      kkt_info info{};
      if (kkt_release_factor(h) == KKT_OK && kkt_get_info(h, &info) == KKT_OK) keep = info.device_bytes;
      else whole = true;  // a failed release: drop the whole instance instead
    }
    if (whole) kkt_destroy(h);
    lk.lock();
    if (whole) {
      victim->h = nullptr;
      used -= victim->charged;
      victim->charged = 0;
      st.evictions++;
    } else {
      keep = std::min(victim->charged, std::max(0LL, keep));
      used -= victim->charged - keep;
      victim->charged = keep;
      st.releases++;
    }
    victim->factored = false;
    victim->busy = false;
    freeing--;
    cv.notify_all();
    return true;
  }

  // reserve `need` more bytes for instance `in` (freeing idle instances'
  // memory, or waiting for busy ones); false if it can never fit
  bool reserve(Instance* in, long long need, std::unique_lock<std::mutex>& lk) {
    if (in->charged + need > cfg.budget_bytes) return false;
    while (used + need > cfg.budget_bytes) {
      if (!free_one(in, lk)) cv.wait_for(lk, std::chrono::milliseconds(20));
    }
    used += need;
    in->charged += need;
    peak = std::max(peak, used);
    return true;
  }

  // analysis of a non-resident instance (structure only, no fronts yet)
  int admit(Instance* in, std::unique_lock<std::mutex>& lk, std::string& why) {
    lk.unlock();
    kkt_solver* h = kkt_create();
    int rc = h ? KKT_OK : KKT_ERR_INIT;
    if (!rc) rc = kkt_set_precision(h, cfg.precision);
    if (!rc) rc = kkt_set_refine_tol(h, cfg.refine_tol, 10);
    if (!rc) rc = kkt_set_pattern(h, in->n, in->nnz, in->ro.data(), in->ci.data(), in->mtype);
    if (!rc) rc = kkt_analyze_values(h, in->vals0.data());
    kkt_info info{};
    if (!rc) rc = kkt_get_info(h, &info);
    if (!rc) {
      std::vector<int> rg((size_t)std::max(info.nlevels, 1), kServeRowGroups);
      rc = kkt_set_forward_row_groups(h, rg.data());
    }
    if (rc) {
      const char* m = h ? kkt_last_error(h) : nullptr;
      why = std::string("admission of instance ") + std::to_string(in->iid) + " failed: " + (m ? m : "");
      if (h) kkt_destroy(h);
      lk.lock();
      return rc;
    }
    lk.lock();
    const long long need = info.factored_bytes;
    if (!reserve(in, need, lk)) {
      in->too_large = true;
      why = "instance " + std::to_string(in->iid) + " needs " + std::to_string(need) +
            " bytes of device memory, more than the whole budget of " + std::to_string(cfg.budget_bytes);
      lk.unlock();
      kkt_destroy(h);
      lk.lock();
      return KKT_ERR_ALLOC;
    }
    in->h = h;
    in->factored = false;
    st.admissions++;
    return KKT_OK;
  }

  // an instance whose factor was released: reserve the factor's memory again
  // before the request's factor allocates it
  int recharge(Instance* in, std::unique_lock<std::mutex>& lk, std::string& why) {
    kkt_info info{};
    lk.unlock();
    const int rc = kkt_get_info(in->h, &info);
    lk.lock();
    if (rc) {
      why = "instance " + std::to_string(in->iid) + ": kkt_get_info failed";
      return rc;
    }
    const long long need = info.factored_bytes - in->charged;
    if (need > 0 && !reserve(in, need, lk)) {
      in->too_large = true;
      why = "instance " + std::to_string(in->iid) + " needs " + std::to_string(info.factored_bytes) +
            " bytes of device memory, more than the whole budget of " + std::to_string(cfg.budget_bytes);
      return KKT_ERR_ALLOC;
    }
    return KKT_OK;
  }

  // one request on its (resident, busy) instance, lock released
  void run(Request* r, Instance* in) {
    kkt_solver* h = in->h;
    const double tol = std::pow(10.0, -cfg.tol_exp);
    int rc = -1;
    const double* vals = r->device ? r->d_vals : r->vals.data();
    const double* b = r->device ? r->d_b : r->b.data();
    if (in->factored && !r->group_failed) {
      rc = r->device ? kkt_iterative_solve_device(h, vals, b, r->x, tol, cfg.max_iter)
                     : kkt_iterative_solve(h, vals, b, r->x, tol, cfg.max_iter);
    }
    if (rc > 0) {
      r->res.kind = 1;
      r->res.iterations = rc;
      r->res.status = KKT_OK;
    } else {
      int st_ = r->device ? kkt_factor_device(h, vals) : kkt_factor(h, vals);
      if (!st_) st_ = r->device ? kkt_solve_device(h, b, r->x) : kkt_solve(h, b, r->x);
      r->res.kind = 2;
      r->res.iterations = 0;
      r->res.status = st_;
      in->factored = st_ == KKT_OK;
    }
    kkt_info info{};
    if (kkt_get_info(h, &info) == KKT_OK) r->res.rel_residual = info.last_rel_residual;
    if (r->res.status) {
      const char* m = kkt_last_error(h);
      snprintf(r->res.message, sizeof r->res.message, "%s", m ? m : "solve failed");
    }
  }

  void worker() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      Request* r = nullptr;
      cv.wait(lk, [&] { return (r = next_locked()) != nullptr || stop; });
      if (!r) return;  // stopping with nothing runnable
      Instance* in = inst[r->iid].get();
      in->busy = true;
      if (!r->group_failed) r->res.start_ms = now_ms() - t0;
      int rc = KKT_OK;
      std::string why;
      if (in->too_large) {
        rc = KKT_ERR_ALLOC;
        why = "instance " + std::to_string(in->iid) + " does not fit in the budget";
      } else if (!in->h) {
        rc = admit(in, lk, why);
      } else if (!in->factored && !r->group_failed) {
        rc = recharge(in, lk, why);
      }
      if (!rc) {
        lk.unlock();
        const double ts = now_ms();
        run(r, in);
        const double te = now_ms();
        kkt_info info{};
        const bool fell_back = kkt_get_info(in->h, &info) == KKT_OK && info.fp64_fallbacks > 0;
        lk.lock();
        service_ms += te - ts;
        // an FP64 refactor after a stall now holds FP64 fronts as well: charged
        // once it happened (the memory is already held), idle instances are
        // evicted until the budget holds again
        if (fell_back) {
          const long long extra = info.device_bytes - in->charged;
          if (extra > 0) {
            used += extra;
            in->charged += extra;
            peak = std::max(peak, used);
            while (used > cfg.budget_bytes && free_one(in, lk)) {
            }
          }
        }
      } else {
        r->res.status = rc;
        snprintf(r->res.message, sizeof r->res.message, "%s", why.c_str());
      }
      finish_locked(r, in);
    }
  }

  void finish_locked(Request* r, Instance* in) {
    r->res.done_ms = now_ms() - t0;
    in->busy = false;
    in->last_used = now_ms();
    in->requests++;
    st.completed++;
    if (r->res.status) st.errors++;
    else if (r->res.kind == 1) st.iterative++;
    else st.direct++;
    r->done = true;
    inflight--;
    done_cv.notify_all();
    cv.notify_all();
  }

  // group runner: solves what next_group_locked takes in one group call.  Its
  // group object (own stream) is made here; without one (no device) the runner
  // ends and the workers take every request.
  void group_runner() {
    kkt_group* g = kkt_group_create();
    if (!g) return;
    const double tol = std::pow(10.0, -cfg.tol_exp);
    std::vector<Request*> batch;
    std::vector<kkt_solver*> hs;
    std::vector<const double*> vs, bs;
    std::vector<double*> xs;
    std::vector<int> status;
    std::vector<double> relres;
    std::unique_lock<std::mutex> lk(mu);
    group_live++;
    for (;;) {
      batch.clear();
      cv.wait(lk, [&] {
        if (stop) return true;
        next_group_locked(batch);
        return !batch.empty();
      });
      if (batch.empty()) break;  // stopping
      const int K = (int)batch.size();
      // device-array requests first, then host-array ones: one group call each
      std::stable_partition(batch.begin(), batch.end(), [](const Request* r) { return r->device; });
      hs.resize(K);
      vs.resize(K);
      bs.resize(K);
      xs.resize(K);
      status.assign(K, 0);
      relres.assign(K, 0.0);
      const double ts = now_ms();
      int nd = 0;
      for (int k = 0; k < K; k++) {
        Request* r = batch[k];
        r->res.start_ms = ts - t0;
        hs[k] = inst[r->iid]->h;
        vs[k] = r->device ? r->d_vals : r->vals.data();
        bs[k] = r->device ? r->d_b : r->b.data();
        xs[k] = r->x;
        nd += r->device ? 1 : 0;
      }
      st.group_calls += (nd > 0) + (nd < K);
      st.group_requests += K;
      lk.unlock();
      if (nd > 0)
        (void)kkt_group_iterative_solve_device(g, nd, hs.data(), vs.data(), bs.data(), xs.data(), tol, cfg.max_iter,
                                               status.data());
      if (nd < K)
        (void)kkt_group_iterative_solve(g, K - nd, hs.data() + nd, vs.data() + nd, bs.data() + nd, xs.data() + nd,
                                        tol, cfg.max_iter, status.data() + nd);
      for (int k = 0; k < K; k++) {
        kkt_info info{};
        if (kkt_get_info(hs[k], &info) == KKT_OK) relres[k] = info.last_rel_residual;
      }
      const double te = now_ms();
      lk.lock();
      service_ms += K * (te - ts);
      for (int k = 0; k < K; k++) {
        Request* r = batch[k];
        Instance* in = inst[r->iid].get();
        if (status[k] > 0) {
          r->res.kind = 1;
          r->res.iterations = status[k];
          r->res.status = KKT_OK;
          r->res.rel_residual = relres[k];
          finish_locked(r, in);
        } else {
          r->group_failed = true;  // refactor + direct solve on a worker, instance still held
          fallback.push_back(r);
        }
      }
      cv.notify_all();
    }
    group_live--;
    lk.unlock();
    kkt_group_destroy(g);
  }
};

extern "C" {

void kkt_serve_default_config(kkt_serve_config* c) {
  if (!c) return;
  c->budget_bytes = 200LL << 30;
  c->nworkers = 12;  // best of 4-24 on the 64-instance mix with the default 4 hardware queues
  c->precision = KKT_PREC_FP32_REFINE;
  c->tol_exp = 10;
  c->max_iter = 30;
  c->refine_tol = 1e-10;
  c->group_runners = 2;
  c->group_max = 64;
}

kkt_server* kkt_serve_create(const kkt_serve_config* cfg) {
  if (!cfg || cfg->budget_bytes <= 0 || cfg->nworkers < 1 || cfg->nworkers > 64 || cfg->max_iter < 1 ||
      cfg->tol_exp < 1 || !(cfg->refine_tol > 0) ||
      (cfg->precision != KKT_PREC_FP64 && cfg->precision != KKT_PREC_FP32_REFINE) || cfg->group_runners < 0 ||
      cfg->group_runners > 8 || (cfg->group_runners > 0 && (cfg->group_max < 1 || cfg->group_max > KKT_GROUP_MAX)))
    return nullptr;
  kkt_server* s = nullptr;
  try {
    s = new kkt_server();
    s->cfg = *cfg;
    s->st.budget_bytes = cfg->budget_bytes;
    for (int w = 0; w < cfg->nworkers; w++) s->workers.emplace_back([s] { s->worker(); });
    for (int q = 0; q < cfg->group_runners; q++) s->workers.emplace_back([s] { s->group_runner(); });
  } catch (...) {
    delete s;
    return nullptr;
  }
  return s;
}

void kkt_serve_destroy(kkt_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // finish what was submitted, then stop the workers
    s->done_cv.wait(lk, [&] { return s->queue.empty() && s->fallback.empty() && s->inflight == 0; });
    s->stop = true;
    s->cv.notify_all();
  }
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->inst)
    if (kv.second->h) kkt_destroy(kv.second->h);
  delete s;
}

int kkt_serve_register(kkt_server* s, int iid, int n, int nnz, const int* rowOffs, const int* colIdxs,
                       const double* vals0, int mtype) {
  if (!s) return KKT_ERR_ARGUMENT;
  if (n <= 0 || nnz <= 0 || !rowOffs || !colIdxs || !vals0)
    return s->fail(KKT_ERR_ARGUMENT, "bad instance arrays");
  if (mtype != KKT_MT_SYMMETRIC && mtype != KKT_MT_SPD) return s->fail(KKT_ERR_UNKNOWN_MTYPE, "symmetric types only");
  if (rowOffs[0] != 0 || rowOffs[n] != nnz) return s->fail(KKT_ERR_PATTERN, "rowOffs must run from 0 to nnz");
  auto in = std::make_unique<Instance>();
  in->iid = iid;
  in->n = n;
  in->nnz = nnz;
  in->mtype = mtype;
  in->ro.assign(rowOffs, rowOffs + n + 1);
  in->ci.assign(colIdxs, colIdxs + nnz);
  in->vals0.assign(vals0, vals0 + nnz);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->inst.count(iid)) {
    s->err = "instance " + std::to_string(iid) + " is already registered";
    s->has_err = true;
    return KKT_ERR_ARGUMENT;
  }
  s->inst[iid] = std::move(in);
  s->has_err = false;
  return KKT_OK;
}

static long long submit(kkt_server* s, int iid, bool device, const double* vals, const double* b, double* x) {
  if (!s) return KKT_ERR_ARGUMENT;
  if (!vals || !b || !x) return s->fail(KKT_ERR_ARGUMENT, "null array");
  auto r = std::make_unique<Request>();
  r->iid = iid;
  r->device = device;
  r->x = x;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->inst.find(iid);
  if (it == s->inst.end()) {
    s->err = "instance " + std::to_string(iid) + " is not registered";
    s->has_err = true;
    return KKT_ERR_ARGUMENT;
  }
  const Instance* in = it->second.get();
  if (device) {
    r->d_vals = vals;
    r->d_b = b;
  } else {
    r->vals.assign(vals, vals + in->nnz);
    r->b.assign(b, b + in->n);
  }
  r->id = s->next_id++;
  r->res.submit_ms = now_ms() - s->t0;
  Request* rp = r.get();
  s->reqs[r->id] = std::move(r);
  s->queue.push_back(rp);
  s->st.submitted++;
  s->st.queue_max = std::max<long long>(s->st.queue_max, (long long)s->queue.size());
  // workers and group runners wait on the same condition and each takes only
  // its own kind of request: wake them all, or the one woken may not be able to
  // take it
  s->cv.notify_all();
  s->has_err = false;
  return rp->id;
}

long long kkt_serve_submit(kkt_server* s, int iid, const double* vals, const double* b, double* x) {
  return submit(s, iid, false, vals, b, x);
}

long long kkt_serve_submit_device(kkt_server* s, int iid, const double* d_vals, const double* d_b, double* d_x) {
  return submit(s, iid, true, d_vals, d_b, d_x);
}

int kkt_serve_wait(kkt_server* s, long long req, kkt_serve_result* out) {
  if (!s) return KKT_ERR_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) {
    s->err = "unknown request " + std::to_string(req);
    s->has_err = true;
    return KKT_ERR_ARGUMENT;
  }
  Request* r = it->second.get();
  s->done_cv.wait(lk, [&] { return r->done; });
  if (out) *out = r->res;
  const int rc = r->res.status;
  s->reqs.erase(it);
  return rc;
}

int kkt_serve_get_stats(kkt_server* s, kkt_serve_stats* out) {
  if (!s || !out) return KKT_ERR_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->bytes_resident = s->used;
  out->bytes_peak = s->peak;
  out->queue_depth = (long long)s->queue.size();
  int res = 0;
  for (auto& kv : s->inst) res += kv.second->h != nullptr;
  out->instances_resident = res;
  out->service_ms = s->service_ms;
  out->elapsed_ms = now_ms() - s->t0;
  return KKT_OK;
}

const char* kkt_serve_last_error(kkt_server* s) {
  if (!s || !s->has_err) return nullptr;
  return s->err.c_str();
}

}  // extern "C"
