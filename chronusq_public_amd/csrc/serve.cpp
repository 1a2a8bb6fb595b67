// Serving loop: SCF jobs of different sizes on one GPU with continuous
// admission.  Jobs arrive at given times; the dispatcher admits the oldest
// waiting job as soon as an execution slot is free and its HBM footprint
// fits the budget next to the jobs already running (a job whose footprint
// alone exceeds the budget is refused).  Each admitted job sets up its SCF
// (context, Schwarz bounds, one-electron integrals, guess) on a thread of
// its own and keeps its state (shell pairs, densities, DIIS history)
// resident on the device until it finishes.
//
// Two ways of running the iterations:
//  * per-job streams: each job iterates on its own context and stream, so
//    the Fock builds, eigensolves and GEMMs of the jobs in flight overlap on
//    the device;
//  * cross-request batching (opts->batch_fock): the Fock builds of the
//    jobs in flight go into one batched contraction per round, one launch
//    per class pair for all of them, over a context merged from the jobs'
//    own pair data (rebuilt only when the set of jobs changes); each job
//    then finishes its iteration (energy, convergence, CDIIS, eigensolve,
//    density) on its own stream, concurrently with the others.  A job that
//    finishes set-up joins the next round; a converged job leaves.  Jobs
//    are batched with jobs of similar size only (basis sizes within a
//    factor of two, kBucketBase below), each size class in rounds of its
//    own on a coordinator thread of its own, so that a small job does not
//    iterate at the pace of the largest one in flight.
// The reference has no serving layer; each job is its SingleSlater SCF.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cqx.h"
#include "ctx_access.hpp"
#include "kernels.hpp"
#include "scf_access.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Device bytes a job holds while it runs: the SCF's NB x NB buffers (about
// 40 of them with an 8-deep DIIS history), the J/K engine's per-call
// planes and copies, the bounded quartet lists and the pair data (twice
// when its Fock builds also go through a merged batched context).
double job_footprint(int nb, int nshell, long long list_cap, bool mixed, bool batched) {
  const double nb2 = (double)nb * nb;
  const double pairs = 0.5 * (double)nshell * (nshell + 1);
  return 8.0 * nb2 * 48.0 + 8.0 * (double)list_cap * (mixed ? 2.0 : 1.0) +
         pairs * (72.0 + 64.0 * 9.0) * (batched ? 2.0 : 1.0) + (64 << 20);
}

// size classes of cross-request batching: class b holds the jobs with
// kBucketBase 2^b <= NB < kBucketBase 2^(b+1) (smaller ones in class 0)
constexpr int kBucketBase = 96;
constexpr int kBuckets = 8;
int size_bucket(int nb) {
  int b = 0;
  while (b + 1 < kBuckets && nb >= (kBucketBase << (b + 1))) ++b;
  return b;
}

// State shared by the dispatcher, the job threads and (batched) the round
// coordinators; every field is guarded by mu.
struct Served {
  const cqx_job *jobs;
  int njobs;
  const cqx_serve_opts *opts;
  cqx_job_result *out;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  int inflight = 0, done = 0;
  double used = 0.0;
  std::vector<cqx_basis *> basis;
  std::vector<double> foot;
  // batched rounds
  std::vector<cqx_ctx *> ctx;
  std::vector<cqx_scf *> scf;
  std::vector<char> active;      // set up, taking part in the rounds
  std::vector<char> fock_ready;  // this round's J, K are built
  std::vector<char> post_done;   // this round's iteration is finished
  std::vector<char> leaving;     // finished (converged or failed) in this round
  std::vector<int> round_rc;     // status of this round's batched build for the job
  std::vector<double> round_ms;  // this round's build time
  std::vector<int> bucket;       // size class of the job
  bool stop = false;

  void finish(int j, int rc) {  // with mu held
    cqx_job_result &R = out[j];
    R.status = rc;
    R.t_done = seconds_since(t0);
    if (opts->verbose)
      std::fprintf(stderr, "[serve] job %d natom %d: status %d, %d iterations, E %.10f, latency %.3f s\n", j,
                   jobs[j].natom, rc, R.iterations, R.e_total, R.t_done - R.t_arrival);
    used -= foot[j];
    --inflight;
    ++done;
    cv.notify_all();
  }
};

// The job's results and requested matrices from its SCF.
int collect(Served &S, int j, cqx_scf *scf, int rc) {
  const cqx_job &J = S.jobs[j];
  cqx_job_result &R = S.out[j];
  cqx_scf_result r{};
  if (scf && cqx_scf_result_get(scf, &r) == CQX_OK) {
    R.e_total = r.e_total;
    R.iterations = r.iterations;
    R.converged = r.converged;
    R.ms_fock = r.ms_fock;
  }
  if (rc == CQX_OK && J.density_out) rc = cqx_scf_matrix(scf, CQX_SCF_DENSITY, J.density_out);
  if (rc == CQX_OK && J.fock_out) rc = cqx_scf_matrix(scf, CQX_SCF_FOCK, J.fock_out);
  return rc;
}

// Context and SCF of one job (its own stream), the guess density formed.
int setup_job(Served &S, int j, cqx_ctx **ctx, cqx_scf **scf) {
  const cqx_job &J = S.jobs[j];
  int rc = cqx_ctx_create(S.basis[j], S.opts->thresh, S.opts->device, ctx);
  if (rc == CQX_OK) rc = cqx_set_option(*ctx, "list_cap", (double)S.opts->list_cap);
  if (rc == CQX_OK && S.opts->graph && !S.opts->batch_fock) rc = cqx_set_option(*ctx, "graph", 1.0);
  if (rc == CQX_OK) {
    std::vector<double> Z(J.Z, J.Z + J.natom);
    cqx_scf_opts so = S.opts->scf;
    so.nocc = J.nocc;
    rc = cqx_scf_create(*ctx, J.natom, Z.data(), J.xyz, &so, scf);
  }
  return rc;
}

// Per-job streams: the whole SCF of job j on its own context.
void run_job(Served &S, int j) {
  cqx_ctx *ctx = nullptr;
  cqx_scf *scf = nullptr;
  int rc = setup_job(S, j, &ctx, &scf);
  if (rc == CQX_OK) {
    cqx_scf_result r{};
    rc = cqx_scf_run(scf, &r);
  }
  rc = collect(S, j, scf, rc);
  if (scf) cqx_scf_destroy(scf);
  if (ctx) cqx_ctx_destroy(ctx);
  std::lock_guard<std::mutex> g(S.mu);
  S.finish(j, rc);
}

// Cross-request batching, job side: set up, then for every round wait for
// the batched build of J and K and finish the iteration.
void run_job_batched(Served &S, int j) {
  cqx_ctx *ctx = nullptr;
  cqx_scf *scf = nullptr;
  int rc = setup_job(S, j, &ctx, &scf);
  if (rc == CQX_OK) {
    std::unique_lock<std::mutex> lk(S.mu);
    S.ctx[j] = ctx;
    S.scf[j] = scf;
    S.active[j] = 1;  // joins the next round
    S.cv.notify_all();
    for (;;) {
      S.cv.wait(lk, [&] { return S.fock_ready[j] || S.stop; });
      if (!S.fock_ready[j]) {  // stopped without a build
        rc = CQX_ERR_HIP;
        break;
      }
      S.fock_ready[j] = 0;
      const int brc = S.round_rc[j];
      const double ms = S.round_ms[j];
      lk.unlock();
      int fin = 0, iter = 0, max_iter = 0;
      rc = brc;
      if (rc == CQX_OK) rc = cqx::scf_after_jk(scf, ms, &fin);
      if (rc == CQX_OK) rc = cqx::scf_iterations(scf, &iter, &max_iter);
      if (rc == CQX_OK && !fin && iter >= max_iter) rc = CQX_ERR_NOT_CONVERGED;
      lk.lock();
      const bool leave = rc != CQX_OK || fin;
      if (leave) S.leaving[j] = 1;
      S.post_done[j] = 1;
      S.cv.notify_all();
      if (leave) break;
    }
  }
  rc = collect(S, j, scf, rc);
  {
    // the coordinator no longer touches this job's buffers once it has seen
    // it leave (it waits for every post_done of a round first)
    std::unique_lock<std::mutex> lk(S.mu);
    S.cv.wait(lk, [&] { return !S.active[j]; });
  }
  if (scf) cqx_scf_destroy(scf);
  if (ctx) cqx_ctx_destroy(ctx);
  std::lock_guard<std::mutex> g(S.mu);
  S.finish(j, rc);
}

// Cross-request batching, coordinator of size class b: one batched J/K
// build per round for every active job of the class, over a context merged
// from their pair data.
void coordinate(Served &S, int b) {
  (void)hipSetDevice(S.opts->device);
  cqx_ctx *merged = nullptr;
  std::vector<int> prev;
  std::unique_lock<std::mutex> lk(S.mu);
  for (;;) {
    S.cv.wait(lk, [&] {
      if (S.stop) return true;
      for (int j = 0; j < S.njobs; ++j)
        if (S.active[j] && S.bucket[j] == b) return true;
      return false;
    });
    if (S.stop) break;
    std::vector<int> P;
    for (int j = 0; j < S.njobs; ++j)
      if (S.active[j] && S.bucket[j] == b) P.push_back(j);
    std::vector<cqx_ctx *> parts;
    std::vector<const double *> D(P.size());
    std::vector<double *> J(P.size()), K(P.size());
    std::vector<int> nb(P.size());
    for (size_t i = 0; i < P.size(); ++i) {
      parts.push_back(S.ctx[P[i]]);
      cqx::scf_fock_buffers(S.scf[P[i]], &D[i], &J[i], &K[i], &nb[i]);
    }
    lk.unlock();
    // the merged context follows the set of jobs; a new set is merged anew
    int rc = CQX_OK;
    if (P != prev || !merged) {
      if (merged) cqx_ctx_destroy(merged);
      merged = nullptr;
      rc = cqx::ctx_merge(parts.data(), (int)parts.size(), &merged);
      if (rc == CQX_OK) rc = cqx_set_option(merged, "list_cap", (double)S.opts->list_cap * (double)P.size());
      if (rc == CQX_OK && S.opts->scf.fp32_cut > 0.0) rc = cqx_set_option(merged, "fp32_cut", S.opts->scf.fp32_cut);
      if (rc == CQX_OK && S.opts->graph) rc = cqx_set_option(merged, "graph", 1.0);
      prev = P;
      if (rc != CQX_OK && merged) {
        cqx_ctx_destroy(merged);
        merged = nullptr;
      }
    }
    const auto tb = Clock::now();
    if (rc == CQX_OK) {
      hipStream_t s = cqx::ctx_stream(merged);
      std::vector<cqx_contraction> list;
      for (size_t i = 0; i < P.size(); ++i) {
        const size_t n2 = (size_t)nb[i] * nb[i];
        if (cqx::launch_zero(J[i], n2, s) != hipSuccess || cqx::launch_zero(K[i], n2, s) != hipSuccess)
          rc = CQX_ERR_HIP;
        list.push_back({D[i], J[i], 1, CQX_COULOMB});
        list.push_back({D[i], K[i], 1, CQX_EXCHANGE});
      }
      if (rc == CQX_OK) rc = cqx_two_body_contract(merged, 1, (int)list.size(), list.data(), 0, CQX_DEVICE);
      if (rc != CQX_OK && S.opts->verbose)
        std::fprintf(stderr, "[serve] batched Fock build of %zu jobs failed: %s\n", P.size(), cqx_last_error(merged));
    }
    const double ms = 1e3 * seconds_since(tb);
    lk.lock();
    for (int j : P) {
      S.round_rc[j] = rc;
      S.round_ms[j] = ms;
      S.post_done[j] = 0;
      S.fock_ready[j] = 1;
    }
    S.cv.notify_all();
    S.cv.wait(lk, [&] {
      for (int j : P)
        if (!S.post_done[j]) return false;
      return true;
    });
    for (int j : P)
      if (S.leaving[j]) S.active[j] = 0;
    S.cv.notify_all();
  }
  lk.unlock();
  if (merged) cqx_ctx_destroy(merged);
}

}  // namespace

extern "C" {

void cqx_serve_default_opts(cqx_serve_opts *o) {
  if (!o) return;
  o->max_inflight = 8;
  o->hbm_budget_bytes = 200e9;
  o->device = 0;
  o->thresh = 1e-12;
  o->list_cap = 1ll << 25;
  cqx_scf_default_opts(&o->scf);
  o->verbose = 0;
  o->graph = 0;
  o->batch_fock = 0;
}

int cqx_serve(const cqx_job *jobs, int njobs, const cqx_serve_opts *opts, cqx_job_result *out) {
  if (njobs < 0 || (njobs && (!jobs || !out)) || !opts || opts->max_inflight <= 0 || opts->list_cap < 1)
    return CQX_ERR_ARG;
  Served S;
  S.jobs = jobs;
  S.njobs = njobs;
  S.opts = opts;
  S.out = out;
  S.t0 = Clock::now();
  S.basis.assign(njobs, nullptr);
  S.foot.assign(njobs, 0.0);
  S.ctx.assign(njobs, nullptr);
  S.scf.assign(njobs, nullptr);
  S.active.assign(njobs, 0);
  S.fock_ready.assign(njobs, 0);
  S.post_done.assign(njobs, 0);
  S.leaving.assign(njobs, 0);
  S.round_rc.assign(njobs, CQX_OK);
  S.round_ms.assign(njobs, 0.0);
  S.bucket.assign(njobs, 0);
  std::vector<int> order(njobs);
  for (int i = 0; i < njobs; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
  for (int i = 0; i < njobs; ++i) {
    std::memset(&out[i], 0, sizeof(cqx_job_result));
    out[i].status = -1;
  }
  const bool batched = opts->batch_fock != 0;
  std::vector<std::thread> threads;
  std::vector<std::thread> coordinators;
  if (batched)
    for (int b = 0; b < kBuckets; ++b) coordinators.emplace_back(coordinate, std::ref(S), b);
  std::deque<int> waiting;
  size_t next = 0;
  int rc_all = CQX_OK;
  {
    std::unique_lock<std::mutex> lk(S.mu);
    while (S.done < njobs) {
      const double now = seconds_since(S.t0);
      // arrivals: parse the basis on the host (cheap) to size the job
      while (next < order.size() && jobs[order[next]].arrival_s <= now) {
        const int j = order[next++];
        const cqx_job &J = jobs[j];
        out[j].t_arrival = J.arrival_s;
        int rc = cqx_basis_from_g94(J.g94, J.natom, J.Z, J.xyz, 0, &S.basis[j]);
        int nsh = 0, nb = 0, mp = 0, ml = 0;
        if (rc == CQX_OK) rc = cqx_basis_dims(S.basis[j], &nsh, &nb, &mp, &ml);
        if (rc == CQX_OK) {
          S.bucket[j] = size_bucket(nb);
          S.foot[j] = job_footprint(nb, nsh, opts->list_cap, opts->scf.fp32_cut > 0.0, batched);
          out[j].footprint_bytes = S.foot[j];
          if (S.foot[j] > opts->hbm_budget_bytes) rc = CQX_ERR_MEM;  // refused: beyond the whole budget
        }
        if (rc != CQX_OK) {
          out[j].status = rc;
          out[j].t_admit = out[j].t_done = now;
          ++S.done;
          rc_all = rc_all == CQX_OK ? rc : rc_all;
          if (opts->verbose)
            std::fprintf(stderr, "[serve] job %d refused: status %d (footprint %.3g B, budget %.3g B)\n", j, rc,
                         S.foot[j], opts->hbm_budget_bytes);
          continue;
        }
        waiting.push_back(j);
      }
      // admission: oldest first, while a slot is free and the footprint fits
      while (!waiting.empty() && S.inflight < opts->max_inflight &&
             S.used + S.foot[waiting.front()] <= opts->hbm_budget_bytes) {
        const int j = waiting.front();
        waiting.pop_front();
        S.used += S.foot[j];
        ++S.inflight;
        out[j].t_admit = seconds_since(S.t0);
        threads.emplace_back(batched ? run_job_batched : run_job, std::ref(S), j);
      }
      if (S.done >= njobs) break;
      if (next < order.size()) {
        const double wait = jobs[order[next]].arrival_s - seconds_since(S.t0);
        if (wait > 0) S.cv.wait_for(lk, std::chrono::duration<double>(wait));
      } else {
        S.cv.wait(lk);
      }
    }
    S.stop = true;
    S.cv.notify_all();
  }
  for (auto &t : threads) t.join();
  for (auto &t : coordinators) t.join();
  for (auto *b : S.basis)
    if (b) cqx_basis_destroy(b);
  for (int i = 0; i < njobs && rc_all == CQX_OK; ++i)
    if (out[i].status != CQX_OK) rc_all = out[i].status;
  return rc_all;
}

}  // extern "C"
