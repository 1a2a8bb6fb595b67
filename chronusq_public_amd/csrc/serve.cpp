// Serving loop: SCF jobs of different sizes on one GPU with continuous
// admission.  Jobs arrive at given times; the dispatcher admits the oldest
// waiting job as soon as an execution slot is free and its HBM footprint
// fits the budget next to the jobs already running; each admitted job runs
// its whole SCF (context, Schwarz, one-electron integrals, guess,
// iterations) on its own HIP stream, so the Fock builds, eigensolves and
// GEMMs of the jobs in flight overlap on the device.  A job's state
// (shell pairs, densities, DIIS history) stays resident until it finishes.
// The reference has no serving layer; each job is its SingleSlater SCF.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cqx.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Device bytes a job holds while it runs: the SCF's NB x NB buffers (about
// 40 of them with an 8-deep DIIS history), the J/K engine's per-call
// planes and copies, the bounded quartet lists and the pair data.
double job_footprint(int nb, int nshell, long long list_cap, bool mixed) {
  const double nb2 = (double)nb * nb;
  const double pairs = 0.5 * (double)nshell * (nshell + 1);
  return 8.0 * nb2 * 48.0 + 8.0 * (double)list_cap * (mixed ? 2.0 : 1.0) + pairs * (72.0 + 64.0 * 9.0) +
         (64 << 20);
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
}

int cqx_serve(const cqx_job *jobs, int njobs, const cqx_serve_opts *opts, cqx_job_result *out) {
  if (njobs < 0 || (njobs && (!jobs || !out)) || !opts || opts->max_inflight <= 0 || opts->list_cap < 1)
    return CQX_ERR_ARG;
  const auto t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv;
  int inflight = 0, done = 0;
  double used = 0.0;
  std::vector<int> order(njobs);
  for (int i = 0; i < njobs; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
  std::vector<cqx_basis *> basis(njobs, nullptr);
  std::vector<double> foot(njobs, 0.0);
  std::vector<std::thread> threads;
  std::deque<int> waiting;
  for (int i = 0; i < njobs; ++i) {
    std::memset(&out[i], 0, sizeof(cqx_job_result));
    out[i].status = -1;
  }

  auto run_job = [&](int j) {
    const cqx_job &J = jobs[j];
    cqx_job_result &R = out[j];
    cqx_ctx *ctx = nullptr;
    cqx_scf *scf = nullptr;
    int rc = cqx_ctx_create(basis[j], opts->thresh, opts->device, &ctx);
    if (rc == CQX_OK) rc = cqx_set_option(ctx, "list_cap", (double)opts->list_cap);
    if (rc == CQX_OK && opts->graph) rc = cqx_set_option(ctx, "graph", 1.0);
    if (rc == CQX_OK) {
      std::vector<double> Z(J.Z, J.Z + J.natom);
      cqx_scf_opts so = opts->scf;
      so.nocc = J.nocc;
      rc = cqx_scf_create(ctx, J.natom, Z.data(), J.xyz, &so, &scf);
    }
    if (rc == CQX_OK) {
      cqx_scf_result r{};
      rc = cqx_scf_run(scf, &r);
      if (rc == CQX_OK && J.density_out) rc = cqx_scf_matrix(scf, CQX_SCF_DENSITY, J.density_out);
      if (rc == CQX_OK && J.fock_out) rc = cqx_scf_matrix(scf, CQX_SCF_FOCK, J.fock_out);
      R.e_total = r.e_total;
      R.iterations = r.iterations;
      R.converged = r.converged;
      R.ms_fock = r.ms_fock;
    }
    if (scf) cqx_scf_destroy(scf);
    if (ctx) cqx_ctx_destroy(ctx);
    std::lock_guard<std::mutex> g(mu);
    R.status = rc;
    R.t_done = seconds_since(t0);
    if (opts->verbose)
      std::fprintf(stderr, "[serve] job %d natom %d: status %d, %d iterations, E %.10f, latency %.3f s\n", j,
                   J.natom, rc, R.iterations, R.e_total, R.t_done - R.t_arrival);
    used -= foot[j];
    --inflight;
    ++done;
    cv.notify_all();
  };

  size_t next = 0;
  int rc_all = CQX_OK;
  {
    std::unique_lock<std::mutex> lk(mu);
    while (done < njobs) {
      const double now = seconds_since(t0);
      // arrivals: parse the basis on the host (cheap) to size the job
      while (next < order.size() && jobs[order[next]].arrival_s <= now) {
        const int j = order[next++];
        const cqx_job &J = jobs[j];
        out[j].t_arrival = J.arrival_s;
        int rc = cqx_basis_from_g94(J.g94, J.natom, J.Z, J.xyz, 0, &basis[j]);
        int nsh = 0, nb = 0, mp = 0, ml = 0;
        if (rc == CQX_OK) rc = cqx_basis_dims(basis[j], &nsh, &nb, &mp, &ml);
        if (rc != CQX_OK) {
          out[j].status = rc;
          out[j].t_admit = out[j].t_done = now;
          ++done;
          rc_all = rc;
          continue;
        }
        foot[j] = job_footprint(nb, nsh, opts->list_cap, opts->scf.fp32_cut > 0.0);
        out[j].footprint_bytes = foot[j];
        waiting.push_back(j);
      }
      // admission: oldest first, while a slot is free and the footprint fits
      // (a job larger than the whole budget still runs, alone)
      while (!waiting.empty() && inflight < opts->max_inflight &&
             (inflight == 0 || used + foot[waiting.front()] <= opts->hbm_budget_bytes)) {
        const int j = waiting.front();
        waiting.pop_front();
        used += foot[j];
        ++inflight;
        out[j].t_admit = seconds_since(t0);
        threads.emplace_back(run_job, j);
      }
      if (done >= njobs) break;
      if (next < order.size()) {
        const double wait = jobs[order[next]].arrival_s - seconds_since(t0);
        if (wait > 0)
          cv.wait_for(lk, std::chrono::duration<double>(wait));
      } else {
        cv.wait(lk);
      }
    }
  }
  for (auto &t : threads) t.join();
  for (auto *b : basis)
    if (b) cqx_basis_destroy(b);
  for (int i = 0; i < njobs && rc_all == CQX_OK; ++i)
    if (out[i].status != CQX_OK) rc_all = out[i].status;
  return rc_all;
}

}  // extern "C"
