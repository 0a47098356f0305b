// server.cpp -- the serving loop: XC builds from many concurrent SCFs on one
// GPU, with continuous batching and admission by device-memory footprint.
//
// OCC runs one SCF per process and calls DFT::compute_vxc once per iteration
// (scf_impl.h:539 -> dft.h:395-429).  Here many SCFs share the device:
//   - an SCF is registered with its basis, grid and functional (inputs copied);
//     its context (grid leaves, AO values) is created when it is admitted,
//     i.e. when its footprint fits in the memory budget beside the resident
//     ones, and stays resident until the SCF is closed;
//   - builds are submitted with a host density matrix and run on `slots`
//     worker threads, each driving the context's own HIP stream, so builds
//     of different SCFs run on the device at the same time; a slot that
//     finishes takes the oldest runnable request at once (no batch barrier);
//   - builds of one SCF run one after another in submission order (they
//     share the context's buffers), as the SCF iterations they come from do;
//   - when the SCF at the head of the admission queue does not fit, idle
//     resident SCFs (no build pending or running) are evicted, least recently
//     used first; an evicted SCF is admitted again (AO values re-evaluated)
//     when its next build arrives.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "occxc.h"
#include "common.hpp"
#include "context.hpp"
#include "server_policy.hpp"

using namespace occxc;
using occxc::serving::ScfState;

namespace {

using Clock = std::chrono::steady_clock;

// Deep copy of the inputs of one SCF (the caller's arrays may go away before
// the SCF is admitted).
struct ScfInputs {
  std::vector<int> l, pure, nprim, ids;
  std::vector<double> origin, exps, coefs, points, weights;
  occxc_basis basis{};
  occxc_grid grid{};
  occxc_functional func{};
  occxc_options opt{};

  ScfInputs(const occxc_basis *b, const occxc_grid *g, const occxc_functional *f,
            const occxc_options *o) {
    OCCXC_REQUIRE(b && g && f && b->nshell > 0 && b->l && b->pure && b->nprim && b->origin &&
                      b->exponents && b->coefficients && f->ids && f->nfunc > 0 &&
                      g->npts >= 0 && (g->npts == 0 || (g->points && g->weights)),
                  OCCXC_ERR_INVALID_ARGUMENT, "server: null or empty basis, grid or functional");
    l.assign(b->l, b->l + b->nshell);
    pure.assign(b->pure, b->pure + b->nshell);
    nprim.assign(b->nprim, b->nprim + b->nshell);
    long long np = 0;
    for (int v : nprim) {
      OCCXC_REQUIRE(v > 0, OCCXC_ERR_INVALID_ARGUMENT, "basis: shell without primitives");
      np += v;
    }
    origin.assign(b->origin, b->origin + 3 * b->nshell);
    exps.assign(b->exponents, b->exponents + np);
    coefs.assign(b->coefficients, b->coefficients + np);
    points.assign(g->points, g->points + 3 * g->npts);
    weights.assign(g->weights, g->weights + g->npts);
    ids.assign(f->ids, f->ids + f->nfunc);
    basis = {b->nshell, l.data(), pure.data(), nprim.data(), origin.data(), exps.data(),
             coefs.data()};
    grid = {g->npts, points.data(), weights.data()};
    func = {(int)ids.size(), ids.data()};
    if (o) opt = *o;
    else occxc_default_options(&opt);
  }
};

struct Request {
  long long ticket;
  long long scf;
  std::vector<double> D;
  double *V, *exc, *nel;
  Clock::time_point arrival;
};

struct Ticket {
  bool done = false;
  occxc_status status = OCCXC_OK;
  std::string message;
  double latency_ms = 0.0;
};

struct Scf {
  long long id = 0;
  std::unique_ptr<ScfInputs> in;
  occxc_context *ctx = nullptr;
  ScfState state = ScfState::Waiting;
  size_t bytes = 0;       // device bytes held (resident) or needed (known after a try)
  bool busy = false;      // a build or the admission is running
  bool closing = false;   // closed by the caller: free once drained
  Clock::time_point last_used = Clock::now();
  std::deque<Request> pending;
  occxc_status admit_status = OCCXC_OK;
  std::string admit_message;
  DevBuf<double> d_D, d_V, d_r;  // per-SCF device staging of D, Vxc, {Exc, N}
  std::vector<double> h_r;
};

} // namespace

struct occxc_server {
  occxc_server_options opt{};
  size_t budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<long long, std::unique_ptr<Scf>> scfs;
  std::map<long long, Ticket> tickets;
  long long next_scf = 1, next_ticket = 1;
  size_t resident_bytes = 0;
  bool admitting = false;  // one admission at a time: its limit is exact
  bool stop = false;
  std::vector<std::thread> workers;
  occxc_server_stats stats{};
};

namespace {

Scf *next_admission(occxc_server *s) {
  return serving::next_admission(s->scfs, s->admitting, s->resident_bytes, s->budget);
}

Scf *next_eviction(occxc_server *s) {
  return serving::next_eviction(s->scfs, s->admitting, s->resident_bytes, s->budget);
}

Scf *next_build(occxc_server *s) { return serving::next_build(s->scfs); }

Scf *next_release(occxc_server *s) { return serving::next_release(s->scfs); }

void finish(occxc_server *s, const Request &r, occxc_status st, const std::string &msg) {
  Ticket &t = s->tickets[r.ticket];
  t.done = true;
  t.status = st;
  t.message = msg;
  t.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - r.arrival).count();
  if (st == OCCXC_OK) s->stats.builds_completed += 1;
  else if (st == OCCXC_ERR_CANCELLED) s->stats.builds_cancelled += 1;
  else s->stats.builds_failed += 1;
}

// Completes every queued (not started) request of an SCF as cancelled.
void cancel_pending(occxc_server *s, Scf *f) {
  while (!f->pending.empty()) {
    finish(s, f->pending.front(), OCCXC_ERR_CANCELLED, "build cancelled before it started");
    f->pending.pop_front();
  }
}

long long scf_nbf(const ScfInputs &in) {
  long long nbf = 0;
  for (int i = 0; i < in.basis.nshell; ++i) {
    const int l = in.l[i];
    nbf += in.pure[i] ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
  }
  return nbf;
}

// device bytes of an SCF's staging buffers (D, Vxc, {Exc, N_elec})
size_t staging_bytes(const ScfInputs &in) {
  const long long nbf = scf_nbf(in);
  return (size_t)(2 * nbf * nbf + 2) * sizeof(double);
}

void admit(occxc_server *s, Scf *f, std::unique_lock<std::mutex> &lk) {
  f->busy = true;
  f->state = ScfState::Admitting;
  s->admitting = true;
  const size_t limit = s->budget - s->resident_bytes;
  s->resident_bytes += f->bytes;  // reserve what is known so far
  const size_t reserved = f->bytes;
  lk.unlock();
  // the context may use what the staging buffers leave of the limit (a
  // limit of 0 would mean "no limit" to context_setup: keep it at least 1)
  const size_t staging = staging_bytes(*f->in);
  const size_t ctx_limit = limit > staging ? limit - staging : 1;
  auto c = new occxc_context();
  occxc_status st = OCCXC_OK;
  std::string msg;
  st = guarded([&] {
    try {
      context_setup(c, 1, &f->in->basis, &f->in->grid, &f->in->func, &f->in->opt, ctx_limit);
      DeviceGuard g(c->device);
      f->d_D.alloc((size_t)c->nbf * c->nbf);
      f->d_V.alloc((size_t)c->nbf * c->nbf);
      f->d_r.alloc(2);
      f->h_r.assign(2, 0.0);
    } catch (...) {
      if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = nullptr;
      }
      throw;
    }
  });
  if (st != OCCXC_OK) msg = occxc_last_error_message();
  const size_t need = c->device_bytes ? c->device_bytes + staging : 0;
  if (st != OCCXC_OK) {
    delete c;
    c = nullptr;
  }
  lk.lock();
  s->resident_bytes -= reserved;
  s->admitting = false;
  f->busy = false;
  if (st == OCCXC_OK) {
    f->ctx = c;
    f->bytes = need;
    f->state = ScfState::Resident;
    s->resident_bytes += need;
    s->stats.scfs_admitted += 1;
  } else if (st == OCCXC_ERR_OUT_OF_MEMORY && need > 0 && need <= s->budget &&
             s->resident_bytes > 0) {
    // does not fit beside the resident SCFs now: wait for memory
    f->bytes = need;
    f->state = ScfState::Waiting;
    s->stats.admissions_deferred += 1;
  } else {
    // cannot be served: fail its requests
    f->state = ScfState::Waiting;
    f->admit_status = st == OCCXC_ERR_OUT_OF_MEMORY ? OCCXC_ERR_OUT_OF_MEMORY : st;
    f->admit_message = need > s->budget ? "SCF needs more device memory than the server budget"
                                        : msg;
    while (!f->pending.empty()) {
      finish(s, f->pending.front(), f->admit_status, f->admit_message);
      f->pending.pop_front();
    }
  }
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

void build(occxc_server *s, Scf *f, std::unique_lock<std::mutex> &lk) {
  f->busy = true;
  Request r = std::move(f->pending.front());
  f->pending.pop_front();
  lk.unlock();
  occxc_context *c = f->ctx;
  const occxc_status st = guarded([&] {
    DeviceGuard g(c->device);
    const size_t nn = (size_t)c->nbf * c->nbf;
    f->d_D.upload(r.D.data(), nn, c->stream);
    launch_build(c, f->d_D.get(), f->d_V.get(), f->d_r.get(), c->stream);
    f->d_V.download(r.V, nn, c->stream);
    f->d_r.download(f->h_r.data(), 2, c->stream);
    OCCXC_HIP(hipStreamSynchronize(c->stream));
    *r.exc = f->h_r[0];
    *r.nel = f->h_r[1];
  });
  const std::string msg = st == OCCXC_OK ? std::string() : occxc_last_error_message();
  lk.lock();
  f->busy = false;
  f->last_used = Clock::now();
  finish(s, r, st, msg);
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

// Frees the context of an idle resident SCF; it stays registered.
void evict(occxc_server *s, Scf *f, std::unique_lock<std::mutex> &lk) {
  f->busy = true;
  occxc_context *c = f->ctx;
  lk.unlock();
  {
    DeviceGuard g(c->device);
    f->d_D.release();
    f->d_V.release();
    f->d_r.release();
    context_destroy(c);
  }
  lk.lock();
  f->ctx = nullptr;
  f->busy = false;
  f->state = ScfState::Waiting;
  s->resident_bytes -= f->bytes;
  s->stats.evictions += 1;
  s->cv_work.notify_all();
}

void release(occxc_server *s, Scf *f, std::unique_lock<std::mutex> &lk) {
  const long long id = f->id;
  occxc_context *c = f->ctx;
  const size_t held = f->state == ScfState::Resident ? f->bytes : 0;
  std::unique_ptr<Scf> own = std::move(s->scfs[id]);
  s->scfs.erase(id);
  lk.unlock();
  if (c) {
    DeviceGuard g(c->device);
    own->d_D.release();
    own->d_V.release();
    own->d_r.release();
    context_destroy(c);
  }
  own.reset();
  lk.lock();
  s->resident_bytes -= held;
  s->stats.scfs_closed += 1;
  s->cv_work.notify_all();
  s->cv_done.notify_all();
}

void worker(occxc_server *s) {
  (void)hipSetDevice(s->opt.device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    Scf *f = nullptr;
    s->cv_work.wait(lk, [&] {
      return s->stop || next_release(s) || next_build(s) || next_admission(s) ||
             next_eviction(s);
    });
    if ((f = next_release(s))) {
      release(s, f, lk);
      continue;
    }
    if ((f = next_build(s))) {
      build(s, f, lk);
      continue;
    }
    if ((f = next_admission(s))) {
      admit(s, f, lk);
      continue;
    }
    if ((f = next_eviction(s))) {
      evict(s, f, lk);
      continue;
    }
    if (s->stop) return;
  }
}

} // namespace

extern "C" {

void occxc_server_default_options(occxc_server_options *opt) {
  if (!opt) return;
  opt->device = 0;
  opt->slots = 4;
  opt->memory_budget = 0.0;
}

occxc_status occxc_server_create(const occxc_server_options *opt, occxc_server **srv) {
  return guarded([&] {
    OCCXC_REQUIRE(srv, OCCXC_ERR_INVALID_ARGUMENT, "srv: null pointer");
    *srv = nullptr;
    occxc_server_options o;
    occxc_server_default_options(&o);
    if (opt) o = *opt;
    OCCXC_REQUIRE(o.slots >= 1 && o.slots <= 64, OCCXC_ERR_INVALID_ARGUMENT,
                  "server: slots must be 1..64");
    OCCXC_REQUIRE(o.memory_budget >= 0.0, OCCXC_ERR_INVALID_ARGUMENT,
                  "server: negative memory budget");
    DeviceGuard g(o.device);
    size_t free_b = 0, total_b = 0;
    OCCXC_HIP(hipMemGetInfo(&free_b, &total_b));
    auto s = std::make_unique<occxc_server>();
    s->opt = o;
    s->budget = o.memory_budget > 0.0 ? (size_t)o.memory_budget : (size_t)(0.9 * free_b);
    for (int i = 0; i < o.slots; ++i) s->workers.emplace_back(worker, s.get());
    *srv = s.release();
  });
}

occxc_status occxc_server_destroy(occxc_server *srv) {
  return guarded([&] {
    if (!srv) return;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      // builds not yet started are cancelled, running ones finish, then
      // every SCF is released
      for (auto &kv : srv->scfs) {
        cancel_pending(srv, kv.second.get());
        kv.second->closing = true;
      }
      srv->cv_done.notify_all();
      srv->cv_work.notify_all();
      srv->cv_done.wait(lk, [&] { return srv->scfs.empty(); });
      srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers) t.join();
    delete srv;
  });
}

occxc_status occxc_server_open(occxc_server *srv, const occxc_basis *basis,
                               const occxc_grid *grid, const occxc_functional *func,
                               const occxc_options *opt, long long *scf) {
  return guarded([&] {
    OCCXC_REQUIRE(srv && scf, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    auto f = std::make_unique<Scf>();
    f->in = std::make_unique<ScfInputs>(basis, grid, func, opt);
    f->in->opt.device = srv->opt.device;
    std::lock_guard<std::mutex> lk(srv->mu);
    f->id = srv->next_scf++;
    *scf = f->id;
    srv->scfs[f->id] = std::move(f);
    srv->stats.scfs_opened += 1;
  });
}

occxc_status occxc_server_submit(occxc_server *srv, long long scf, const double *D,
                                 double *Vxc, double *exc, double *nelec, long long *ticket) {
  return guarded([&] {
    OCCXC_REQUIRE(srv && D && Vxc && exc && nelec && ticket, OCCXC_ERR_INVALID_ARGUMENT,
                  "null pointer");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->scfs.find(scf);
    OCCXC_REQUIRE(it != srv->scfs.end() && !it->second->closing, OCCXC_ERR_INVALID_ARGUMENT,
                  "server: unknown or closed SCF");
    Scf *f = it->second.get();
    // nbf from the basis inputs (the context may not exist yet)
    const long long nbf = scf_nbf(*f->in);
    Request r;
    r.ticket = srv->next_ticket++;
    r.scf = scf;
    r.D.assign(D, D + nbf * nbf);
    for (double v : r.D)
      OCCXC_REQUIRE(std::isfinite(v), OCCXC_ERR_INVALID_ARGUMENT, "D: non-finite entry");
    r.V = Vxc;
    r.exc = exc;
    r.nel = nelec;
    r.arrival = Clock::now();
    *ticket = r.ticket;
    srv->tickets[r.ticket] = Ticket{};
    srv->stats.builds_submitted += 1;
    if (f->admit_status != OCCXC_OK) {
      finish(srv, r, f->admit_status, f->admit_message);
    } else {
      f->pending.push_back(std::move(r));
      srv->cv_work.notify_one();
    }
  });
}

occxc_status occxc_server_wait(occxc_server *srv, long long ticket, double *latency_ms) {
  occxc_status result = OCCXC_OK;
  std::string msg;
  occxc_status st = guarded([&] {
    OCCXC_REQUIRE(srv, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    OCCXC_REQUIRE(it != srv->tickets.end(), OCCXC_ERR_INVALID_ARGUMENT, "server: unknown ticket");
    srv->cv_done.wait(lk, [&] { return srv->tickets[ticket].done; });
    Ticket t = srv->tickets[ticket];
    srv->tickets.erase(ticket);
    if (latency_ms) *latency_ms = t.latency_ms;
    result = t.status;
    msg = t.message;
  });
  if (st != OCCXC_OK) return st;
  if (result != OCCXC_OK) set_last_error(msg);
  return result;
}

occxc_status occxc_server_poll(occxc_server *srv, long long ticket, int *done) {
  return guarded([&] {
    OCCXC_REQUIRE(srv && done, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    OCCXC_REQUIRE(it != srv->tickets.end(), OCCXC_ERR_INVALID_ARGUMENT, "server: unknown ticket");
    *done = it->second.done ? 1 : 0;
  });
}

occxc_status occxc_server_cancel(occxc_server *srv, long long ticket, int *cancelled) {
  return guarded([&] {
    OCCXC_REQUIRE(srv && cancelled, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    *cancelled = 0;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    OCCXC_REQUIRE(it != srv->tickets.end(), OCCXC_ERR_INVALID_ARGUMENT, "server: unknown ticket");
    if (it->second.done) return;
    for (auto &kv : srv->scfs) {
      auto &q = kv.second->pending;
      for (auto r = q.begin(); r != q.end(); ++r) {
        if (r->ticket != ticket) continue;
        finish(srv, *r, OCCXC_ERR_CANCELLED, "build cancelled before it started");
        q.erase(r);
        *cancelled = 1;
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
        return;
      }
    }
    // not queued: the build is running and completes normally
  });
}

occxc_status occxc_server_close(occxc_server *srv, long long scf) {
  return guarded([&] {
    OCCXC_REQUIRE(srv, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->scfs.find(scf);
    OCCXC_REQUIRE(it != srv->scfs.end(), OCCXC_ERR_INVALID_ARGUMENT, "server: unknown SCF");
    cancel_pending(srv, it->second.get());
    it->second->closing = true;
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  });
}

occxc_status occxc_server_get_stats(occxc_server *srv, occxc_server_stats *stats) {
  return guarded([&] {
    OCCXC_REQUIRE(srv && stats, OCCXC_ERR_INVALID_ARGUMENT, "null pointer");
    std::lock_guard<std::mutex> lk(srv->mu);
    *stats = srv->stats;
    stats->resident_bytes = (double)srv->resident_bytes;
    stats->budget_bytes = (double)srv->budget;
    long long res = 0, wait = 0;
    for (auto &kv : srv->scfs) {
      if (kv.second->state == ScfState::Resident) ++res;
      else ++wait;
    }
    stats->scfs_resident = res;
    stats->scfs_waiting = wait;
  });
}

} // extern "C"
