// The serving loop's scheduling decisions (occ_amd/csrc/server_policy.hpp)
// on a mock device: SCFs with a footprint, a memory budget, requests with
// arrival times.  The loop below takes decisions in the server worker's order
// (release, build, admission, eviction) and does on the mock device what the
// server does on the GPU.  Exit status = number of failed checks.
#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "server_policy.hpp"

using occxc::serving::ScfState;
namespace sv = occxc::serving;

struct Req {
  long long ticket;
  int arrival;
};

struct FScf {
  long long id = 0;
  ScfState state = ScfState::Waiting;
  bool busy = false, closing = false, failed = false;
  size_t bytes = 0;     // known footprint (0 before the first admission try)
  size_t need = 0;      // what the mock device's context set-up needs
  int last_used = 0;
  std::deque<Req> pending;
};
using Table = std::map<long long, std::unique_ptr<FScf>>;

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);            \
      ++fails;                                                    \
    }                                                             \
  } while (0)

struct Mock {
  Table scfs;
  size_t budget, resident = 0, peak = 0;
  int tick = 0, evictions = 0, deferred = 0, admitted = 0;
  bool admitting = false;
  std::vector<long long> done, failed;
  std::vector<long long> evicted;
  explicit Mock(size_t b) : budget(b) {}
  FScf *add(long long id, size_t need) {
    auto f = std::make_unique<FScf>();
    f->id = id;
    f->need = need;
    FScf *p = f.get();
    scfs[id] = std::move(f);
    return p;
  }
  void submit(long long id, long long ticket) { scfs[id]->pending.push_back({ticket, ++tick}); }
  // one decision, as the server's worker takes it; false when idle
  bool step() {
    FScf *f;
    if ((f = sv::next_release(scfs))) {
      if (f->state == ScfState::Resident) resident -= f->bytes;
      scfs.erase(f->id);
      return true;
    }
    if ((f = sv::next_build(scfs))) {
      done.push_back(f->pending.front().ticket);
      f->pending.pop_front();
      f->last_used = ++tick;
      return true;
    }
    if ((f = sv::next_admission(scfs, admitting, resident, budget))) {
      // context set-up refuses beyond budget - resident (server.cpp admit)
      const size_t limit = budget - resident;
      if (f->need <= limit) {
        f->state = ScfState::Resident;
        f->bytes = f->need;
        resident += f->need;
        peak = std::max(peak, resident);
        ++admitted;
      } else if (f->need <= budget && resident > 0) {
        f->bytes = f->need;  // wait for memory
        ++deferred;
      } else {
        f->failed = true;
        while (!f->pending.empty()) {
          failed.push_back(f->pending.front().ticket);
          f->pending.pop_front();
        }
      }
      return true;
    }
    if ((f = sv::next_eviction(scfs, admitting, resident, budget))) {
      resident -= f->bytes;
      f->state = ScfState::Waiting;
      evicted.push_back(f->id);
      ++evictions;
      return true;
    }
    return false;
  }
  void run() {
    for (int k = 0; k < 10000 && step();) ++k;
  }
};

int main() {
  // 1. budget for one large SCF at a time: every request completes, the
  //    resident bytes never exceed the budget, idle SCFs are evicted
  {
    Mock m(100);
    m.add(1, 60);
    m.add(2, 60);
    m.add(3, 30);
    m.submit(1, 11);
    m.submit(2, 21);
    m.submit(3, 31);
    m.submit(1, 12);
    m.run();
    CHECK(m.done.size() == 4 && m.failed.empty());
    CHECK(m.peak <= 100);
    CHECK(m.evictions >= 1 && m.deferred >= 1);
    CHECK(m.done[0] == 11);  // the first SCF is admitted and built first
  }
  // 2. the least recently used idle SCF is the one evicted
  {
    Mock m(100);
    FScf *x = m.add(1, 40), *y = m.add(2, 40);
    for (FScf *f : {x, y}) {
      f->state = ScfState::Resident;
      f->bytes = f->need;
      m.resident += f->need;
    }
    x->last_used = 5;
    y->last_used = 3;
    FScf *z = m.add(3, 50);
    z->bytes = 50;  // known from an earlier try
    m.submit(3, 31);
    CHECK(sv::next_admission(m.scfs, false, m.resident, m.budget) == nullptr);
    CHECK(sv::next_eviction(m.scfs, false, m.resident, m.budget) == y);
    // a busy or a queued SCF is never evicted
    y->busy = true;
    CHECK(sv::next_eviction(m.scfs, false, m.resident, m.budget) == x);
    y->busy = false;
    m.submit(2, 21);
    CHECK(sv::next_eviction(m.scfs, false, m.resident, m.budget) == x);
    // evicting every idle SCF would not make room: no eviction
    m.submit(1, 11);
    CHECK(sv::next_eviction(m.scfs, false, m.resident, m.budget) == nullptr);
  }
  // 3. a free slot takes the oldest runnable request
  {
    Mock m(1000);
    FScf *a = m.add(1, 10), *b = m.add(2, 10);
    for (FScf *f : {a, b}) {
      f->state = ScfState::Resident;
      f->bytes = 10;
    }
    m.submit(2, 21);
    m.submit(1, 11);
    CHECK(sv::next_build(m.scfs) == b);
    b->busy = true;  // its build runs: the other SCF's request goes next
    CHECK(sv::next_build(m.scfs) == a);
  }
  // 4. one admission at a time, in registration order
  {
    Mock m(1000);
    m.add(1, 10);
    m.add(2, 10);
    m.submit(2, 21);
    m.submit(1, 11);
    CHECK(sv::next_admission(m.scfs, true, 0, m.budget) == nullptr);
    CHECK(sv::next_admission(m.scfs, false, 0, m.budget) == m.scfs[1].get());
  }
  // 5. an SCF larger than the whole budget fails its requests; the others run
  {
    Mock m(100);
    m.add(1, 500);
    m.add(2, 20);
    m.submit(1, 11);
    m.submit(2, 21);
    m.submit(1, 12);
    m.run();
    CHECK(m.failed.size() == 2 && m.done.size() == 1 && m.done[0] == 21);
  }
  // 6. a closed SCF is released once drained, and its memory returns
  {
    Mock m(100);
    FScf *a = m.add(1, 70);
    m.submit(1, 11);
    m.run();
    a->closing = true;
    m.add(2, 70);
    m.submit(2, 21);
    m.run();
    CHECK(m.scfs.count(1) == 0 && m.done.size() == 2 && m.peak <= 100 && m.evictions == 0);
  }
  std::printf("%d failed checks\n", fails);
  return fails;
}
