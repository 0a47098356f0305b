// server_policy.hpp -- the scheduling decisions of the serving loop
// (server.cpp), free of any device work, so that the same code drives the
// server and its CPU tests.  Each function is a template over the SCF table:
// a map from SCF id to (a pointer to) a record with the members
//   state (ScfState), busy, closing, bytes, last_used, pending
// where pending is a queue of requests with an `arrival` member.
#pragma once

#include <cstddef>

namespace occxc {
namespace serving {

enum class ScfState { Waiting, Admitting, Resident };

// The first SCF in registration order that waits for admission with work
// queued.  Admission is in that order: a large SCF at the head is not
// overtaken forever.
template <class Table> auto admission_head(Table &scfs) -> decltype(scfs.begin()->second.get()) {
  for (auto &kv : scfs) {
    auto *f = kv.second.get();
    if (f->state == ScfState::Waiting && !f->busy && !f->pending.empty()) return f;
  }
  return nullptr;
}

// The admission head if no admission is running and its footprint (when
// known from an earlier try) fits beside the resident SCFs.
template <class Table>
auto next_admission(Table &scfs, bool admitting, size_t resident, size_t budget)
    -> decltype(scfs.begin()->second.get()) {
  if (admitting) return nullptr;
  auto *f = admission_head(scfs);
  if (f && f->bytes && resident + f->bytes > budget) return nullptr;
  return f;
}

// The idle resident SCF (no build running or queued) used least recently,
// when evicting idle SCFs can make room for the admission head.
template <class Table>
auto next_eviction(Table &scfs, bool admitting, size_t resident, size_t budget)
    -> decltype(scfs.begin()->second.get()) {
  if (admitting) return nullptr;
  auto *h = admission_head(scfs);
  if (!h || !h->bytes || resident + h->bytes <= budget) return nullptr;
  decltype(h) victim = nullptr;
  size_t idle = 0;
  for (auto &kv : scfs) {
    auto *f = kv.second.get();
    if (f->state != ScfState::Resident || f->busy || !f->pending.empty()) continue;
    idle += f->bytes;
    if (!victim || f->last_used < victim->last_used) victim = f;
  }
  if (!victim || resident - idle + h->bytes > budget) return nullptr;
  return victim;
}

// The resident, idle SCF whose oldest queued request arrived first
// (continuous batching: a free slot takes the oldest runnable request).
template <class Table> auto next_build(Table &scfs) -> decltype(scfs.begin()->second.get()) {
  decltype(scfs.begin()->second.get()) best = nullptr;
  for (auto &kv : scfs) {
    auto *f = kv.second.get();
    if (f->state != ScfState::Resident || f->busy || f->pending.empty()) continue;
    if (!best || f->pending.front().arrival < best->pending.front().arrival) best = f;
  }
  return best;
}

// A closed SCF with nothing running or queued: ready to be freed.
template <class Table> auto next_release(Table &scfs) -> decltype(scfs.begin()->second.get()) {
  for (auto &kv : scfs) {
    auto *f = kv.second.get();
    if (f->closing && !f->busy && f->pending.empty()) return f;
  }
  return nullptr;
}

} // namespace serving
} // namespace occxc
