// Test infrastructure only: the oracle's restatement of the moving window's
// pending admission queue (pending_particle_queue.h:25-76,
// pending_particle_queue.cpp:77-244). It is the checker for the library's
// queue and is never linked into it.
//
// Pinning: repeatHeldAddition is restated by its definition, count literal
// binary64 additions (pending_particle_queue.h:11-13), so it is exact by
// construction and checks the library's jumping version and the reference
// build's. The queue's due steps come from a step-by-step scan of the same
// literal additions (no estimate, no search), checked against the reference
// build's queue over whole runs.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <vector>

#include "orc_particles.hpp"

namespace orc {

// pending_particle_queue.h:11-14 by definition: one rounded addition per count.
inline double repeatHeldAddition(double initial, double increment, std::uint64_t count) {
  if (!std::isfinite(initial) || !std::isfinite(increment))
    throw std::invalid_argument("Held position/increment must be finite");
  double x = initial;
  for (std::uint64_t n = 0; n < count; ++n) {
    x = x + increment;
    if (!std::isfinite(x)) throw std::overflow_error("Held position overflow");
  }
  return x;
}

struct PendingGeometry {
  double referenceOriginZ = 0.;
  std::int64_t initialOffset = 0;
  double dz = 0.;
  double heldDistancePerBeta = 0.;  // c * dt
  std::uint64_t maximumStep = 0;
};

// PendingParticleQueue (pending_particle_queue.cpp:99-244).
class PendingQueue {
 public:
  PendingQueue(const PendingGeometry& g, std::uint64_t initialStep) : g_(g), s0_(initialStep), last_(initialStep) {
    if (!(g.dz > 0.) || !(g.heldDistancePerBeta > 0.) || !std::isfinite(g.dz) || !std::isfinite(g.heldDistancePerBeta) ||
        !std::isfinite(g.referenceOriginZ) || initialStep > g.maximumStep)
      throw std::invalid_argument("Invalid pending-window geometry/horizon");
  }

  // origin (.cpp:110-116)
  double origin(std::uint64_t step) const {
    if (step > g_.maximumStep) throw std::out_of_range("Pending origin exceeds horizon");
    return std::fma(static_cast<double>(g_.initialOffset - static_cast<std::int64_t>(step)), g_.dz,
                    g_.referenceOriginZ);
  }

  // extract (.cpp:148-175): frozen, upstream of the initial inlet, moving
  // downstream and not taking part in the initial step.
  void extract(std::vector<Particle>& active) {
    if (done_) throw std::logic_error("Pending queue can only be initialized once");
    const double inlet = origin(s0_);
    std::vector<Particle> kept;
    std::vector<std::uint64_t> due;
    for (const Particle& p : active) {
      const double inc = g_.heldDistancePerBeta * p.betaHeld;
      if (p.frozen && p.position[2] < inlet && p.betaHeld >= 0. &&
          !(repeatHeldAddition(p.position[2], inc, 1) >= inlet)) {
        due.push_back(firstDue(p));
        pending_.push_back(p);
      } else {
        kept.push_back(p);
      }
    }
    active.swap(kept);
    // admission order: by due step, extraction order within a step
    for (std::size_t i = 0; i < due.size(); ++i) order_.push_back(i);
    std::stable_sort(order_.begin(), order_.end(), [&](std::size_t a, std::size_t b) { return due[a] < due[b]; });
    for (std::size_t i : order_) due_.push_back(due[i]);
    done_ = true;
  }

  // admit (.cpp:184-206): appends the due entries at their start-of-step z.
  std::size_t admit(std::uint64_t step, std::vector<Particle>& active) {
    if (!done_) throw std::logic_error("Pending queue has not been initialized");
    if (step < last_ || step > g_.maximumStep || step > nextDueStep())
      throw std::logic_error("Pending admission skipped an event or exceeds its horizon");
    last_ = step;
    std::size_t n = 0;
    while (cursor_ < order_.size() && due_[cursor_] <= step) {
      active.push_back(at(cursor_, step));
      ++cursor_;
      ++n;
    }
    return n;
  }

  // visitNear (.cpp:221-226) as a list
  std::vector<Particle> near(std::uint64_t step, std::uint64_t lookAhead) const {
    if (!done_ || step < s0_ || step > g_.maximumStep) throw std::logic_error("Invalid pending support view");
    const std::uint64_t until = std::min(g_.maximumStep, step + std::min(lookAhead, g_.maximumStep - step));
    std::vector<Particle> out;
    for (std::size_t c = cursor_; c < order_.size() && due_[c] <= until; ++c) out.push_back(at(c, step));
    return out;
  }

  std::size_t size() const { return order_.size() - cursor_; }
  std::uint64_t nextDueStep() const {
    return cursor_ < due_.size() ? std::min(due_[cursor_], g_.maximumStep + 1) : g_.maximumStep + 1;
  }
  std::size_t admitted() const { return cursor_; }

 private:
  // first step whose held end reaches its inlet, scanning step by step
  std::uint64_t firstDue(const Particle& p) const {
    const double inc = g_.heldDistancePerBeta * p.betaHeld;
    double z = p.position[2] + inc;  // end of the initial step
    for (std::uint64_t s = s0_ + 1; s <= g_.maximumStep; ++s) {
      z = z + inc;
      if (!std::isfinite(z)) throw std::overflow_error("Held position overflow");
      if (z >= origin(s)) return s;
    }
    return g_.maximumStep + 1;
  }
  Particle at(std::size_t c, std::uint64_t step) const {
    Particle p = pending_[order_[c]];
    p.position[2] = repeatHeldAddition(p.position[2], g_.heldDistancePerBeta * p.betaHeld, step - s0_);
    return p;
  }

  PendingGeometry g_;
  std::uint64_t s0_, last_;
  bool done_ = false;
  std::vector<Particle> pending_;
  std::vector<std::size_t> order_;
  std::vector<std::uint64_t> due_;
  std::size_t cursor_ = 0;
};

}  // namespace orc
