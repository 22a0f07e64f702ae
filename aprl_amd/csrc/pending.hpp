// Pending admission queue of the moving window: held (frozen) particles
// upstream of the inlet wait on the host until they are due and enter the
// device store then (PendingParticleQueue, pending_particle_queue.h:25-76),
// positioned by the exact repeated held addition
// (repeatHeldAddition, pending_particle_queue.h:11-14).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

namespace aprl_gpu {

// The value of `count` separately rounded binary64 additions
// x <- x + increment from x = initial (round to nearest, ties to even),
// without performing every addition: inside one binade the additions that
// round the same way are replaced by one integer step of the binade's
// units. Throws invalid_argument for nonfinite input or another rounding
// mode and overflow_error when a sum overflows.
double repeatHeldAddition(double initial, double increment, std::uint64_t count);

// Window geometry of the queue (PendingWindowGeometry,
// pending_particle_queue.h:16-23): origin(step) =
// fma(initialOffset - step, dz, referenceOriginZ); a held particle moves by
// heldDistancePerBeta * beta per step (pass c * dt).
struct PendingGeometry {
  double referenceOriginZ = 0.0;
  std::int64_t initialOffset = 0;
  double dz = 0.0;
  double heldDistancePerBeta = 0.0;
  std::uint64_t maximumStep = 0;
};

// One waiting particle: the state it had at the queue's initial step.
struct PendingParticle {
  double x[3], u[3], q, m, w, beta;
  std::uint64_t id;
};

class PendingQueue {
 public:
  PendingQueue(const PendingGeometry& g, std::uint64_t initialStep);

  // Moves every held particle that is upstream of the inlet, moves
  // downstream (beta >= 0) and does not yet take part in the initial step
  // out of the arrays, which are compacted in place with the order of the
  // rest unchanged; returns the number kept (extract, .h:33-36). Once only.
  std::size_t extract(std::size_t n, double* x3, double* u3, double* q, double* m, double* w, std::uint64_t* id,
                      int* frozen, double* beta);

  // The waiting particles due at or before min(step + lookAhead,
  // maximumStep), in admission order (due step, then extraction order),
  // with z at the start of `step`. take() hands them to `sink` and removes
  // them from the queue once the sink returns (a sink that throws leaves
  // the queue as it was); it requires steps in order that skip no due step
  // (admit, .h:38-43). view() leaves the queue as it is (visitNear,
  // .h:48-49).
  std::size_t take(std::uint64_t step, std::uint64_t lookAhead,
                   const std::function<void(const std::vector<PendingParticle>&)>& sink);
  std::vector<PendingParticle> view(std::uint64_t step, std::uint64_t lookAhead) const;

  std::size_t size() const { return order_.size() - cursor_; }
  // due step of the next waiting particle; maximumStep + 1 when none is
  // due within the horizon
  std::uint64_t nextDueStep() const;
  std::size_t admittedCount() const { return cursor_; }
  double origin(std::uint64_t step) const;
  bool extracted() const { return extracted_; }

 private:
  double heldIncrement(const PendingParticle& p) const { return g_.heldDistancePerBeta * p.beta; }
  bool participates(double z0, double increment, std::uint64_t step) const;
  std::uint64_t firstDue(double z0, double increment) const;
  std::size_t endOfDue(std::uint64_t until) const;
  PendingParticle at(std::size_t n, std::uint64_t step) const;

  PendingGeometry g_;
  std::uint64_t initialStep_, lastStep_;
  bool extracted_ = false;
  std::vector<PendingParticle> pending_;  // extraction order
  std::vector<std::uint32_t> order_;      // admission order
  std::vector<std::uint64_t> due_;        // due step per admission slot
  std::size_t cursor_ = 0;
};

}  // namespace aprl_gpu
