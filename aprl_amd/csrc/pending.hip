// Host side of the moving window's pending admission queue
// (pending_particle_queue.cpp) and the exact repeated held addition.
#include "pending.hpp"

#include <algorithm>
#include <cfenv>
#include <cfloat>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace aprl_gpu {

namespace {

constexpr std::uint64_t kUnits52 = std::uint64_t(1) << 52;  // significand of 1.0 in units of its ulp
constexpr std::uint64_t kUnits53 = std::uint64_t(1) << 53;

// One literal binary64 addition (a lone add: nothing to contract).
double addOnce(double x, double increment) {
  const double y = x + increment;
  if (!std::isfinite(y)) throw std::overflow_error("Held position overflow");
  return y;
}

// A run of additions of increment >= 0 from x that all round on the grid
// of x's binade in the same way. Let u be the binade's unit in the last
// place, A = |x| / u in [2^52, 2^53) and r = increment / u = m + f. Each
// addition moves A by r rounded to an integer: m for f < 1/2, m + 1 for
// f > 1/2, and for f = 1/2 the even neighbour, which from an even A is
// m (m even) or m + 1 (m odd) again. The run stops while every exact sum
// still lies inside the binade (upwards for x > 0, downwards in magnitude
// for x < 0), so the rounding grid cannot change within it. Returns the
// number of additions replaced (0: take one literal addition instead).
std::uint64_t constantRun(double& x, double increment, std::uint64_t count) {
  if (!(std::fabs(x) >= DBL_MIN)) return 0;  // zero and subnormals step literally
  const int e = std::ilogb(x);
  const double r = std::ldexp(increment, 52 - e);
  if (!(r >= 0.5) || !(r < 0x1p51)) return 0;
  const double whole = std::floor(r);
  const double frac = r - whole;
  const std::uint64_t m = static_cast<std::uint64_t>(whole);
  const std::uint64_t a = static_cast<std::uint64_t>(std::ldexp(std::fabs(x), 52 - e));
  std::uint64_t stride;
  if (frac < 0.5) {
    stride = m;
  } else if (frac > 0.5) {
    stride = m + 1;
  } else {
    if (a & 1) return 0;  // one literal tie settles the even phase
    stride = (m & 1) ? m + 1 : m;
  }
  if (stride == 0) return 0;
  // the largest k with every exact sum of the run inside the binade:
  // x > 0: a + (k - 1) stride + m + 1 <= 2^53 - 1;
  // x < 0: a - (k - 1) stride - m - 1 >= 2^52
  std::uint64_t room;
  if (x > 0.0) {
    if (a + m + 2 > kUnits53) return 0;
    room = kUnits53 - 2 - m - a;
  } else {
    if (a < kUnits52 + m + 1) return 0;
    room = a - m - 1 - kUnits52;
  }
  const std::uint64_t k = std::min(count, room / stride + 1);
  const std::uint64_t moved = x > 0.0 ? a + k * stride : a - k * stride;
  x = std::copysign(std::ldexp(static_cast<double>(moved), e - 52), x);
  return k;
}

double repeatNonNegative(double x, double increment, std::uint64_t count) {
  while (count) {
    const double y = addOnce(x, increment);
    --count;
    if (y == x) return y;  // a fixed point stays fixed
    x = y;
    if (count) count -= constantRun(x, increment, count);
  }
  return x;
}

}  // namespace

double repeatHeldAddition(double initial, double increment, std::uint64_t count) {
  static_assert(std::numeric_limits<double>::is_iec559 && std::numeric_limits<double>::digits == 53,
                "IEEE binary64 required");
  if (!std::isfinite(initial) || !std::isfinite(increment))
    throw std::invalid_argument("Held position/increment must be finite");
  if (std::fegetround() != FE_TONEAREST) throw std::invalid_argument("Held jump requires round-to-nearest/even");
  if (!count) return initial;
  if (increment < 0.0) {
    // round to nearest is symmetric; an exact cancellation is +0 either way
    const double r = -repeatNonNegative(-initial, -increment, count);
    return r == 0.0 ? 0.0 : r;
  }
  return repeatNonNegative(initial, increment, count);
}

PendingQueue::PendingQueue(const PendingGeometry& g, std::uint64_t initialStep)
    : g_(g), initialStep_(initialStep), lastStep_(initialStep) {
  const std::uint64_t maxSigned = static_cast<std::uint64_t>(std::numeric_limits<std::int64_t>::max());
  if (!(g.dz > 0.0) || !std::isfinite(g.dz) || !(g.heldDistancePerBeta > 0.0) ||
      !std::isfinite(g.heldDistancePerBeta) || !std::isfinite(g.referenceOriginZ) || initialStep > g.maximumStep ||
      g.maximumStep >= maxSigned ||
      g.initialOffset < std::numeric_limits<std::int64_t>::min() + static_cast<std::int64_t>(g.maximumStep))
    throw std::invalid_argument("Invalid pending-window geometry/horizon");
  (void)repeatHeldAddition(0.0, 0.0, 0);  // rejects another rounding mode up front
  (void)origin(initialStep_);
  (void)origin(g.maximumStep);
}

double PendingQueue::origin(std::uint64_t step) const {
  if (step > g_.maximumStep) throw std::out_of_range("Pending origin exceeds horizon");
  const std::int64_t offset = g_.initialOffset - static_cast<std::int64_t>(step);
  const double value = std::fma(static_cast<double>(offset), g_.dz, g_.referenceOriginZ);
  if (!std::isfinite(value)) throw std::overflow_error("Pending window origin overflow");
  return value;
}

// The held path's end after the additions of steps initialStep..step
// reaches the inlet of that step.
bool PendingQueue::participates(double z0, double increment, std::uint64_t step) const {
  return repeatHeldAddition(z0, increment, step - initialStep_ + 1) >= origin(step);
}

// First step after the initial one in which the particle takes part
// (participation only turns on: the held end moves downstream and the
// inlet upstream), or maximumStep + 1. A straight-line estimate picks the
// first probe; the answer is the first step whose exact predicate holds.
std::uint64_t PendingQueue::firstDue(double z0, double increment) const {
  const std::uint64_t none = g_.maximumStep + 1;
  std::uint64_t lo = initialStep_ + 1, hi = none;  // predicate false below lo, true at hi (or hi = none)
  if (lo > g_.maximumStep) return none;
  const double gap = origin(initialStep_) - z0 - increment;
  const double estimate = std::ceil(gap / (g_.dz + increment));
  if (std::isfinite(estimate)) {
    const double span = static_cast<double>(g_.maximumStep - initialStep_);
    const std::uint64_t guess =
        initialStep_ + static_cast<std::uint64_t>(std::max(1.0, std::min(estimate, span)));
    if (participates(z0, increment, guess)) {
      hi = guess;
      if (guess - 1 >= lo && !participates(z0, increment, guess - 1)) return guess;
    } else {
      lo = guess + 1;
      if (lo <= g_.maximumStep && participates(z0, increment, lo)) return lo;
    }
  }
  while (lo < hi) {
    const std::uint64_t mid = lo + (hi - lo) / 2;
    if (participates(z0, increment, mid)) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

std::size_t PendingQueue::extract(std::size_t n, double* x3, double* u3, double* q, double* m, double* w,
                                  std::uint64_t* id, int* frozen, double* beta) {
  if (extracted_) throw std::logic_error("Pending queue can only be initialized once");
  if (n >= std::numeric_limits<std::uint32_t>::max()) throw std::length_error("Too many particles for the pending queue");
  const double inlet = origin(initialStep_);
  std::vector<std::uint64_t> due;
  std::size_t kept = 0;
  for (std::size_t i = 0; i < n; ++i) {
    const double z = x3[3 * i + 2];
    PendingParticle p{};
    p.beta = beta[i];
    const double inc = heldIncrement(p);
    if (frozen[i] && z < inlet && p.beta >= 0.0 && !participates(z, inc, initialStep_)) {
      for (int a = 0; a < 3; ++a) {
        p.x[a] = x3[3 * i + a];
        p.u[a] = u3[3 * i + a];
      }
      p.q = q[i];
      p.m = m[i];
      p.w = w[i];
      p.id = id[i];
      due.push_back(firstDue(z, inc));
      pending_.push_back(p);
      continue;
    }
    if (kept != i) {
      for (int a = 0; a < 3; ++a) {
        x3[3 * kept + a] = x3[3 * i + a];
        u3[3 * kept + a] = u3[3 * i + a];
      }
      q[kept] = q[i];
      m[kept] = m[i];
      w[kept] = w[i];
      id[kept] = id[i];
      frozen[kept] = frozen[i];
      beta[kept] = beta[i];
    }
    ++kept;
  }
  order_.resize(pending_.size());
  for (std::size_t i = 0; i < order_.size(); ++i) order_[i] = static_cast<std::uint32_t>(i);
  std::stable_sort(order_.begin(), order_.end(), [&](std::uint32_t a, std::uint32_t b) { return due[a] < due[b]; });
  due_.resize(order_.size());
  for (std::size_t i = 0; i < order_.size(); ++i) due_[i] = due[order_[i]];
  extracted_ = true;
  return kept;
}

std::size_t PendingQueue::endOfDue(std::uint64_t until) const {
  return static_cast<std::size_t>(std::upper_bound(due_.begin() + cursor_, due_.end(), until) - due_.begin());
}

PendingParticle PendingQueue::at(std::size_t n, std::uint64_t step) const {
  PendingParticle p = pending_[order_[n]];
  p.x[2] = repeatHeldAddition(p.x[2], heldIncrement(p), step - initialStep_);
  return p;
}

std::uint64_t PendingQueue::nextDueStep() const {
  return cursor_ < due_.size() ? std::min(due_[cursor_], g_.maximumStep + 1) : g_.maximumStep + 1;
}

std::size_t PendingQueue::take(std::uint64_t step, std::uint64_t lookAhead,
                               const std::function<void(const std::vector<PendingParticle>&)>& sink) {
  if (!extracted_) throw std::logic_error("Pending queue has not been initialized");
  if (step < lastStep_ || step > g_.maximumStep || step > nextDueStep())
    throw std::logic_error("Pending admission skipped an event or exceeds its horizon");
  const std::uint64_t until = step + std::min(lookAhead, g_.maximumStep - step);
  const std::size_t end = endOfDue(until);
  std::vector<PendingParticle> out;
  out.reserve(end - cursor_);
  for (std::size_t n = cursor_; n < end; ++n) out.push_back(at(n, step));
  sink(out);
  cursor_ = end;
  lastStep_ = step;
  return out.size();
}

std::vector<PendingParticle> PendingQueue::view(std::uint64_t step, std::uint64_t lookAhead) const {
  if (!extracted_ || step < initialStep_ || step > g_.maximumStep)
    throw std::logic_error("Invalid pending support view");
  const std::uint64_t until = step + std::min(lookAhead, g_.maximumStep - step);
  const std::size_t end = endOfDue(until);
  std::vector<PendingParticle> out;
  out.reserve(end - cursor_);
  for (std::size_t n = cursor_; n < end; ++n) out.push_back(at(n, step));
  return out;
}

}  // namespace aprl_gpu
