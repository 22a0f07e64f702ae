// C-ABI of the moving window's pending admission queue (declared in
// include/aprl_gpu.h).
#include <memory>

#include "handles.hpp"
#include "pending.hpp"

struct aprl_pending {
  aprl_gpu::PendingQueue q;
  explicit aprl_pending(const aprl_gpu::PendingGeometry& g, std::uint64_t step) : q(g, step) {}
};

using namespace aprl_gpu;

extern "C" {

aprl_status aprl_repeat_held_addition(double initial, double increment, uint64_t count, double* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("Null output");
    *out = repeatHeldAddition(initial, increment, count);
  });
}

aprl_status aprl_pending_create(const aprl_pending_geometry* geometry, uint64_t initial_step, aprl_pending** out) {
  return guarded([&] {
    if (!geometry || !out) throw std::invalid_argument("Null pending geometry or output");
    PendingGeometry g;
    g.referenceOriginZ = geometry->reference_origin_z;
    g.initialOffset = geometry->initial_offset;
    g.dz = geometry->dz;
    g.heldDistancePerBeta = geometry->held_distance_per_beta;
    g.maximumStep = geometry->maximum_step;
    *out = new aprl_pending(g, initial_step);
  });
}

void aprl_pending_destroy(aprl_pending* queue) { delete queue; }

aprl_status aprl_pending_extract(aprl_pending* queue, size_t n, double* x3, double* u3, double* q, double* m,
                                 double* w, uint64_t* id, int* frozen, double* beta, size_t* retained) {
  return guarded([&] {
    if (!queue || !retained) throw std::invalid_argument("Null pending queue or output");
    if (n && (!x3 || !u3 || !q || !m || !w || !id || !frozen || !beta))
      throw std::invalid_argument("Null particle array");
    *retained = queue->q.extract(n, x3, u3, q, m, w, id, frozen, beta);
  });
}

aprl_status aprl_pending_admit(aprl_pending* queue, uint64_t step, uint64_t look_ahead, aprl_particles* particles,
                               size_t* admitted) {
  return guarded([&] {
    if (!queue || !particles || !admitted) throw std::invalid_argument("Null pending queue, store or output");
    *admitted = queue->q.take(step, look_ahead, [&](const std::vector<PendingParticle>& due) {
      if (due.empty()) return;
      // the store may grow (a reallocation copies what the last step's
      // kernels may still write): wait for them, only when entries come in
      hipCheck(hipStreamSynchronize(particles->stream()), "hipStreamSynchronize");
      const std::size_t k = due.size();
      std::vector<double> x(3 * k), u(3 * k), q(k), m(k), w(k), beta(k);
      std::vector<uint64_t> id(k);
      std::vector<int> state(k, kStateHeld);
      for (std::size_t i = 0; i < k; ++i) {
        for (int a = 0; a < 3; ++a) {
          x[3 * i + a] = due[i].x[a];
          u[3 * i + a] = due[i].u[a];
        }
        q[i] = due[i].q;
        m[i] = due[i].m;
        w[i] = due[i].w;
        beta[i] = due[i].beta;
        id[i] = due[i].id;
        if (!(std::isfinite(beta[i]) && beta[i] >= 0.0 && beta[i] < 1.0))
          throw std::invalid_argument("Held beta must be finite in [0, 1)");
      }
      if (particles->fields->fp32)
        appendSplitParticles(particles->sp, particles->fields->s32, k, x.data(), u.data(), q.data(), m.data(),
                             w.data(), id.data(), state.data(), beta.data());
      else
        appendParticles(particles->p, k, x.data(), u.data(), q.data(), m.data(), w.data(), id.data(), state.data(),
                        beta.data());
    });
  });
}

aprl_status aprl_pending_view(const aprl_pending* queue, uint64_t step, uint64_t look_ahead, size_t capacity,
                              double* x3, uint64_t* id, size_t* n) {
  return guarded([&] {
    if (!queue || !n) throw std::invalid_argument("Null pending queue or output");
    const std::vector<PendingParticle> near = queue->q.view(step, look_ahead);
    *n = near.size();
    const std::size_t k = std::min(capacity, near.size());
    if (k && (!x3 || !id)) throw std::invalid_argument("Null view array");
    for (std::size_t i = 0; i < k; ++i) {
      for (int a = 0; a < 3; ++a) x3[3 * i + a] = near[i].x[a];
      id[i] = near[i].id;
    }
  });
}

aprl_status aprl_pending_state(const aprl_pending* queue, uint64_t* waiting, uint64_t* next_due, uint64_t* admitted) {
  return guarded([&] {
    if (!queue) throw std::invalid_argument("Null pending queue");
    if (waiting) *waiting = queue->q.size();
    if (next_due) *next_due = queue->q.nextDueStep();
    if (admitted) *admitted = queue->q.admittedCount();
  });
}

aprl_status aprl_pending_origin(const aprl_pending* queue, uint64_t step, double* z) {
  return guarded([&] {
    if (!queue || !z) throw std::invalid_argument("Null pending queue or output");
    *z = queue->q.origin(step);
  });
}

}  // extern "C"
