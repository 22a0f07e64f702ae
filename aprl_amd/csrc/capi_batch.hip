// C-ABI entry point that steps many independent window instances together
// (declared in include/aprl_gpu.h).
#include <cmath>

#include "handles.hpp"

using namespace aprl_gpu;

namespace {
// Rethrow a nested C-ABI failure with the exception type its status maps to.
void rethrow(aprl_status s) {
  if (s == APRL_OK) return;
  const std::string m = aprl_last_error();
  switch (s) {
    case APRL_INVALID_ARGUMENT: throw std::invalid_argument(m);
    case APRL_OUT_OF_RANGE: throw std::out_of_range(m);
    case APRL_LOGIC_ERROR: throw std::logic_error(m);
    case APRL_DEVICE_ERROR: throw DeviceError(m);
    default: throw std::runtime_error(m);
  }
}
}  // namespace

extern "C" {

aprl_status aprl_window_step_batch(aprl_instance_step* steps, size_t count) {
  return guarded([&] {
    if (count && !steps) throw std::invalid_argument("Null instance step array");
    for (size_t i = 0; i < count; ++i) {
      const aprl_instance_step& s = steps[i];
      if (!s.fields || !s.coupling || !s.particles) throw std::invalid_argument("Null instance handle");
      if (s.coupling->fields != s.fields || s.particles->fields != s.fields)
        throw std::invalid_argument("Instance coupling, particles and field grid do not belong together");
      for (size_t j = 0; j < i; ++j)
        if (steps[j].fields == s.fields) throw std::invalid_argument("An instance appears twice in one batch");
    }
    // Phase 1: field half step and, for the FP32-split form, the sort and
    // the preview, for every instance before any instance waits. (A sort
    // prepared on a second stream per instance was measured slower here:
    // the instances' streams already keep the queues busy.)
    for (size_t i = 0; i < count; ++i) {
      aprl_instance_step& s = steps[i];
      CouplingState& C = s.coupling->c;
      rethrow(aprl_fields_clear_current(s.fields));
      rethrow(aprl_window_begin_step(s.fields, s.step, s.offset, nullptr));
      rethrow(aprl_fields_advance_magnetic(s.fields, nullptr));
      rethrow(aprl_window_set_inlet(s.fields, nullptr, nullptr));
      if (s.fields->fp32) splitStepBegin(C, s.particles->sp, s.origin, s.time, s.substeps);
    }
    // Phase 2: the trials (the FP32-split form's fixed count with its
    // commit and window shift; the plain form's first batch, whose
    // convergence the device decides), with no wait on an instance's trials.
    for (size_t i = 0; i < count; ++i) {
      aprl_instance_step& s = steps[i];
      if (s.fields->fp32) {
        splitStepMiddle(s.coupling->c, s.particles->sp);
        rethrow(aprl_window_finish_energy_step(s.fields));
        rethrow(aprl_window_shift(s.fields, s.step + 1, s.offset - 1, nullptr));
      } else {
        plainStepBegin(s.coupling->c, s.particles->p, s.origin, s.time, s.substeps);
      }
    }
    // Phase 3: wait for every instance and check its errors; the plain
    // form's commit and window shift follow its convergence.
    for (size_t i = 0; i < count; ++i) {
      aprl_instance_step& s = steps[i];
      if (s.fields->fp32) {
        splitStepEnd(s.coupling->c, s.particles->sp, &s.report);
      } else {
        plainStepFinish(s.coupling->c, s.particles->p, &s.report);
        rethrow(aprl_window_finish_energy_step(s.fields));
        rethrow(aprl_window_shift(s.fields, s.step + 1, s.offset - 1, nullptr));
      }
    }
  });
}

aprl_status aprl_window_run_batch(aprl_instance_run* runs, size_t count, size_t steps) {
  return guarded([&] {
    if (count && !runs) throw std::invalid_argument("Null instance run array");
    for (size_t i = 0; i < count; ++i) {
      const aprl_instance_step& s = runs[i].state;
      if (!s.fields || !s.coupling || !s.particles) throw std::invalid_argument("Null instance handle");
      if (s.coupling->fields != s.fields || s.particles->fields != s.fields)
        throw std::invalid_argument("Instance coupling, particles and field grid do not belong together");
      for (size_t j = 0; j < i; ++j)
        if (runs[j].state.fields == s.fields) throw std::invalid_argument("An instance appears twice in one batch");
    }
    // One instance's step, split where the host would wait for the device.
    auto begin = [&](aprl_instance_run& r) {
      aprl_instance_step& s = r.state;
      if (r.queue) {
        size_t admitted = 0;
        rethrow(aprl_pending_admit(r.queue, s.step, r.look_ahead, s.particles, &admitted));
      }
      r.entry_steps += s.fields->fp32 ? s.particles->sp.count : s.particles->p.count;
      rethrow(aprl_fields_clear_current(s.fields));
      rethrow(aprl_window_begin_step(s.fields, s.step, s.offset, nullptr));
      rethrow(aprl_fields_advance_magnetic(s.fields, nullptr));
      rethrow(aprl_window_set_inlet(s.fields, nullptr, nullptr));
      if (s.fields->fp32) splitStepBegin(s.coupling->c, s.particles->sp, s.origin, s.time, s.substeps);
    };
    auto trials = [&](aprl_instance_run& r) {
      aprl_instance_step& s = r.state;
      if (s.fields->fp32) {
        splitStepMiddle(s.coupling->c, s.particles->sp);
        rethrow(aprl_window_finish_energy_step(s.fields));
        rethrow(aprl_window_shift(s.fields, s.step + 1, s.offset - 1, nullptr));
      } else {
        plainStepBegin(s.coupling->c, s.particles->p, s.origin, s.time, s.substeps);
      }
    };
    auto finish = [&](aprl_instance_run& r) {
      aprl_instance_step& s = r.state;
      if (s.fields->fp32) {
        splitStepEnd(s.coupling->c, s.particles->sp, &s.report);
      } else {
        plainStepFinish(s.coupling->c, s.particles->p, &s.report);
        rethrow(aprl_window_finish_energy_step(s.fields));
        rethrow(aprl_window_shift(s.fields, s.step + 1, s.offset - 1, nullptr));
      }
      r.iterations += s.report.iterations;
      // Simulation::finishMovingWindowStep (simulation_moving_window.cpp:
      // 351): the box origin follows the shifted window
      s.step += 1;
      s.offset -= 1;
      const aprl_grid_geometry& g = s.fields->fp32 ? s.fields->s32.g : s.fields->s.g;
      s.time += g.dt;
      if (r.queue) rethrow(aprl_pending_origin(r.queue, s.step, &s.origin[2]));
      else s.origin[2] = std::fma(static_cast<double>(s.offset), g.dz, r.reference_z);
      if (r.plane) {
        int taken = 0;
        rethrow(aprl_lab_plane_sample(r.plane, s.time, s.origin, &taken));
        r.samples += static_cast<uint64_t>(taken);
      }
    };
    if (!steps) return;
    for (size_t i = 0; i < count; ++i) begin(runs[i]);
    for (size_t n = 0; n < steps; ++n) {
      for (size_t i = 0; i < count; ++i) trials(runs[i]);
      // An instance's next step is enqueued as soon as its own step is
      // done, so the device is not left to drain between steps.
      for (size_t i = 0; i < count; ++i) {
        finish(runs[i]);
        if (n + 1 < steps) begin(runs[i]);
      }
    }
  });
}

}  // extern "C"
