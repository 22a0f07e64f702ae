// Opaque handle types of the C-ABI.
#pragma once

#include "particles.hpp"

struct aprl_fields {
  bool fp32 = false;
  aprl_gpu::FieldState s;      // plain FP64 form
  aprl_gpu::FieldState32 s32;  // FP32 form
};
struct aprl_particles {
  aprl_fields* fields;
  aprl_gpu::ParticleStore p;
};
struct aprl_coupling {
  aprl_gpu::CouplingState c;
};
