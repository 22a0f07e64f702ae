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
  aprl_gpu::ParticleStore p;  // plain FP64 form
  aprl_gpu::SplitStore sp;    // FP32-split form
  hipStream_t stream() const { return fields->fp32 ? fields->s32.stream : fields->s.stream; }
};
struct aprl_coupling {
  aprl_gpu::CouplingState c;
  aprl_fields* fields = nullptr;
};
