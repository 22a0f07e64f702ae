// Opaque handle types of the C-ABI.
#pragma once

#include "particles.hpp"

struct aprl_fields {
  aprl_gpu::FieldState s;
};
struct aprl_particles {
  aprl_fields* fields;
  aprl_gpu::ParticleStore p;
};
struct aprl_coupling {
  aprl_gpu::CouplingState c;
};
