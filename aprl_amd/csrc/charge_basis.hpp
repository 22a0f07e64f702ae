// Longitudinal node weights of the source shapes, shared by the charge
// diagnostics and the charge-matched inlet.
#pragma once

#include <hip/hip_runtime.h>

namespace aprl_gpu {

// longitudinal_shape::basis (longitudinal_shape.h:25-55) for the binomial3
// and compensated quintic nodes.
__device__ inline int chargeBasis(int shape, double z, double v[12]) {
  const double cell = floor(z), f = z - cell;
  for (int k = 0; k < 12; ++k) v[k] = 0.0;
  if (shape == 2) {
    const double c[3] = {.25, .5, .25};
    for (int k = -1; k <= 1; ++k) {
      v[k + 1] += c[k + 1] * (1. - f);
      v[k + 2] += c[k + 1] * f;
    }
    return static_cast<int>(cell) - 1;
  }
  double n[6] = {1., 0., 0., 0., 0., 0.};
  for (unsigned j = 1; j <= 5; ++j) {
    double saved = 0.;
    for (unsigned k = 0; k < j; ++k) {
      const double t = n[k] / j;
      n[k] = saved + (k + 1 - f) * t;
      saved = (f + j - k - 1) * t;
    }
    n[j] = saved;
  }
  const double p = 5., a = p / 6., b = 3 * p / 40. + p * (p - 1) / 72.,
               d = 5 * p / 112. + p * (p - 1) / 80. + p * (p - 1) * (p - 2) / 1296.;
  const double c[7] = {-d / 64., b / 16. + 3 * d / 32., -a / 4. - b / 4. - 15 * d / 64.,
                       1 + a / 2. + 3 * b / 8. + 5 * d / 16., -a / 4. - b / 4. - 15 * d / 64., b / 16. + 3 * d / 32.,
                       -d / 64.};
  for (unsigned j = 0; j <= 5; ++j)
    for (unsigned k = 0; k < 7; ++k) v[j + k] += n[j] * c[k];
  return static_cast<int>(cell) - 5;
}

}  // namespace aprl_gpu
