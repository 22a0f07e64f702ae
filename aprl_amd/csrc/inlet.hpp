// Charge-matched moving-window inlet (charge_matched_inlet.h:16-72) and
// the centered normal reference of the inlet Ez halo
// (moving_window_fields.cpp:225-234).
#pragma once

#include "particles.hpp"

namespace aprl_gpu {

// After a shift to `origin`: the new layer's Ez on the interior aperture
// from Ez(k=1), div_xy E(k=1) and the shaped charge of the store's active
// particles near the inlet (shape 0 quintic, 2 linear-binomial3; xBinomial
// selects the x-binomial3 transverse weights).
void completeInletPlain(FieldState& F, const ParticleStore& P, const double origin[3], int shape, bool xBinomial);
void completeInletSplit(FieldState32& F, const SplitStore& P, const double origin[3], int shape, bool xBinomial);

// halo -= Ez(k=0) - ref on the interior aperture nodes of the Ez plane.
template <class R>
void centerInletHalo(FieldStateT<R>& F, R* halo, const R* ref);

}  // namespace aprl_gpu
