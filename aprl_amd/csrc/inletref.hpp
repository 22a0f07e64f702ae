// The moving window's inlet reference on the device (BlockedCowanReference,
// blocked_cowan_reference.h:68-129, and movingWindowFieldValue,
// simulation_moving_window.cpp:280-294): the native E(0), B(-dt/2) of the
// rigidly translating source on the global reference grid, kept in device
// memory, and the planes the window takes from it every step.
#pragma once

#include <cstddef>

#include "common.hpp"

struct aprl_inlet_reference {
  aprl_grid_geometry g{};
  double origin[3] = {0, 0, 0};
  std::size_t nx[6] = {}, ny[6] = {}, nz[6] = {};
  aprl_gpu::DeviceArray<double> field[6];  // native Ex Ey Ez Bx By Bz, x fastest
  hipStream_t stream = nullptr;
  aprl_gpu::Stream ownStream;
};

namespace aprl_gpu {

// reference(a, i, j, k, step) over component a's x-y nodes into out (the
// window's precision), on stream st after everything already on it:
// z = k - translationBeta * step, linear between the two nearest planes;
// std::out_of_range outside the prepared planes, as the reference.
template <class R>
void referencePlane(const aprl_inlet_reference& ref, unsigned a, std::ptrdiff_t k, std::size_t step,
                    double translationBeta, R* out, hipStream_t st);

}  // namespace aprl_gpu
