// Prescribed reference surface current of the moving window
// (TransverseReferenceCurrent::apply, transverse_reference_current.cpp:
// 104-120): the equivalent surface current of the finite reference field at
// the four transverse interfaces of the aperture, added to Jz every Picard
// trial and at the commit, interpolated linearly between the table rows at
// z = offset + k - beta (step + 1).
#include <cmath>

#include "particles.hpp"

namespace aprl_gpu {

namespace {

template <class R>
__global__ void kReferenceCurrent(R* jz, R* esum, const R* estar, const R* eold, double negS, int writeEsum,
                                  const unsigned long long* tileCounts, int anyActive, int ntx, int nty, int ntz,
                                  const double* table, const int* pts, int count, int rows, int nx, int ny, int nz,
                                  double offset, double shiftZ, const unsigned* stop) {
  if (stop && *stop) return;
  const long long total = static_cast<long long>(count) * nz;
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int p = static_cast<int>(t % count), k = static_cast<int>(t / count);
    const double z = offset + static_cast<double>(k) - shiftZ;
    const int lo = static_cast<int>(floor(z)), hi = min(lo + 1, rows - 1);
    const double fraction = z - lo;
    const double a = table[static_cast<size_t>(lo) * count + p], b = table[static_cast<size_t>(hi) * count + p];
    const double value = fraction == 0. ? a : (1. - fraction) * a + fraction * b;
    const int i = pts[2 * p], j = pts[2 * p + 1];
    const size_t idx = (static_cast<size_t>(k) * (ny + 1) + j) * (nx + 1) + i;
    const int tx = min(i / kTileX, ntx - 1), ty = min(j / kTileY, nty - 1), tz = min(k / kTileZ, ntz - 1);
    const bool active = anyActive && activeField(tileCounts[(static_cast<size_t>(tz) * nty + ty) * ntx + tx]) != 0;
    const double current = (active ? static_cast<double>(jz[idx]) : 0.0) + value;
    jz[idx] = static_cast<R>(current);
    if (active && writeEsum) {
      // applyCurrentSource(-1) on the preview for Ez (interior i, j)
      double e = static_cast<double>(estar[idx]);
      if (i > 0 && i < nx && j > 0 && j < ny && current != 0.) e += negS * current;
      esum[idx] = static_cast<R>(e + static_cast<double>(eold[idx]));
    }
  }
}

}  // namespace

template <class R>
void referenceCurrentTrial(CouplingState& C, FieldStateT<R>& F, R* esumZ, const R* estarZ, bool writeEsum,
                           bool anyActive, int ntx, int nty, int ntz, const unsigned* stop) {
  if (!C.refCurrent) return;
  const aprl_grid_geometry& g = F.g;
  const double offset = static_cast<double>(F.offset);
  const double shiftZ = C.refBeta * static_cast<double>(F.step + 1);
  // the queried rows of k = 0 and k = nz - 1 bound all others
  const double lo = offset + 0.0 - shiftZ, hi = offset + static_cast<double>(g.nz - 1) - shiftZ;
  if (!std::isfinite(lo) || !std::isfinite(hi) || lo < 0. || hi > static_cast<double>(C.refRows - 1))
    throw std::out_of_range("Reference surface-current query outside domain");
  const long long total = static_cast<long long>(C.refCount) * static_cast<long long>(g.nz);
  const int blocks = static_cast<int>(std::min<long long>((total + 255) / 256, 4096));
  kReferenceCurrent<R><<<std::max(blocks, 1), 256, 0, F.stream>>>(
      F.base(8), esumZ, estarZ, F.base(2), -1.0 * g.dt / kEps0, writeEsum ? 1 : 0, C.tileCounts.get(), anyActive ? 1 : 0, ntx, nty, ntz,
      C.refTable.get(), C.refPoints.get(), static_cast<int>(C.refCount), static_cast<int>(C.refRows),
      static_cast<int>(g.nx), static_cast<int>(g.ny), static_cast<int>(g.nz), offset, shiftZ, stop);
  launchCheck("reference current");
  // J now holds values outside the particles' tiles: the next clear is full
  F.footprint.valid = false;
  F.currentZero = false;
}

template void referenceCurrentTrial<double>(CouplingState&, FieldStateT<double>&, double*, const double*, bool, bool,
                                            int, int, int, const unsigned*);
template void referenceCurrentTrial<float>(CouplingState&, FieldStateT<float>&, float*, const float*, bool, bool, int,
                                           int, int, const unsigned*);

}  // namespace aprl_gpu
