// C-ABI entry points for the field grid, its CPML boundary and the moving
// window (declared in include/aprl_gpu.h). A grid is created in the plain
// FP64 form or the FP32 form; every entry point dispatches on it. Host
// arrays are always double: the FP32 form rounds them once on upload and
// widens them on download.
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>

#include "diagnostics.hpp"
#include "handles.hpp"
#include "inlet.hpp"
#include "inletref.hpp"

namespace aprl_gpu {
namespace {
thread_local std::string lastErrorText;

template <class R>
void toDevice(hipStream_t st, R* dst, const double* host, size_t n) {
  if constexpr (std::is_same<R, double>::value) {
    hipCheck(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  } else {
    std::vector<R> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = static_cast<R>(host[i]);
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    hipCheck(hipMemcpy(dst, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice), "hipMemcpy");
  }
}

template <class R>
void toHost(hipStream_t st, double* host, const R* src, size_t n) {
  if constexpr (std::is_same<R, double>::value) {
    hipCheck(hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
  } else {
    std::vector<R> tmp(n);
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    hipCheck(hipMemcpy(tmp.data(), src, n * sizeof(R), hipMemcpyDeviceToHost), "hipMemcpy");
    for (size_t i = 0; i < n; ++i) host[i] = static_cast<double>(tmp[i]);
  }
}

template <class S>
using RealOf = typename std::remove_pointer<decltype(std::declval<S&>().comp[0].get())>::type;

// Upload caller planes; a NULL array or entry selects the zero plane (the
// persistent zero plane of an E component, or NULL where the kernel itself
// treats NULL as zero). out[a] receives the device plane to use.
template <class R>
void uploadPlanes(FieldStateT<R>& s, DeviceArray<R>* dst, const double* const* host, int first, int count,
                  const R** out) {
  for (int a = 0; a < count; ++a) {
    const int c = first + a;
    if (host && host[a]) {
      if (dst[a].size() != s.plane(c)) dst[a].resize(s.plane(c));
      toDevice(s.stream, dst[a].get(), host[a], s.plane(c));
      out[a] = dst[a].get();
    } else {
      out[a] = c < 3 ? s.zeroPlane[c].get() : nullptr;
    }
  }
}

// The window's planes of components first.. at fixed-world plane k and
// step: the caller's where given, else the device reference's
// (movingWindowFieldValue), else the zero reference.
template <class R>
void windowPlanes(FieldStateT<R>& s, DeviceArray<R>* dst, const double* const* host, int first, int count,
                  const R** out, std::ptrdiff_t k, std::size_t step) {
  uploadPlanes(s, dst, host, first, count, out);
  if (!s.reference) return;
  for (int a = 0; a < count; ++a) {
    if (host && host[a]) continue;
    const int c = first + a;
    if (dst[a].size() != s.plane(c)) dst[a].resize(s.plane(c));
    referencePlane<R>(*s.reference, static_cast<unsigned>(c), k, step, s.referenceBeta, dst[a].get(), s.stream);
    out[a] = dst[a].get();
  }
}

// The ledger's step start: a pending inserted-plane measurement (a
// charge-matched window whose inlet completion did not follow the shift)
// is taken first, then the start power.
template <class R>
void beginLedger(FieldStateT<R>& s) {
  if (s.ledger.insertPending) {
    ledgerAfterShift(s);
    s.ledger.insertPending = false;
  }
  ledgerBegin(s);
}

// Run fn on the grid's state of its precision.
template <class Fn>
void withState(aprl_fields* f, Fn&& fn) {
  if (!f) throw std::invalid_argument("Null field grid handle");
  if (f->fp32) fn(f->s32);
  else fn(f->s);
}

void checkComponent(int c) {
  if (c < 0 || c > 8) throw std::out_of_range("Yee component index must be in [0, 9)");
}

template <class R>
void createState(FieldStateT<R>& s, const aprl_grid_geometry& g, aprl_precision precision) {
  s.g = g;
  s.precision = precision;
  if (s.g.solver != 0 && s.g.solver != 1) throw std::invalid_argument("Unknown Maxwell solver");
  s.cw = s.g.solver == 1 ? cowanForZDispersion(s.g.dx, s.g.dy, s.g.dz)
                         : CowanCoef{{1, 1, 1}, {0, 0, 0}, 0, 0, 0, {1, 1, 1}};
  validateGeometry(s.g, s.cw);
  hipCheck(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
  allocateFields(s, 0);
  for (int a = 0; a < 3; ++a) {
    s.zeroPlane[a].resize(s.plane(a));
    hipCheck(hipMemsetAsync(s.zeroPlane[a].get(), 0, s.plane(a) * sizeof(R), s.stream), "hipMemsetAsync");
  }
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
}

// Exactly one of hostOut (download) and hostIn (upload) is used.
template <class R>
void historyCopy(FieldStateT<R>& s, size_t index, double* hostOut, const double* hostIn, size_t count) {
  if (!s.hasCpml) throw std::logic_error("No CPML boundary is installed");
  if (index >= 12) throw std::out_of_range("CPML checkpoint history index is out of range");
  const size_t block = s.histLayers[index] * s.histRow[index];
  if (count != s.histSlots[index] * block) throw std::invalid_argument("CPML history size mismatch");
  for (size_t slot = 0; slot < s.histSlots[index]; ++slot) {
    R* dev = s.hist[index].get() + s.historyDeviceOffset(index, slot);
    if (hostOut) toHost(s.stream, hostOut + slot * block, dev, block);
    else toDevice(s.stream, dev, hostIn + slot * block, block);
  }
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
}
}  // namespace

void setLastError(const std::string& message) { lastErrorText = message; }
}  // namespace aprl_gpu

using namespace aprl_gpu;

extern "C" {

const char* aprl_last_error(void) { return lastErrorText.c_str(); }

aprl_status aprl_fields_create(const aprl_grid_geometry* geometry, aprl_precision precision, aprl_fields** out) {
  return guarded([&] {
    if (!geometry || !out) throw std::invalid_argument("Null geometry or output handle");
    if (precision != APRL_FP64 && precision != APRL_FP32_SPLIT)
      throw std::invalid_argument("Field grid precision must be APRL_FP64 or APRL_FP32_SPLIT");
    std::unique_ptr<aprl_fields> f(new aprl_fields);
    f->fp32 = precision == APRL_FP32_SPLIT;
    if (f->fp32) createState(f->s32, *geometry, precision);
    else createState(f->s, *geometry, precision);
    *out = f.release();
  });
}

void aprl_fields_destroy(aprl_fields* fields) {
  if (!fields) return;
  hipStream_t stream = fields->fp32 ? fields->s32.stream : fields->s.stream;
  if (stream) (void)hipStreamSynchronize(stream);
  delete fields;
  if (stream) (void)hipStreamDestroy(stream);
}

size_t aprl_fields_component_size(const aprl_fields* fields, int component) {
  if (!fields || component < 0 || component > 8) return 0;
  return fields->fp32 ? fields->s32.compSize(component) : fields->s.compSize(component);
}

aprl_status aprl_fields_upload(aprl_fields* fields, int component, const double* host, size_t count) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      checkComponent(component);
      if (!host || count != s.compSize(component)) throw std::invalid_argument("Component upload size mismatch");
      toDevice(s.stream, s.base(component), host, count);
      if (component >= 6) {
        s.currentZero = false;
        s.footprint.valid = false;
      }
      hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    });
  });
}

aprl_status aprl_fields_download(aprl_fields* fields, int component, double* host, size_t count) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      checkComponent(component);
      if (!host || count != s.compSize(component)) throw std::invalid_argument("Component download size mismatch");
      toHost(s.stream, host, s.base(component), count);
    });
  });
}

aprl_status aprl_fields_download_plane(aprl_fields* fields, int component, size_t k, double* host, size_t count) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      checkComponent(component);
      if (k >= s.dims[component][2]) throw std::out_of_range("Plane index is outside the component");
      if (!host || count != s.plane(component)) throw std::invalid_argument("Plane download size mismatch");
      toHost(s.stream, host, s.base(component) + k * s.plane(component), count);
    });
  });
}

aprl_status aprl_fields_set_cpml(aprl_fields* fields, const aprl_cpml_parameters* p, size_t globalNz,
                                 size_t zOffset, int lowerZ, int upperZ) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      if (!p) throw std::invalid_argument("Null CPML parameters");
      if (s.window) throw std::logic_error("Moving-window boundary already installed");
      installCpml(s, *p, globalNz, zOffset, lowerZ != 0, upperZ != 0);
      hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    });
  });
}

size_t aprl_fields_history_size(const aprl_fields* fields, size_t index) {
  if (!fields || index >= 12) return 0;
  auto size = [&](const auto& s) -> size_t {
    return s.hasCpml ? s.histSlots[index] * s.histLayers[index] * s.histRow[index] : 0;
  };
  return fields->fp32 ? size(fields->s32) : size(fields->s);
}

aprl_status aprl_fields_history_download(aprl_fields* fields, size_t index, double* host, size_t count) {
  return guarded([&] {
    if (!host && count) throw std::invalid_argument("Null history buffer");
    withState(fields, [&](auto& s) { historyCopy(s, index, host, nullptr, count); });
  });
}

aprl_status aprl_fields_history_upload(aprl_fields* fields, size_t index, const double* host, size_t count) {
  return guarded([&] {
    if (!host && count) throw std::invalid_argument("Null history buffer");
    withState(fields, [&](auto& s) { historyCopy(s, index, nullptr, host, count); });
  });
}

aprl_status aprl_fields_clear_current(aprl_fields* fields) {
  return guarded([&] { withState(fields, [&](auto& s) { launchClearCurrent(s); }); });
}

aprl_status aprl_fields_advance_magnetic(aprl_fields* fields, const double* const* halo) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      auto f = s.ptrs();
      if (s.window) {
        if (halo) throw std::invalid_argument("The moving window supplies its own electric halos");
        if (s.phase != 1) throw std::logic_error("Moving-window magnetic phase mismatch");
        for (int a = 0; a < 3; ++a) { f.lo[a] = s.loPlane[a]; f.up[a] = s.zeroPlane[a].get(); }
        launchMagnetic(s, f);
        s.phase = 2;
        return;
      }
      if (halo) {
        for (int a = 0; a < 3; ++a) {
          if (halo[a]) uploadPlanes(s, &s.haloLo[a], &halo[a], a, 1, &f.lo[a]);
          if (halo[3 + a]) uploadPlanes(s, &s.haloUp[a], &halo[3 + a], a, 1, &f.up[a]);
        }
      }
      launchMagnetic(s, f);
    });
  });
}

aprl_status aprl_fields_advance_electric(aprl_fields* fields) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      using R = RealOf<std::remove_reference_t<decltype(s)>>;
      const auto f = s.ptrs();
      R* const none[3] = {nullptr, nullptr, nullptr};
      const R* noPlane = nullptr;
      if (s.window) {
        if (s.phase != 2) throw std::logic_error("Moving-window electric phase mismatch");
        if (!s.inletSet) throw std::logic_error("Moving-window inlet planes were not supplied");
        windowCommitElectric(s, f);
        s.phase = 3;
        return;
      }
      launchElectric(s, f, 0, none, noPlane, noPlane);
    });
  });
}

aprl_status aprl_fields_synchronize(aprl_fields* fields) {
  return guarded([&] {
    withState(fields, [&](auto& s) { hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize"); });
  });
}

aprl_status aprl_window_install(aprl_fields* fields, const aprl_cpml_parameters* p) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      using R = RealOf<std::remove_reference_t<decltype(s)>>;
      if (!p) throw std::invalid_argument("Null CPML parameters");
      if (s.window || s.hasCpml) throw std::logic_error("Moving-window boundary already installed");
      const aprl_grid_geometry& g = s.g;
      if (g.solver != 1 || p->cells[2] != 0 || p->cells[0] < 2 || p->cells[1] < 2 || p->cells[0] >= g.nx ||
          p->cells[0] >= g.nx - p->cells[0] || p->cells[1] >= g.ny || p->cells[1] >= g.ny - p->cells[1])
        throw std::invalid_argument(
            "Moving-window requires valid Cowan integer z slabs, xy-only CPML and an inlet reference");
      if (!std::isfinite(g.dt) || !std::isfinite(g.dz) || !(g.dz > 0) ||
          std::abs(kC * g.dt - g.dz) > 64 * std::numeric_limits<double>::epsilon() * g.dz)
        throw std::invalid_argument("Moving-window requires c*dt=dz");
      // Re-home the existing E/B/J contents into arrays with z slack.
      DeviceArray<R> old[9];
      for (int a = 0; a < 9; ++a) old[a].swap(s.comp[a]);
      allocateFields(s, std::max<std::size_t>(64, g.nz));
      for (int a = 0; a < 9; ++a)
        hipCheck(hipMemcpyAsync(s.base(a), old[a].get(), s.compSize(a) * sizeof(R), hipMemcpyDeviceToDevice,
                                s.stream),
                 "hipMemcpyAsync");
      s.currentZero = false;
      s.footprint.valid = false;
      hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
      installCpml(s, *p, g.nz, 0, false, false);
      s.window = true;
      s.phase = 0;
      s.established = false;
      hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    });
  });
}

aprl_status aprl_window_begin_step(aprl_fields* fields, size_t step, ptrdiff_t offset, const double* const* lower) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      if (!s.window) throw std::logic_error("No moving window is installed");
      if (s.chargeMatch) throw std::logic_error("A charge-matched inlet begins its steps with the centered reference");
      if (s.phase != 0 || (s.established && (step != s.step || offset != s.offset)))
        throw std::logic_error("Moving-window beginStep phase/origin mismatch");
      windowPlanes(s, s.haloLo, lower, 0, 3, s.loPlane, offset - 1, step);
      beginLedger(s);
      s.step = step;
      s.offset = offset;
      s.established = true;
      s.inletSet = false;
      s.phase = 1;
    });
  });
}

aprl_status aprl_window_set_charge_match(aprl_fields* fields, int enable) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      if (!s.window) throw std::logic_error("No moving window is installed");
      s.chargeMatch = enable != 0;
    });
  });
}

aprl_status aprl_window_begin_step_centered(aprl_fields* fields, size_t step, ptrdiff_t offset,
                                            const double* const* lower, const double* ez_at_offset) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      using R = RealOf<std::remove_reference_t<decltype(s)>>;
      if (!s.window) throw std::logic_error("No moving window is installed");
      if (!s.chargeMatch) throw std::logic_error("The window has no charge-matched inlet");
      if (s.phase != 0 || (s.established && (step != s.step || offset != s.offset)))
        throw std::logic_error("Moving-window beginStep phase/origin mismatch");
      // the halo planes are private copies: the Ez plane is corrected in place
      for (int a = 0; a < 3; ++a) {
        if (s.haloLo[a].size() != s.plane(a)) s.haloLo[a].resize(s.plane(a));
        if (lower && lower[a]) toDevice(s.stream, s.haloLo[a].get(), lower[a], s.plane(a));
        else if (s.reference)
          referencePlane<R>(*s.reference, static_cast<unsigned>(a), offset - 1, step, s.referenceBeta,
                            s.haloLo[a].get(), s.stream);
        else hipCheck(hipMemsetAsync(s.haloLo[a].get(), 0, s.plane(a) * sizeof(R), s.stream), "hipMemsetAsync");
        s.loPlane[a] = s.haloLo[a].get();
      }
      if (s.refEz.size() != s.plane(2)) s.refEz.resize(s.plane(2));
      if (ez_at_offset) toDevice(s.stream, s.refEz.get(), ez_at_offset, s.plane(2));
      else if (s.reference)
        referencePlane<R>(*s.reference, 2u, offset, step, s.referenceBeta, s.refEz.get(), s.stream);
      else hipCheck(hipMemsetAsync(s.refEz.get(), 0, s.plane(2) * sizeof(R), s.stream), "hipMemsetAsync");
      centerInletHalo(s, s.haloLo[2].get(), s.refEz.get());
      beginLedger(s);
      s.step = step;
      s.offset = offset;
      s.established = true;
      s.inletSet = false;
      s.phase = 1;
    });
  });
}

aprl_status aprl_window_set_inlet(aprl_fields* fields, const double* ex, const double* ey) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      if (!s.window) throw std::logic_error("No moving window is installed");
      const double* const host[2] = {ex, ey};
      windowPlanes(s, s.inlet, host, 0, 2, s.inletPlane, s.offset, s.step + 1);
      s.inletSet = true;
    });
  });
}

aprl_status aprl_window_set_reference(aprl_fields* fields, aprl_inlet_reference* reference, double translation_beta) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      if (!s.window) throw std::logic_error("No moving window is installed");
      if (reference) {
        const aprl_grid_geometry& r = reference->g;
        if (r.nx != s.g.nx || r.ny != s.g.ny || r.dx != s.g.dx || r.dy != s.g.dy || r.dz != s.g.dz ||
            r.dt != s.g.dt || r.solver != s.g.solver)
          throw std::invalid_argument("Blocked reference fill geometry/offset mismatch");
        if (!std::isfinite(translation_beta)) throw std::invalid_argument("Reference translation beta must be finite");
      }
      s.reference = reference;
      s.referenceBeta = translation_beta;
    });
  });
}

aprl_status aprl_window_shift(aprl_fields* fields, size_t newStep, ptrdiff_t newOffset, const double* const* planes) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      using R = RealOf<std::remove_reference_t<decltype(s)>>;
      if (!s.window) throw std::logic_error("No moving window is installed");
      if (s.phase != 3 || newStep != s.step + 1 || newOffset != s.offset - 1)
        throw std::logic_error("Moving-window requires one left shift after each complete step");
      const R* dev[6];
      windowPlanes(s, s.shiftPlanes, planes, 0, 6, dev, newOffset, newStep);
      ledgerBeforeShift(s);
      windowShift(s, dev);
      // the charge-matched inlet completes the new plane first
      if (s.chargeMatch && s.ledger.enabled) s.ledger.insertPending = true;
      else ledgerAfterShift(s);
      s.step = newStep;
      s.offset = newOffset;
      s.phase = 0;
    });
  });
}

aprl_status aprl_window_enable_energy_ledger(aprl_fields* fields, int enable) {
  return guarded([&] { withState(fields, [&](auto& s) { ledgerEnable(s, enable != 0); }); });
}

aprl_status aprl_window_finish_energy_step(aprl_fields* fields) {
  return guarded([&] {
    withState(fields, [&](auto& s) {
      if (!s.window) throw std::logic_error("No moving window is installed");
      ledgerFinish(s);
    });
  });
}

aprl_status aprl_window_energy_ledger(aprl_fields* fields, double* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("Null ledger output");
    withState(fields, [&](auto& s) {
      if (s.ledger.insertPending && s.phase == 0) {
        ledgerAfterShift(s);
        s.ledger.insertPending = false;
      }
      ledgerRead(s, out);
    });
  });
}

aprl_status aprl_profile_enable(aprl_fields* fields, int enable) {
  return guarded([&] { withState(fields, [&](auto& s) { s.profiler.enabled = enable != 0; }); });
}

aprl_status aprl_profile_query(aprl_fields* fields, const char* kernel, uint64_t* launches, double* totalMs) {
  return guarded([&] {
    if (!kernel || !launches || !totalMs) throw std::invalid_argument("Null profile query argument");
    withState(fields, [&](auto& s) {
      const Profiler::Stat st = s.profiler.query(kernel);
      *launches = st.launches;
      *totalMs = st.ms;
    });
  });
}

aprl_status aprl_profile_reset(aprl_fields* fields) {
  return guarded([&] { withState(fields, [&](auto& s) { s.profiler.reset(); }); });
}

}  // extern "C"
