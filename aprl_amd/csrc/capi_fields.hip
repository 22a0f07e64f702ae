// C-ABI entry points for the field grid, its CPML boundary and the moving
// window (declared in include/aprl_gpu.h).
#include <cstring>
#include <memory>

#include "handles.hpp"

namespace aprl_gpu {
namespace {
thread_local std::string lastErrorText;
}
void setLastError(const std::string& message) { lastErrorText = message; }

static void upload(hipStream_t s, double* dst, const double* host, size_t n) {
  hipCheck(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
}

// Upload caller planes; a NULL array or entry selects the zero plane (the
// persistent zero plane of an E component, or NULL where the kernel itself
// treats NULL as zero). out[a] receives the device plane to use.
static void uploadPlanes(FieldState& s, DeviceArray<double>* dst, const double* const* host, int first, int count,
                         const double** out) {
  for (int a = 0; a < count; ++a) {
    const int c = first + a;
    if (host && host[a]) {
      if (dst[a].size() != s.plane(c)) dst[a].resize(s.plane(c));
      upload(s.stream, dst[a].get(), host[a], s.plane(c));
      out[a] = dst[a].get();
    } else {
      out[a] = c < 3 ? s.zeroPlane[c].get() : nullptr;
    }
  }
}

static FieldState& S(aprl_fields* f) {
  if (!f) throw std::invalid_argument("Null field grid handle");
  return f->s;
}

static void checkComponent(int c) {
  if (c < 0 || c > 8) throw std::out_of_range("Yee component index must be in [0, 9)");
}
}  // namespace aprl_gpu

using namespace aprl_gpu;

extern "C" {

const char* aprl_last_error(void) { return lastErrorText.c_str(); }

aprl_status aprl_fields_create(const aprl_grid_geometry* geometry, aprl_precision precision, aprl_fields** out) {
  return guarded([&] {
    if (!geometry || !out) throw std::invalid_argument("Null geometry or output handle");
    if (precision != APRL_FP64) throw std::invalid_argument("Field grid precision must be APRL_FP64 for this entry point");
    std::unique_ptr<aprl_fields> f(new aprl_fields);
    FieldState& s = f->s;
    s.g = *geometry;
    s.precision = precision;
    if (s.g.solver != 0 && s.g.solver != 1) throw std::invalid_argument("Unknown Maxwell solver");
    s.cw = s.g.solver == 1 ? cowanForZDispersion(s.g.dx, s.g.dy, s.g.dz) : CowanCoef{{1, 1, 1}, {0, 0, 0}, 0, 0, 0, {1, 1, 1}};
    validateGeometry(s.g, s.cw);
    hipCheck(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
    allocateFields(s, 0);
    for (int a = 0; a < 3; ++a) {
      s.zeroPlane[a].resize(s.plane(a));
      hipCheck(hipMemsetAsync(s.zeroPlane[a].get(), 0, s.plane(a) * sizeof(double), s.stream), "hipMemsetAsync");
    }
    hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    *out = f.release();
  });
}

void aprl_fields_destroy(aprl_fields* fields) {
  if (!fields) return;
  (void)hipStreamSynchronize(fields->s.stream);
  hipStream_t stream = fields->s.stream;
  delete fields;
  if (stream) (void)hipStreamDestroy(stream);
}

size_t aprl_fields_component_size(const aprl_fields* fields, int component) {
  if (!fields || component < 0 || component > 8) return 0;
  return fields->s.compSize(component);
}

aprl_status aprl_fields_upload(aprl_fields* fields, int component, const double* host, size_t count) {
  return guarded([&] {
    FieldState& s = S(fields);
    checkComponent(component);
    if (!host || count != s.compSize(component)) throw std::invalid_argument("Component upload size mismatch");
    upload(s.stream, s.base(component), host, count);
    if (component >= 6) {
      s.currentZero = false;
      s.footprint.valid = false;
    }
    hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  });
}

aprl_status aprl_fields_download(aprl_fields* fields, int component, double* host, size_t count) {
  return guarded([&] {
    FieldState& s = S(fields);
    checkComponent(component);
    if (!host || count != s.compSize(component)) throw std::invalid_argument("Component download size mismatch");
    hipCheck(hipMemcpyAsync(host, s.base(component), count * sizeof(double), hipMemcpyDeviceToHost, s.stream),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  });
}

aprl_status aprl_fields_set_cpml(aprl_fields* fields, const aprl_cpml_parameters* p, size_t globalNz,
                                 size_t zOffset, int lowerZ, int upperZ) {
  return guarded([&] {
    FieldState& s = S(fields);
    if (!p) throw std::invalid_argument("Null CPML parameters");
    if (s.window) throw std::logic_error("Moving-window boundary already installed");
    installCpml(s, *p, globalNz, zOffset, lowerZ != 0, upperZ != 0);
    hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  });
}

size_t aprl_fields_history_size(const aprl_fields* fields, size_t index) {
  if (!fields || index >= 12 || !fields->s.hasCpml) return 0;
  return fields->s.histSlots[index] * fields->s.histLayers[index] * fields->s.histRow[index];
}

// Exactly one of hostOut (download) and hostIn (upload) is used.
static void historyCopy(FieldState& s, size_t index, double* hostOut, const double* hostIn, size_t count) {
  if (!s.hasCpml) throw std::logic_error("No CPML boundary is installed");
  if (index >= 12) throw std::out_of_range("CPML checkpoint history index is out of range");
  const size_t block = s.histLayers[index] * s.histRow[index];
  if (count != s.histSlots[index] * block) throw std::invalid_argument("CPML history size mismatch");
  for (size_t slot = 0; slot < s.histSlots[index]; ++slot) {
    double* dev = s.hist[index].get() + s.historyDeviceOffset(index, slot);
    if (hostOut)
      hipCheck(hipMemcpyAsync(hostOut + slot * block, dev, block * sizeof(double), hipMemcpyDeviceToHost, s.stream),
               "hipMemcpyAsync");
    else
      hipCheck(hipMemcpyAsync(dev, hostIn + slot * block, block * sizeof(double), hipMemcpyHostToDevice, s.stream),
               "hipMemcpyAsync");
  }
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
}

aprl_status aprl_fields_history_download(aprl_fields* fields, size_t index, double* host, size_t count) {
  return guarded([&] {
    if (!host && count) throw std::invalid_argument("Null history buffer");
    historyCopy(S(fields), index, host, nullptr, count);
  });
}

aprl_status aprl_fields_history_upload(aprl_fields* fields, size_t index, const double* host, size_t count) {
  return guarded([&] {
    if (!host && count) throw std::invalid_argument("Null history buffer");
    historyCopy(S(fields), index, nullptr, host, count);
  });
}

aprl_status aprl_fields_clear_current(aprl_fields* fields) {
  return guarded([&] { launchClearCurrent(S(fields)); });
}

aprl_status aprl_fields_advance_magnetic(aprl_fields* fields, const double* const* halo) {
  return guarded([&] {
    FieldState& s = S(fields);
    FieldPtrs f = s.ptrs();
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
}

aprl_status aprl_fields_advance_electric(aprl_fields* fields) {
  return guarded([&] {
    FieldState& s = S(fields);
    const FieldPtrs f = s.ptrs();
    double* const none[3] = {nullptr, nullptr, nullptr};
    if (s.window) {
      if (s.phase != 2) throw std::logic_error("Moving-window electric phase mismatch");
      if (!s.inletSet) throw std::logic_error("Moving-window inlet planes were not supplied");
      launchElectric(s, f, 0, none, s.inletPlane[0], s.inletPlane[1]);
      s.phase = 3;
      return;
    }
    launchElectric(s, f, 0, none, nullptr, nullptr);
  });
}

aprl_status aprl_fields_synchronize(aprl_fields* fields) {
  return guarded([&] { hipCheck(hipStreamSynchronize(S(fields).stream), "hipStreamSynchronize"); });
}

aprl_status aprl_window_install(aprl_fields* fields, const aprl_cpml_parameters* p) {
  return guarded([&] {
    FieldState& s = S(fields);
    if (!p) throw std::invalid_argument("Null CPML parameters");
    if (s.window || s.hasCpml) throw std::logic_error("Moving-window boundary already installed");
    const aprl_grid_geometry& g = s.g;
    if (g.solver != 1 || p->cells[2] != 0 || p->cells[0] < 2 || p->cells[1] < 2 || p->cells[0] >= g.nx ||
        p->cells[0] >= g.nx - p->cells[0] || p->cells[1] >= g.ny || p->cells[1] >= g.ny - p->cells[1])
      throw std::invalid_argument("Moving-window requires valid Cowan integer z slabs, xy-only CPML and an inlet reference");
    if (!std::isfinite(g.dt) || !std::isfinite(g.dz) || !(g.dz > 0) ||
        std::abs(kC * g.dt - g.dz) > 64 * std::numeric_limits<double>::epsilon() * g.dz)
      throw std::invalid_argument("Moving-window requires c*dt=dz");
    // Re-home the existing E/B/J contents into arrays with z slack.
    DeviceArray<double> old[9];
    for (int a = 0; a < 9; ++a) old[a].swap(s.comp[a]);
    allocateFields(s, std::max<std::size_t>(64, g.nz));
    for (int a = 0; a < 9; ++a)
      hipCheck(hipMemcpyAsync(s.base(a), old[a].get(), s.compSize(a) * sizeof(double), hipMemcpyDeviceToDevice,
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
}

aprl_status aprl_window_begin_step(aprl_fields* fields, size_t step, ptrdiff_t offset, const double* const* lower) {
  return guarded([&] {
    FieldState& s = S(fields);
    if (!s.window) throw std::logic_error("No moving window is installed");
    if (s.phase != 0 || (s.established && (step != s.step || offset != s.offset)))
      throw std::logic_error("Moving-window beginStep phase/origin mismatch");
    uploadPlanes(s, s.haloLo, lower, 0, 3, s.loPlane);
    s.step = step;
    s.offset = offset;
    s.established = true;
    s.inletSet = false;
    s.phase = 1;
  });
}

aprl_status aprl_window_set_inlet(aprl_fields* fields, const double* ex, const double* ey) {
  return guarded([&] {
    FieldState& s = S(fields);
    if (!s.window) throw std::logic_error("No moving window is installed");
    const double* const host[2] = {ex, ey};
    uploadPlanes(s, s.inlet, host, 0, 2, s.inletPlane);
    s.inletSet = true;
  });
}

aprl_status aprl_window_shift(aprl_fields* fields, size_t newStep, ptrdiff_t newOffset, const double* const* planes) {
  return guarded([&] {
    FieldState& s = S(fields);
    if (!s.window) throw std::logic_error("No moving window is installed");
    if (s.phase != 3 || newStep != s.step + 1 || newOffset != s.offset - 1)
      throw std::logic_error("Moving-window requires one left shift after each complete step");
    const double* dev[6];
    uploadPlanes(s, s.shiftPlanes, planes, 0, 6, dev);
    windowShift(s, dev);
    s.step = newStep;
    s.offset = newOffset;
    s.phase = 0;
  });
}

aprl_status aprl_profile_enable(aprl_fields* fields, int enable) {
  return guarded([&] { S(fields).profiler.enabled = enable != 0; });
}

aprl_status aprl_profile_query(aprl_fields* fields, const char* kernel, uint64_t* launches, double* totalMs) {
  return guarded([&] {
    if (!kernel || !launches || !totalMs) throw std::invalid_argument("Null profile query argument");
    const Profiler::Stat st = S(fields).profiler.query(kernel);
    *launches = st.launches;
    *totalMs = st.ms;
  });
}

aprl_status aprl_profile_reset(aprl_fields* fields) {
  return guarded([&] { S(fields).profiler.reset(); });
}

}  // extern "C"
