/* aprl_gpu.h — C-ABI of the MI355X (gfx950) implementation of APRL's
 * per-timestep electromagnetic PIC loop.
 *
 * Plain pointers and sizes only. Every call is stream-ordered on the
 * instance's own HIP stream and returns an aprl_status; aprl_last_error()
 * gives the message of the last failure on the calling thread. Status codes
 * map one-to-one onto the reference's exception types so that a C++ shim can
 * rethrow the same type and text (eb_field.cpp:880-941, eb_field.cpp:976-990,
 * simulation_source_coupling.cpp:54, moving_window_fields.cpp:208-209).
 *
 * Array layouts are the reference's: each Yee component is x-fastest,
 * index (k*ny + j)*nx + i, with the staggered shapes of eb_field.cpp:217-225:
 *   Ex nx x (ny+1) x (nz+1)   Ey (nx+1) x ny x (nz+1)   Ez (nx+1) x (ny+1) x nz
 *   Bx (nx+1) x ny x nz       By nx x (ny+1) x nz       Bz nx x ny x (nz+1)
 *   J  collocated with E.
 * CPML histories use EBConvolutionalPML::checkpointHistory order and layout
 * (eb_cpml.cpp:546-577).
 */
#ifndef APRL_GPU_H
#define APRL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum aprl_status {
  APRL_OK = 0,
  APRL_INVALID_ARGUMENT = 1, /* std::invalid_argument */
  APRL_OUT_OF_RANGE = 2,     /* std::out_of_range */
  APRL_RUNTIME_ERROR = 3,    /* std::runtime_error */
  APRL_LOGIC_ERROR = 4,      /* std::logic_error */
  APRL_DEVICE_ERROR = 5      /* HIP runtime failure */
} aprl_status;

typedef enum aprl_precision {
  APRL_FP64 = 0,      /* plain form: everything in double */
  APRL_FP32_SPLIT = 1 /* fields/currents/momenta float, positions cell+offset */
} aprl_precision;

/* EBGridGeometry, eb_field.h:54-70. solver: 0 = Yee, 1 = Cowan-z. */
typedef struct aprl_grid_geometry {
  size_t nx, ny, nz;
  double dx, dy, dz, dt;
  int solver;
} aprl_grid_geometry;

/* EBCPMLParameters, eb_cpml.h:12-24. */
typedef struct aprl_cpml_parameters {
  size_t cells[3];
  double polynomial_order;
  double target_reflection;
  double kappa_max;
  double alpha_fraction;
} aprl_cpml_parameters;

typedef struct aprl_fields aprl_fields;

const char* aprl_last_error(void);

/* EBFieldGrid::EBFieldGrid, eb_field.h:204 / eb_field.cpp:211-229: allocates
 * the nine components on the current device, zeroed, with the default PEC
 * boundary; validates the geometry (eb_field.cpp:878-943). */
aprl_status aprl_fields_create(const aprl_grid_geometry* geometry, aprl_precision precision,
                               aprl_fields** out);
void aprl_fields_destroy(aprl_fields* fields);

/* YeeComponent::data() copies (eb_field.h:129-131); component 0..8 =
 * Ex Ey Ez Bx By Bz Jx Jy Jz; count must equal the component size. */
aprl_status aprl_fields_upload(aprl_fields* fields, int component, const double* host, size_t count);
aprl_status aprl_fields_download(aprl_fields* fields, int component, double* host, size_t count);
/* One z plane k of a component (count = nx_a * ny_a values, x fastest): the
 * plane read of LabDetectorManager::sampleFieldPlanes (lab_detectors.cpp:
 * 2426-2487) without a whole-component copy. */
aprl_status aprl_fields_download_plane(aprl_fields* fields, int component, size_t k, double* host, size_t count);
size_t aprl_fields_component_size(const aprl_fields* fields, int component);

/* EBConvolutionalPML::EBConvolutionalPML + EBFieldGrid::setBoundary
 * (eb_cpml.h:33-39, eb_field.h:231). */
aprl_status aprl_fields_set_cpml(aprl_fields* fields, const aprl_cpml_parameters* parameters,
                                 size_t global_nz, size_t local_z_offset, int lower_z_physical,
                                 int upper_z_physical);
/* EBConvolutionalPML::checkpointHistory(index), eb_cpml.h:51-53. */
size_t aprl_fields_history_size(const aprl_fields* fields, size_t index);
aprl_status aprl_fields_history_download(aprl_fields* fields, size_t index, double* host, size_t count);
aprl_status aprl_fields_history_upload(aprl_fields* fields, size_t index, const double* host, size_t count);

/* EBFieldGrid::clearCurrent, eb_field.cpp:271-295. */
aprl_status aprl_fields_clear_current(aprl_fields* fields);
/* EBFieldGrid::advanceMagnetic(const EBElectricHaloView&), eb_field.cpp:302-309:
 * the B curl (Cowan or Yee) fused with the boundary's afterMagneticUpdate.
 * halo[0..2] = lower Ex,Ey,Ez planes, halo[3..5] = upper; NULL array or
 * NULL entries select the PEC parity extension (eb_field.cpp:388-411). */
aprl_status aprl_fields_advance_magnetic(aprl_fields* fields, const double* const* halo);
/* EBFieldGrid::advanceElectric, eb_field.cpp:582-627: curl, -dt/eps0 J, and
 * the boundary's afterElectricUpdate (CPML-E, PEC, window inlet). */
aprl_status aprl_fields_advance_electric(aprl_fields* fields);
aprl_status aprl_fields_synchronize(aprl_fields* fields);

/* MovingWindowFields (moving_window_fields.h:72-134), single owner of the
 * whole window. install: xy-only CPML, no z PEC (moving_window_fields.cpp:193). */
aprl_status aprl_window_install(aprl_fields* fields, const aprl_cpml_parameters* parameters);
/* beginStep + physicalElectricHalos (moving_window_fields.cpp:207-244):
 * lower[a] = reference(a,i,j,offset-1,step) planes of Ex,Ey,Ez; NULL = zero
 * reference. The upper halo is the sacrificial zero plane. */
aprl_status aprl_window_begin_step(aprl_fields* fields, size_t step, ptrdiff_t offset,
                                   const double* const* lower);
/* Charge-matched inlet (inlet_charge_match: ChargeMatchedInlet,
 * charge_matched_inlet.h:16-72, installed by setNewLayerCompletion with a
 * centered normal reference, moving_window_fields.cpp:551-555). When
 * enabled, steps begin with aprl_window_begin_step_centered and every
 * shift is followed by aprl_window_complete_inlet. Not available to
 * aprl_window_step_batch. */
aprl_status aprl_window_set_charge_match(aprl_fields* fields, int enable);
/* beginStep with the centered normal reference (moving_window_fields.cpp:
 * 225-234): as aprl_window_begin_step, then on the interior aperture the
 * lower Ez halo -= Ez(i,j,0) - ez_at_offset, where ez_at_offset =
 * reference(2,i,j,offset,step) (NULL = zero reference). */
aprl_status aprl_window_begin_step_centered(aprl_fields* fields, size_t step, ptrdiff_t offset,
                                            const double* const* lower, const double* ez_at_offset);
/* Inlet overwrite values Ex/Ey(k=0) := reference(a,i,j,offset,step+1)
 * (moving_window_fields.cpp:276-280); NULL = zero reference. */
aprl_status aprl_window_set_inlet(aprl_fields* fields, const double* ex, const double* ey);
/* shift (moving_window_fields.cpp:439-518, resetCurrent = true): every E/B
 * component and every CPML history moves up one z plane; planes[a] =
 * reference(a,i,j,new_offset,new_step) for a = 0..5 (NULL = zero). */
aprl_status aprl_window_shift(aprl_fields* fields, size_t new_step, ptrdiff_t new_offset,
                              const double* const* planes);

/* ---------------------------------------------------------------------------
 * Particles and the conservative-source coupling.
 */
typedef struct aprl_particles aprl_particles;
typedef struct aprl_coupling aprl_coupling;

/* Device store of RelativisticParticleSI (eb_particles.h:12-25) and of the
 * CPML carriers (ParticleCPMLCarrier, particle_boundary.h:76-84) on the
 * field grid's device and stream. Positions are absolute box coordinates in
 * metres, u = gamma v / c. The store's order changes at every coupled step
 * (stable sort by cell tile); ids travel with their particles. A particle
 * that enters the CPML or passes the moving-window exit plane becomes a
 * carrier in place: it keeps its id, position and u (its straight-line
 * velocity is u c / gamma) and its weight slot holds the carrier's current
 * weight from then on. Particles and carriers that leave the domain are
 * removed. Count and downloads cover the live entries (particles and
 * carriers) in store order.
 * On an APRL_FP32_SPLIT grid the store keeps per axis the lattice cell
 * floor(x / h) (int32, lattice anchored at x = 0) and the FP32 offset inside
 * it, and u, charge, mass and weight in FP32; upload and download convert. */
aprl_status aprl_particles_create(aprl_fields* fields, size_t capacity, aprl_particles** out);
void aprl_particles_destroy(aprl_particles* particles);
/* position3/u3: interleaved (x,y,z) per particle; n <= capacity grows it. */
aprl_status aprl_particles_upload(aprl_particles* particles, size_t n, const double* position3, const double* u3,
                                  const double* charge, const double* mass, const double* weight,
                                  const uint64_t* id);
/* Any output pointer may be NULL. */
aprl_status aprl_particles_download(aprl_particles* particles, double* position3, double* u3, double* charge,
                                    double* mass, double* weight, uint64_t* id);
size_t aprl_particles_count(const aprl_particles* particles);
/* Entry state per live entry, in download order: 0 particle, 1 CPML carrier,
 * 2 held particle (no counterpart in the reference, which keeps carriers in
 * a separate list and flags held particles in RelativisticParticleSI). */
aprl_status aprl_particles_download_states(aprl_particles* particles, int* state);
/* RelativisticParticleSI::frozen / frozenBetaZ (eb_particles.h:21-22) for
 * the entries in upload order, before the store's first step: a held
 * particle keeps its momentum and moves without fields (z += c dt beta, or
 * ballistically, aprl_coupling_set_hold_transverse) until its lab z reaches
 * the activation plane (prepareParticle's
 * hold-transverse branch, simulation.cpp:4951-5000); upstream of a moving
 * window it deposits only its shaped support until it crosses the inlet. */
aprl_status aprl_particles_set_held(aprl_particles* particles, const int* frozen, const double* beta);
/* RelativisticParticleSI::frozenBetaZ per live entry in download order (0
 * for entries that are not held), the cached held speed of
 * eb_particles.h:22. */
aprl_status aprl_particles_download_held(aprl_particles* particles, double* beta);

/* The generated-gaussian beam of ParticleInitializer (generatedGaussian,
 * particle_initializer.cpp:60-100), drawn on the device straight into the
 * store, replacing its contents: entry k of `count` (global index
 * g = first + k) gets position center + sigma_position * gaussian(seed, g,
 * axis) and proper velocity mean_u + sigma_u * gaussian(seed ^
 * 0xd6e8feb86659fd93, g, axis) (splitmix64 uniforms and a Box-Muller
 * cosine, particle_initializer.cpp:24-50), charge -e * electrons / total,
 * mass m_e * electrons / total, weight 1 and id g + 1, as an active
 * particle. With boost_gamma > 1 every proper velocity is then mapped into
 * the frame boosted by boost_gamma along z (the light-front form of
 * BoostFrameTransform::properVelocityLabToBox, boostframe.cpp:110-153). The
 * device's log and cos stand in for glibc's, so the draw is deterministic
 * but not bitwise the reference's. */
typedef struct aprl_gaussian_beam {
  double center[3];         /* config.center + positionOffset, m */
  double sigma_position[3]; /* m */
  double mean_u[3];         /* mean proper velocity (gamma beta along the direction) */
  double sigma_u[3];
  double electrons;         /* represented electrons of the whole beam */
  uint64_t seed;            /* beam.input.random_seed */
  uint64_t first;           /* global index of the first entry drawn */
  uint64_t total;           /* macroparticles of the whole beam */
  double boost_gamma;       /* 1: no boost */
} aprl_gaussian_beam;
aprl_status aprl_particles_generate_gaussian(aprl_particles* particles, const aprl_gaussian_beam* beam,
                                             size_t count);

/* Longitudinal source shape (longitudinal_shape.h:16). */
typedef enum aprl_shape {
  APRL_SHAPE_QUINTIC_COMPENSATED = 0,
  APRL_SHAPE_LINEAR_BINOMIAL3 = 2
} aprl_shape;

/* SIMagneticElement (eb_sources.h:107-144); type 0 planar undulator,
 * 1 uniform dipole, 2 quadrupole. Lab frame, SI. */
typedef struct aprl_magnet {
  int type;
  double center[3];
  double length, period, peak_field, polarization_angle;
  int gaussian_fringe;
  double fringe_relative_cutoff, quadrupole_gradient, end_ramp_periods, undulator_phase;
} aprl_magnet;

/* BoostFrameTransform::setOriginsFromGamma (boostframe.h:25-28), c0 = c. */
typedef struct aprl_boost {
  double gamma, lab_time_origin, lab_z_origin, box_z_origin;
} aprl_boost;

typedef struct aprl_step_report {
  unsigned iterations; /* Picard trials */
  double force;        /* self-force defect of the last trial */
  double work;         /* relative fixed-point work defect */
  unsigned tiles;      /* occupied cell tiles (8x8x4 cells) of the step */
  unsigned chunks;     /* particle work chunks (at most 128 particles of one tile) */
  unsigned active_tiles; /* tiles whose nodes the preview and the current reduction
                            touch: an occupied tile within one tile in x and y and
                            the shape's reach in z */
} aprl_step_report;

/* ConservativeSourceCoupling (source_coupling.h:12-41) bound to a field
 * grid. Shapes: APRL_SHAPE_QUINTIC_COMPENSATED, or
 * APRL_SHAPE_LINEAR_BINOMIAL3 with transverse_x_binomial3 = 1 (the
 * production F3 x-smoothing); transverse_x_binomial3 is ignored for the
 * quintic shape. */
aprl_status aprl_coupling_create(aprl_fields* fields, aprl_shape shape, int transverse_x_binomial3,
                                 aprl_coupling** out);
void aprl_coupling_destroy(aprl_coupling* coupling);
aprl_status aprl_coupling_set_frame(aprl_coupling* coupling, const aprl_boost* frame);
/* SIFieldSourceSet::addPrescribedLabMagnet (eb_sources.h:163). */
aprl_status aprl_coupling_add_magnet(aprl_coupling* coupling, const aprl_magnet* magnet);
/* ParticleCPMLRegion cells and the moving-window lab exit plane
 * (particle_boundary.h:36-66; simulation_moving_window.cpp:380-395);
 * a non-finite exit_lab_z disables the exit plane. The carriers' current
 * damping uses the grid's installed CPML grading (polynomial order, target
 * reflection; particle_boundary.cpp:176-241). */
aprl_status aprl_coupling_set_particle_boundary(aprl_coupling* coupling, const size_t* cpml_cells,
                                                double exit_lab_z);
/* reference_surface_current: TransverseReferenceCurrent
 * (transverse_reference_current.cpp:18-26, 104-120) with its prepared table
 * in memory instead of a file: table[row * count + p] is the equivalent
 * surface current at interface point p for reference z cell `row`, points
 * in the reference's order (for i = g0..nx-g0: (i, g1), (i, ny-g1); then for
 * j = g1+1..ny-g1-1: (g0, j), (nx-g0, j); count = 2 (nx - 2 g0 + 1) +
 * 2 (ny - 2 g1 - 1)). Every Picard trial and the commit add to Jz the table
 * interpolated at z = offset + k - translation_beta (step + 1), offset and
 * step being the window's current ones. table = NULL disables it. */
aprl_status aprl_coupling_set_reference_current(aprl_coupling* coupling, const double* table, size_t rows,
                                               size_t count, const size_t* guards, double translation_beta);
/* Simulation::completeMovingInletLayer (simulation_moving_window.cpp:
 * 312-324) after aprl_window_shift to the window origin `origin`: the new
 * layer's Ez on the interior aperture := Ez(k=1) + dz (div_xy E(k=1) -
 * rho / eps0), rho being the shaped inlet charge of the store's particles
 * (carriers excluded) whose longitudinal shape reaches node k = 1. */
aprl_status aprl_window_complete_inlet(aprl_fields* fields, aprl_coupling* coupling, aprl_particles* particles,
                                       const double* origin);
/* Lab z of the held particles' activation plane (config
 * reference.activation_plane_z); NaN disables the release. */
aprl_status aprl_coupling_set_activation(aprl_coupling* coupling, double activation_lab_z);
/* config reference.hold_transverse_until_activation (prepareParticle,
 * simulation.cpp:4951-4965): hold = 1 moves a held particle z += c dt beta
 * with its transverse position held (the library's default); hold = 0 moves
 * it ballistically, x += (c dt / gamma) u (RelativisticBorisPusher::
 * pushPosition, eb_particles.cpp:131-139), the reference's default. */
aprl_status aprl_coupling_set_hold_transverse(aprl_coupling* coupling, int hold);

/* Pending admission queue of the moving window (PendingParticleQueue,
 * pending_particle_queue.h:25-76): held particles upstream of the inlet
 * wait in host memory, each with the step in which its held path first
 * reaches the inlet, and enter a device store as held entries. */
typedef struct aprl_pending aprl_pending;
/* PendingWindowGeometry (pending_particle_queue.h:16-23): the inlet of step
 * s is fma(initial_offset - s, dz, reference_origin_z); a held particle
 * moves held_distance_per_beta * beta per step (pass c * dt). */
typedef struct aprl_pending_geometry {
  double reference_origin_z;
  int64_t initial_offset;
  double dz;
  double held_distance_per_beta;
  uint64_t maximum_step; /* last step of the run */
} aprl_pending_geometry;
/* repeatHeldAddition (pending_particle_queue.h:11-14): the value of
 * `count` separately rounded binary64 additions x += increment. */
aprl_status aprl_repeat_held_addition(double initial, double increment, uint64_t count, double* out);
/* PendingParticleQueue(geometry, initialStep) (pending_particle_queue.h:
 * 30-31); invalid geometry or horizon -> APRL_INVALID_ARGUMENT. */
aprl_status aprl_pending_create(const aprl_pending_geometry* geometry, uint64_t initial_step, aprl_pending** out);
void aprl_pending_destroy(aprl_pending* queue);
/* extract (pending_particle_queue.h:33-36, .cpp:148-175) over host arrays:
 * held particles upstream of the initial inlet with beta >= 0 that do not
 * take part in the initial step move to the queue; the arrays are compacted
 * in place, order kept, and *retained is the number left. Once per queue. */
aprl_status aprl_pending_extract(aprl_pending* queue, size_t n, double* x3, double* u3, double* q, double* m,
                                 double* w, uint64_t* id, int* frozen, double* beta, size_t* retained);
/* admit (pending_particle_queue.h:38-43, .cpp:184-206) into a device store,
 * before the step: appends the waiting particles due at or before
 * min(step + look_ahead, maximum_step) as held entries, in due order, with
 * z at the start of `step`. look_ahead = 0 is the reference's admit(step).
 * look_ahead = 8 also holds in the store the particles whose shaped support
 * the reference deposits through visitNear(step, 8) (simulation.cpp:
 * 4170-4172) and counts in the inlet charge (simulation_moving_window.cpp:
 * 317): an upstream held entry deposits exactly that support. Steps must
 * come in order and skip no due step (APRL_LOGIC_ERROR). */
aprl_status aprl_pending_admit(aprl_pending* queue, uint64_t step, uint64_t look_ahead, aprl_particles* particles,
                               size_t* admitted);
/* visitNear (pending_particle_queue.h:48-49) as arrays: the first
 * `capacity` waiting particles due by step + look_ahead, z at the start of
 * `step`; *n = how many there are. */
aprl_status aprl_pending_view(const aprl_pending* queue, uint64_t step, uint64_t look_ahead, size_t capacity,
                              double* x3, uint64_t* id, size_t* n);
/* size, nextDueStep (maximum_step + 1: none due within the horizon),
 * admittedCount (pending_particle_queue.h:51-62); NULL outputs skipped. */
aprl_status aprl_pending_state(const aprl_pending* queue, uint64_t* waiting, uint64_t* next_due, uint64_t* admitted);
/* origin (pending_particle_queue.h:56): the inlet z of a step. */
aprl_status aprl_pending_origin(const aprl_pending* queue, uint64_t step, double* z);
/* Cumulative particle-boundary counters since the coupling was created:
 * counts[20] / charges[20] (C) hold the CPML entries per face (0-5, faces
 * x- x+ y- y+ z- z+), the outer-face exits of particles (6-11) and of
 * carriers (12-17, charge times end weight), the moving-window exits (18)
 * and the cropped carriers (19) (simulation.h cpmlEntryCount_,
 * directOuterCount_, carrierOuterCount_; simulation_moving_window.cpp:
 * 363-367, 396-398). The carriers beyond a shifted window's upper face are
 * cropped by the next coupled step, which checks the entries against its
 * origin as finishMovingWindowStep does after the shift. */
aprl_status aprl_coupling_boundary_counters(aprl_coupling* coupling, uint64_t* counts, double* charges);
/* One Simulation::solveConservativeSourceStep + commitConservativeElectric
 * (simulation_source_coupling.cpp:16-65) for every entry of the store (the
 * Picard convergence test evaluated on the device, see
 * aprl_coupling_set_trial_batch),
 * after clear_current / begin_step / advance_magnetic / set_inlet. origin =
 * the grid's box origin (m); time = box time at the step start; substeps =
 * particle substeps per field step. Particle segments end at their first
 * boundary event and carriers move and deposit damped current as in
 * pushDepositAndMigrate (simulation.cpp:4310-4917). */
aprl_status aprl_coupling_step(aprl_coupling* coupling, aprl_particles* particles, const double* origin,
                               double time, unsigned substeps, aprl_step_report* report);
/* Optional, no counterpart in the reference (whose particle loop follows
 * the field updates on the same threads): starts the sort phase of the
 * next aprl_coupling_step (tile keys, stable tile sort, state permute,
 * chunk lists) on the coupling's own stream, ordered after everything
 * already enqueued on the grid, so that it runs while the field work
 * enqueued after it (clear_current, begin_step, advance_magnetic)
 * proceeds. Call it before those; the step then uses the prepared sort if
 * it is called with the same store, unwritten since, and the same origin,
 * and otherwise waits for it and sorts again. */
aprl_status aprl_coupling_prepare_step(aprl_coupling* coupling, aprl_particles* particles, const double* origin);
/* FP32-split grids only (no counterpart in the reference, which has no FP32
 * form): the Picard loop of aprl_coupling_step runs exactly `trials` trials
 * (2..64, default 2, the reference's minimum with active particles) in
 * place of the FP64 convergence test, whose tolerances (force < 1e-13) lie
 * below FP32 resolution; a step with no active particle (held entries and
 * carriers only) runs one, as the reference converges in one trial when no
 * particle makes a coupling call. verify = 1 adds one residual gather after the last
 * trial so that the step report carries its force and work (otherwise 0).
 * On such a grid the step's origin must lie on the cell lattice. */
aprl_status aprl_coupling_set_split_trials(aprl_coupling* coupling, unsigned trials, int verify);
/* Plain FP64 grids only (no counterpart in the reference, whose loop tests
 * convergence after every trial on the host): aprl_coupling_step enqueues
 * its first `batch` Picard trials (1..64) before it waits for the device's
 * convergence decision, then one trial at a time; a trial enqueued after
 * the one that converged returns at once, so the result does not depend on
 * the batch. 0 (the default) takes the last step's trial count. */
aprl_status aprl_coupling_set_trial_batch(aprl_coupling* coupling, unsigned batch);

/* ---------------------------------------------------------------------------
 * Batched steps of independent instances.
 */

/* One instance of a batched window step. step/offset/origin/time are the
 * instance's window step index, window offset, grid box origin (m) and box
 * time at the start of the step (the values the single-instance calls take);
 * report is written. */
typedef struct aprl_instance_step {
  aprl_fields* fields;
  aprl_coupling* coupling;
  aprl_particles* particles;
  size_t step;
  ptrdiff_t offset;
  double origin[3];
  double time;
  unsigned substeps;
  aprl_step_report report;
} aprl_instance_step;

/* No counterpart in the reference, which advances one bunch per process:
 * one whole window step of Simulation::solve (simulation.cpp:518-583) --
 * clear J, begin_step, advance_magnetic, set_inlet, the coupled source step,
 * shift -- for each of `count` independent instances, with the planes of
 * each grid's device inlet reference (aprl_window_set_reference) or the
 * zero reference, enqueued phase by phase across the instances so that
 * their kernels overlap on the device (every field grid owns a stream). The
 * FP32-split form runs its step in three phases (sort and preview; trials,
 * commit and shift; wait); the FP64 form enqueues its first Picard trials
 * for every instance before it waits for any instance's device
 * convergence decision, then commits and shifts each. Errors are the
 * single-instance calls' errors. */
aprl_status aprl_window_step_batch(aprl_instance_step* steps, size_t count);
/* No counterpart in the reference: `steps` consecutive window steps of
 * every instance, each as aprl_window_step_batch, then the driver's end of
 * step done in the library: step + 1, offset - 1, time + dt, and the box
 * origin's z from the queue's inlet (aprl_pending_origin) or, without a
 * queue, fma(offset, dz, reference_z) (Simulation::finishMovingWindowStep,
 * simulation_moving_window.cpp:351); before each step the queue admits its
 * due particles (aprl_pending_admit with look_ahead) and after it the lab
 * plane is sampled (aprl_lab_plane_sample; its ring must hold the samples
 * of the call). An instance's next step is enqueued as soon as its own
 * step is committed, so instances do not wait for each other between
 * steps. state is advanced in place; iterations, samples and entry_steps
 * (store entries summed over the steps) accumulate. */
typedef struct aprl_instance_run {
  aprl_instance_step state;
  double reference_z;
  aprl_pending* queue;          /* optional */
  uint64_t look_ahead;
  struct aprl_lab_plane* plane; /* optional */
  uint64_t iterations, samples, entry_steps;
} aprl_instance_run;
aprl_status aprl_window_run_batch(aprl_instance_run* runs, size_t count, size_t steps);

/* ---------------------------------------------------------------------------
 * FP64 diagnostics (compensated, fixed-order device reductions).
 */

/* Field energy: sum of EBFieldGrid::fieldEnergyDensityCell (eb_field.cpp:
 * 715-738) times dx dy dz over the cells [lo, hi) -- Simulation::
 * localInteriorFieldEnergy (simulation.cpp:3308-3342) with lo = CPML cells,
 * hi = n - CPML cells; MovingWindowFields planeEnergy (moving_window_fields.
 * cpp:387-397) with a one-plane box. */
aprl_status aprl_fields_interior_energy(aprl_fields* fields, const size_t* lo, const size_t* hi, double* energy);
/* Outward power through the faces of the cell box [lo, hi) from
 * poyntingComponentCell (eb_field.cpp:740-782): power[2a] = -sum over the
 * lower face of axis a, power[2a+1] = +sum over the upper face, times the
 * face area -- Simulation::localInteriorBoundaryPower (simulation.cpp:
 * 3344-3432) and MovingWindowFields boundaryPower (moving_window_fields.cpp:
 * 398-419). */
aprl_status aprl_fields_face_power(aprl_fields* fields, const size_t* lo, const size_t* hi, double* power);
/* ChargeConservingCurrentDepositor::maxContinuityResidual (eb_deposition.cpp:
 * 195-226) of the last aprl_coupling_step of this store, before the window
 * shift clears J: out[0] = max over vertices of |(rho1 - rho0)/dt + div J|,
 * out[1] = max |(rho1 - rho0)/dt|, out[2] = max max(|rho0|, |rho1|)/dt
 * (the scale of the residual's round-off: rho and div J are each summed
 * from many taps, so a bunch whose charge barely moves has a residual at
 * round-off of rho/dt, far above round-off of the rate), with rho the
 * charge of the coupling's shape at the step's start and end positions
 * (binomial3-x: the shaped charge of
 * verification/transverse_binomial_checks.cpp:33-40). */
aprl_status aprl_coupling_continuity(aprl_coupling* coupling, aprl_particles* particles, double* out);
/* aprl_coupling_continuity without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); it has no counterpart in the
 * reference. Same kernels, same outputs. Its workspace is one file-scope
 * device array of 24 B per vertex for at most 2,400,000 vertices (57.6 MB,
 * held on every GPU the library runs on until the process ends): grids
 * with more vertices ((nx+1)(ny+1)(nz+1)) are rejected with
 * APRL_INVALID_ARGUMENT. Calls on one GPU share the workspace, so calls
 * that overlap on one GPU race; each call clears what must start at zero. */
aprl_status aprl_coupling_continuity_noalloc(aprl_coupling* coupling, aprl_particles* particles, double* out);
/* Gauss residual of eb_initial_field.cpp:774-799 over the vertices strictly
 * inside [lo, hi] (hi <= (nx, ny, nz)) with the shaped charge of the particles at their current
 * positions (grid origin `origin`): out = {sum error^2, sum source^2,
 * max |error|, max |source|}, error = div E - rho / eps0. */
aprl_status aprl_coupling_gauss(aprl_coupling* coupling, aprl_particles* particles, const double* origin,
                                const size_t* lo, const size_t* hi, double* out);

/* Particle sums of the energy ledger (Simulation, simulation.cpp:3591-3625)
 * over the store's particles (active and held entries: the reference's
 * particles_; carriers and removed entries are not in it) at box time
 * `time`, with the coupling's frame: out[0] sum rest u^2 / (gamma + 1),
 * out[1] sum rest gamma, out[2] represented electrons sum m / m_e, out[3]
 * sum represented * lab gamma, out[4] sum represented * box gamma, out[5]
 * sum represented * lab z (labZFromBoxZT), out[6..8] the lab kinetic energy
 * and c p_z in the box and the lab of ParticleEnergyMoment::sample
 * (particle_energy_audit.h:20-40). FP64 compensated sums in a fixed order
 * (the reference accumulates in long double). */
aprl_status aprl_coupling_particle_energy(aprl_coupling* coupling, aprl_particles* particles, double time,
                                          double* out);

/* The moving-window field energy ledger (MovingWindowFieldEnergyLedger,
 * moving_window_fields.h:39-58; enableEnergyLedger, finishEnergyLedgerStep,
 * energyLedger, moving_window_fields.cpp:207-218, 264-288, 387-438,
 * 443-517), kept on the device. Enabled between steps; from then on every
 * begin_step records the outward power through the faces of the physical
 * aperture, the window commit measures the inlet overwrite's plane energy
 * and inlet power change, finish_energy_step (after the commit, before the
 * shift; APRL_LOGIC_ERROR otherwise) adds the trapezoid face flux
 * dt/2 (start + end - overwrite delta), and every shift adds the removed
 * and inserted plane energies (after aprl_window_complete_inlet on a
 * charge-matched window). aprl_window_energy_ledger: out[0] inlet
 * overwrite energy, out[1] inserted, out[2] cropped, out[3] remap energy,
 * out[4..9] outward energy through x-, x+, y-, y+, z-, z+, out[10..12]
 * overwrite events, flux steps, remap steps. FP64 compensated sums (the
 * reference's are long double). */
aprl_status aprl_window_enable_energy_ledger(aprl_fields* fields, int enable);
aprl_status aprl_window_finish_energy_step(aprl_fields* fields);
aprl_status aprl_window_energy_ledger(aprl_fields* fields, double* out);

/* Laboratory field planes: LabDetectorManager::sampleFieldPlanes
 * (lab_detectors.cpp:2426-2487) for one field plane at fixed lab z, on one
 * rank, with samplePlane (2836-2898) and the detector's clamped trilinear
 * stencil (sampleDetectorComponent / interpolationAxis, 33-97); no
 * external background, as the simulation constructs its manager
 * (simulation.cpp:3291-3293). The plane uses the coupling's boost frame
 * and field grid. aprl_lab_plane_sample is called where the simulation
 * calls sampleFieldPlanes: after the window shift, with the step's end box
 * time and the shifted box origin. When the plane lies on the grid
 * (ownerRank) and its lab-time rhythm is due, it samples every cell centre
 * (i, j) into a device ring of `capacity` samples (*taken = 1): records
 * j-major, per point E_lab[3] then B_lab[3] (FieldPointRecord). A full
 * ring is an APRL_LOGIC_ERROR. aprl_lab_plane_fetch synchronizes the grid's
 * stream, copies the pending samples (count <= capacity of the caller's
 * buffers: time_lab[count], records[count * nx * ny * 6]) and empties the
 * ring. rhythm must be positive (config.cpp:1713-1717). */
typedef struct aprl_lab_plane aprl_lab_plane;
aprl_status aprl_lab_plane_create(aprl_coupling* coupling, double z_lab, double rhythm, size_t capacity,
                                  aprl_lab_plane** out);
void aprl_lab_plane_destroy(aprl_lab_plane* plane);
aprl_status aprl_lab_plane_sample(aprl_lab_plane* plane, double time_box, const double* origin_box, int* taken);
aprl_status aprl_lab_plane_fetch(aprl_lab_plane* plane, size_t capacity, double* time_lab, double* records,
                                 size_t* count);

/* Gauss-consistent initial self-field: EBGaussFieldInitializer::initialize
 * (eb_initial_field.cpp:829-1095) for one rank, with the store's active and
 * held entries as the source (positions in m relative to the grid's box
 * origin `origin`). The CIC charge drives a conjugate-gradient solve of the
 * relativistic-Poisson (7-point, 1 - beta_z^2 along z), electrostatic or
 * Cowan relativistic-Poisson (27-point) operator with zero potential on the
 * static-field guard planes (guard_cells = the CPML cells per axis; with
 * include_transverse_cpml_particles the transverse guard is replaced by a
 * two-cell padded lattice and the fields outside the physical region are
 * cleared); E is the (Cowan-smoothed) potential gradient and B the
 * rigid-translation field at -dt/2, stored into the grid's E and B
 * (replacing them). The post-solve discrete Gauss residual must stay within
 * 4 x relative_tolerance. Errors carry the reference's messages: a charge
 * cloud touching the guard, a mean transverse beta above
 * max_mean_transverse_beta, an indefinite Cowan operator, no convergence
 * within maximum_iterations (all APRL_RUNTIME_ERROR), positions outside
 * the grid (APRL_OUT_OF_RANGE). Defaults of the reference's configuration:
 * relativistic-poisson, tolerance 1e-10, max transverse beta 1e-8, 10000
 * iterations (config.cpp:752-757); the moving window requires the Cowan
 * model (config.cpp:2055-2060). No held-transverse mean velocity. */
typedef enum aprl_initial_model {
  APRL_INITIAL_RELATIVISTIC_POISSON = 0,
  APRL_INITIAL_ELECTROSTATIC_POISSON = 1,
  APRL_INITIAL_COWAN_RELATIVISTIC_POISSON = 2
} aprl_initial_model;
typedef struct aprl_gauss_report {
  uint64_t iterations;
  double relative_residual, maximum_gauss_residual, total_charge, mean_beta_z, mean_beta_transverse,
      initial_electric_energy, initial_magnetic_energy;
  double rms_position_cells[3], centre_padding_rms[3];
  uint64_t static_field_guard_cells[3], transverse_poisson_padding_cells;
} aprl_gauss_report;
aprl_status aprl_fields_initialize_gauss(aprl_fields* fields, aprl_particles* particles, const double* origin,
                                         aprl_initial_model model, double max_mean_transverse_beta,
                                         const size_t* guard_cells, double relative_tolerance,
                                         size_t maximum_iterations, int include_transverse_cpml_particles,
                                         aprl_gauss_report* report);

/* The laboratory-plane input route of a moving-window run: the startup
 * sequence of Simulation for HDF5 version-4 plane-arrival records with
 * frozen injection (simulation.cpp:1983-2245), on a device-resident bunch
 * in laboratory coordinates (the reference's particles_ before the boost,
 * RelativisticParticleSI: position[2] holds the arrival time offset of an
 * event until the projection). One lane per macro-particle; the three sums
 * the reference keeps in long double are double-double sums finished in
 * long double on the host, extrema are exact. Errors carry the reference's
 * types and messages. After aprl_bunch_freeze the bunch is the box-frame
 * snapshot at box time zero with every entry held: download it, hand it to
 * the pending queue (aprl_pending_extract) and upload the rest into the
 * store with aprl_particles_set_held. */
typedef struct aprl_bunch aprl_bunch;
aprl_status aprl_bunch_create(size_t capacity, aprl_bunch** out);
void aprl_bunch_destroy(aprl_bunch* bunch);
size_t aprl_bunch_count(const aprl_bunch* bunch);
/* weight may be NULL (unit weights); held speeds start at zero. */
aprl_status aprl_bunch_upload(aprl_bunch* bunch, size_t n, const double* position3, const double* u3,
                              const double* charge, const double* mass, const double* weight, const uint64_t* id);
/* Any output pointer may be NULL; held_beta = frozenBetaZ after the freeze. */
aprl_status aprl_bunch_download(aprl_bunch* bunch, double* position3, double* u3, double* charge, double* mass,
                                double* weight, uint64_t* id, double* held_beta);
/* Synthetic plane-arrival records (the fields of docs/PARTICLE_INPUT_HDF5.md
 * version 4) of the generated-gaussian beam: the draws of
 * aprl_particles_generate_gaussian with axis 2 read as the arrival time
 * offset, so center[2] and sigma_position[2] are in seconds; proper velocity
 * in the laboratory (boost_gamma must be 1); unit weight. Replaces the
 * bunch's contents. */
aprl_status aprl_bunch_generate_plane_events(aprl_bunch* bunch, const aprl_gaussian_beam* beam, size_t count);
/* SILabPlaneProjectionReport (eb_bunch.h:36-47). */
typedef struct aprl_lab_plane_projection {
  double input_plane_lab, reference_position_lab, minimum_forward_distance, maximum_forward_distance,
      recommended_minimum_reference_position, reference_time_offset_lab, mean_longitudinal_beta;
  uint64_t particles;
} aprl_lab_plane_projection;
/* SIBunchPreprocessor::projectLabPlaneEventsToSnapshot (eb_bunch.h:76-88,
 * eb_bunch.cpp:134-289): every event advanced along its own velocity to the
 * common lab time that puts the weighted longitudinal centroid at
 * reference_position_lab; position[2] becomes the lab z relative to it. */
aprl_status aprl_bunch_project_lab_plane_events(aprl_bunch* bunch, double input_plane_lab,
                                                double reference_position_lab, double longitudinal_offset_lab,
                                                int hold_transverse, aprl_lab_plane_projection* report);
/* SIBunchPlacementReport (eb_bunch.h:23-34). */
typedef struct aprl_bunch_placement {
  double relative_tail_lab, relative_head_lab, reference_position_lab, head_after_lab, actual_head_distance,
      recommended_maximum_reference_position;
  uint64_t particles;
} aprl_bunch_placement;
/* SIBunchPreprocessor::placeRelativeLabSnapshot (eb_bunch.h:90-97,
 * eb_bunch.cpp:291-352) with SIBunchPlacement's fields as arguments. */
aprl_status aprl_bunch_place_relative(aprl_bunch* bunch, int has_entrance, double entrance_lab,
                                      double reference_position_lab, double margin_lab,
                                      aprl_bunch_placement* report);
/* The frame of frozen lab-plane injection (simulation.cpp:2126-2186): lab
 * time origin two z-Courant steps before the earliest activation-plane
 * crossing, lab z origin at the activation plane, box z origin 0. */
aprl_status aprl_bunch_frozen_frame_origin(aprl_bunch* bunch, double gamma_boost, double input_plane_lab,
                                           double activation_plane_lab, double dz, int hold_transverse,
                                           aprl_boost* frame);
/* SIBunchBoostReport (eb_bunch.h:49-60). */
typedef struct aprl_bunch_boost_report {
  double earliest_lab_event_time, latest_lab_event_time, maximum_absolute_drift_time, minimum_lab_gamma,
      maximum_lab_gamma, maximum_relative_momentum_round_trip_error, maximum_relative_gamma_round_trip_error;
} aprl_bunch_boost_report;
/* SIBunchPreprocessor::boostLabSnapshotToBoxTimeZero (eb_bunch.h:99-108,
 * eb_bunch.cpp:354-482). */
aprl_status aprl_bunch_boost_to_box_time_zero(aprl_bunch* bunch, const aprl_boost* frame, double snapshot_time_lab,
                                              int hold_transverse, double input_plane_lab,
                                              aprl_bunch_boost_report* report);
/* The frozen flags of frozen injection (simulation.cpp:2228-2245): the held
 * speed frozenBetaZ = heldLongitudinalBetaZ(u) with the transverse hold (0
 * without), and APRL_RUNTIME_ERROR when an entry is at or past the
 * activation plane at box time zero. */
aprl_status aprl_bunch_freeze(aprl_bunch* bunch, const aprl_boost* frame, double activation_plane_lab,
                              int hold_transverse);

/* The moving window's inlet reference (BlockedCowanReference,
 * blocked_cowan_reference.h:68-129) prepared on the device: the source
 * particles' charge on the interior vertices of the global reference grid
 * (CIC, source_shape = -1, or the production longitudinal basis with
 * linear or x-binomial3 transverse weights: APRL_SHAPE_QUINTIC_COMPENSATED,
 * APRL_SHAPE_LINEAR_BINOMIAL3), the 27-point Cowan operator with the
 * translation term at the mean held beta solved exactly by sine modes in
 * x, y and a Thomas sweep per mode along z, and the native E(0), B(-dt/2)
 * of the whole grid kept in device memory. The charge is summed in 64-bit
 * fixed point (order-free), the transforms are direct sums: fields agree
 * with the reference's FFTW-based construction to round-off, not bitwise.
 * The open-wall transverse reference and the block files are not built.
 * Errors carry the reference's types and messages. */
typedef struct aprl_inlet_reference aprl_inlet_reference;
typedef struct aprl_inlet_reference_report {
  double minimum_mode_eigenvalue, minimum_pivot, source_charge, deposited_charge;
  uint64_t interior_nodes[3];
} aprl_inlet_reference_report;
aprl_status aprl_inlet_reference_create(const aprl_grid_geometry* global, const double* global_origin, size_t n,
                                        const double* position3, const double* charge, double beta,
                                        const size_t* guards, int source_shape, int transverse_x_binomial3,
                                        aprl_inlet_reference** out, aprl_inlet_reference_report* report);
void aprl_inlet_reference_destroy(aprl_inlet_reference* reference);
/* BlockedCowanReference::field for a whole native component (0..5 = Ex Ey Ez
 * Bx By Bz, x fastest). */
aprl_status aprl_inlet_reference_download(aprl_inlet_reference* reference, int component, double* host,
                                          size_t count);
/* Simulation::movingWindowFieldValue (simulation_moving_window.cpp:280-294)
 * over component `component`'s x-y nodes at fixed-world plane k and step:
 * the reference at z = k - translation_beta step, linear between planes;
 * APRL_OUT_OF_RANGE outside the prepared planes. */
aprl_status aprl_inlet_reference_plane(aprl_inlet_reference* reference, int component, ptrdiff_t k, size_t step,
                                       double translation_beta, double* host, size_t count);
/* The window's reference callback on the device: with a reference set, the
 * plane arguments of aprl_window_begin_step, aprl_window_begin_step_centered,
 * aprl_window_set_inlet and aprl_window_shift that are NULL are filled from
 * it (movingWindowFieldValue at the window's offset and step, at
 * translation_beta) by a kernel instead of taking the zero reference. The
 * reference's x-y grid and spacings must be the window's. NULL restores the
 * zero reference. The reference must outlive its use by the window. */
aprl_status aprl_window_set_reference(aprl_fields* fields, aprl_inlet_reference* reference, double translation_beta);

/* Launch-shape autotuner (no counterpart in the reference): launch shapes
 * that give bitwise the same results (the Cowan sweep's z-run length, the E
 * update's workgroup size) are timed per kernel and grid shape on their
 * first uses and the fastest is kept; with the environment variable
 * APRL_TUNE_CACHE naming a file, decisions are read from it and appended
 * to it. enable = 0 launches the default shapes. The report lists the
 * decisions taken, one "key value" line each (*length = its size; the
 * buffer receives at most capacity - 1 characters and a terminating 0). */
aprl_status aprl_autotune(int enable);
aprl_status aprl_autotune_report(char* buffer, size_t capacity, size_t* length);

/* Per-kernel timing with HIP events on the instance stream. enable=1 starts
 * recording every launch; query returns launches and total milliseconds for
 * a kernel name (0 launches if unknown). */
aprl_status aprl_profile_enable(aprl_fields* fields, int enable);
aprl_status aprl_profile_query(aprl_fields* fields, const char* kernel, uint64_t* launches, double* total_ms);
aprl_status aprl_profile_reset(aprl_fields* fields);

#ifdef __cplusplus
}
#endif

#endif /* APRL_GPU_H */
