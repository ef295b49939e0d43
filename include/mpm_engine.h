/*
 * mpm_engine.h -- C ABI of the MI355X MLS-MPM step engine.
 *
 * Drop-in boundary for the step API of Summer-Particle-Lab's `MPM3` class
 * (src/three/mpm.js:7-136). Plain C types only: opaque handles, host
 * pointers and sizes. The Node.js N-API addon (summer-particle-lab_amd/node)
 * and the Python ctypes wrapper (summer-particle-lab_amd/mpm_amd) both bind
 * exactly these symbols.
 *
 * Conventions
 *  - Every call returns an int status: MPM_OK (0) or an MPM_ERR_* code;
 *    mpm_last_error() gives the message of the last failure on a handle.
 *  - One engine owns one HIP stream on the current device; a handle is not
 *    thread-safe (callers serialise), distinct handles are independent.
 *  - Host arrays are caller-owned and read/written only inside the call.
 *  - Particle tensors F, C, cached_tau are 9 values per particle, row-major,
 *    particle-major: the layout of MPM3.F / MPM3.C / MPM3.cachedTau
 *    (src/three/mpm.js:9,12).
 *  - Particle volume is the physical particle volume V_p. The reference stores
 *    V_p/64 in `s.volume` and multiplies by 64 in P2G (src/three/mpm.js:68);
 *    a reference-unit caller passes 64*s.volume.
 *  - Device state is FP32 structure-of-arrays; FP64 appears on the device only
 *    in the conservation reductions (mpm_conservation).
 */
#ifndef MPM_ENGINE_H
#define MPM_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 7 and 8 mirror the reference GPU service's status codes for
 * "MPM CFL limit exceeded" and "GPU state became invalid"
 * (gpu/mixed_solver.py:381-396); the state guard of mpm_config sets them
 * (mpm_step returns 7 at once for a dt no state could allow; the rest
 * surfaces at mpm_synchronize / mpm_batch_synchronize). */
enum {
  MPM_OK = 0,
  MPM_ERR_INVALID_ARGUMENT = 1,
  MPM_ERR_HIP = 2,
  MPM_ERR_CAPACITY = 3,
  MPM_ERR_UNSUPPORTED = 4,
  MPM_ERR_NOT_READY = 5,
  MPM_ERR_CFL = 7,
  MPM_ERR_INVALID_STATE = 8
};

/* Grid wall rule per domain face, applied to nodes within one cell of the
 * face. SEPARATE is the reference rule (zero the outward normal component
 * only, src/three/mpm.js:93); STICKY zeroes all components; SLIP zeroes the
 * normal component in both directions. */
enum { MPM_WALL_SEPARATE = 0, MPM_WALL_STICKY = 1, MPM_WALL_SLIP = 2 };

/* Array element type of a particle view. */
enum { MPM_F64 = 0, MPM_F32 = 1 };

/* Constitutive models, numbered as the reference's MODELS table
 * (src/three/constitutive.js:6-19): 0 fixed-corotated elastic,
 * 1 Neo-Hookean, 2 snow, 3 sand (Drucker-Prager), 4 ideal J2,
 * 5 linear-hardening J2, 6 Johnson-Cook, 7 simplified Johnson-Cook,
 * 8 cohesive Drucker-Prager soil, 9 Deshpande-Fleck crushable foam,
 * 10 thermal J2, 11 Norton-Arrhenius creep. */
enum {
  MPM_MODEL_FIXED_COROTATED = 0,
  MPM_MODEL_NEO_HOOKEAN = 1,
  MPM_MODEL_SNOW = 2,
  MPM_MODEL_SAND = 3,
  MPM_MODEL_J2 = 4,
  MPM_MODEL_J2_HARDENING = 5,
  MPM_MODEL_JOHNSON_COOK = 6,
  MPM_MODEL_JOHNSON_COOK_SIMPLE = 7,
  MPM_MODEL_COHESIVE_DP = 8,
  MPM_MODEL_FOAM = 9,
  MPM_MODEL_THERMAL_J2 = 10,
  MPM_MODEL_CREEP = 11
};

typedef struct mpm_engine mpm_engine;

/* Engine geometry and global options; replaces MPM3.configure(domain)
 * (src/three/mpm.js:18) and the domain rules of domainParameters
 * (src/three/discretization.js:4-15): origin = floor(lower/h)*h - 2h,
 * nodes = ceil((upper - origin)/h) + 3 per axis, grid edge = max over axes. */
typedef struct {
  double lower[3], upper[3]; /* domain box */
  double cell_size;          /* h */
  uint32_t capacity;         /* particle capacity */
  int32_t wall_mode[6];      /* faces -x,+x,-y,+y,-z,+z: MPM_WALL_* */
  double gravity[3];         /* constant body acceleration */
  int32_t gravity_on_particles; /* 1: add g*dt to particle v before P2G (the
                                   reference caller's order,
                                   src/three/simulation.js:337-344); 0: add
                                   g*dt to grid velocity after p/m */
  double mass_eps;           /* grid nodes with mass <= eps keep raw momentum
                                (reference: 1e-12) */
  int32_t sort_every;        /* re-bin particles every k steps (>= 1) */
  int32_t strict_domain;     /* 1: apply the reference's domain limits */
  /* State guard: the checks of the reference GPU service's check_state
   * (gpu/mixed_solver.py:381-388), made inside every step's particle kernel.
   * CFL (status 7): the stable step of the state a step starts from,
   * min(cfl_cap, cfl_number h / max(wave, 1), grad_number / max(sum |C|, 1)) with
   * wave = |v| + c max(1, |F|row |F|col, compression) (src/three/mpm.js:43-55;
   * c = sqrt((lambda + 2 mu) hardening / rho)), is below 0.99999 dt.
   * Invalid state (status 8): after G2P, det F outside [1e-6, 1e6] or not a
   * number, a plastic strain that is not a number, a velocity component not
   * finite or above 1e5 in magnitude, or a position not a number or more than
   * one cell outside [lower, upper] (the walls act on nodes within one cell
   * of a face, so a particle can settle up to half a cell past it). */
  int32_t guard_off;         /* 1: no checks */
  double cfl_number;         /* <= 0: the reference's 1.4/8 */
  double cfl_cap;            /* > 0: the stable step's cap term (the reference's
                                h/8/240 in its own units); <= 0: no cap */
  double grad_number;        /* bound of dt sum |C| (the deformation of one
                                step); <= 0: the reference's .08 */
} mpm_config;

/* One material. Strengths are normalised by E as in the reference records
 * (src/three/thermal-plasticity.js:1-4, src/three/metal-plasticity.js:1-4).
 * mu, lambda are absolute Lame parameters; the reference uses
 * mu = lambda = 0.4 E (src/three/mpm.js:37). */
typedef struct {
  int32_t model; /* MPM_MODEL_* */
  float E, mu, lambda;
  /* thermal J2 / creep (models 10, 11) */
  float th_Tref, th_Thot, th_hotRatio, th_A, th_H, th_rate, th_Q, th_n, th_stress, th_kelvin_offset;
  /* J2 / Johnson-Cook (models 4-7) */
  float pl_A, pl_B, pl_n, pl_C, pl_rate0, pl_Tref, pl_Tm, pl_m;
  /* snow (2): stretch yield strain; sand (3): friction angle in degrees.
   * 0 selects the reference defaults .025 and 35; other values are clamped
   * to the reference's ranges [.005, .08] and [15, 50]
   * (src/three/simulation.js:40). */
  float yield_strain, friction;
  /* cohesive Drucker-Prager (8) and crushable foam (9): Cauchy strengths / E,
   * tension positive (src/three/pressure-plasticity.js:2-3). Checked against
   * the reference's ranges and constraints when the model is 8 or 9;
   * mpm_material_defaults fills the reference defaults. */
  float pr_cohesion, pr_slope, pr_dilatancy, pr_tension, pr_alpha, pr_plateau, pr_gamma, pr_densify, pr_beta,
      pr_epsD;
  /* tensile-strain damage (src/three/fracture.js:6-18): when fracture != 0,
   * damage = (max principal log stretch - onset) / (failure - onset), never
   * decreasing; the step stress carries (1 - damage), and at damage 1 the
   * particle steps on as sand (model 3) with F = I. Requires
   * .005 <= onset < failure <= 2. */
  int32_t fracture;
  float fr_onset, fr_failure;
} mpm_material;

/* Host view of particle state (structure of arrays). On upload every pointer
 * except the optional ones must be non-NULL; on download NULL pointers are
 * skipped. dtype selects FP64 or FP32 for all floating arrays. */
typedef struct {
  uint32_t count;
  int32_t dtype;
  void *x, *y, *z, *vx, *vy, *vz; /* [count] */
  void *F, *C;                    /* [count*9] */
  void *mass, *volume, *temperature; /* [count] */
  void *plastic;                  /* accumulated plastic strain ep [count] */
  void *cached_tau;               /* optional [count*9]: stress cached by the
                                     last plastic return (MPM3.cachedTau) */
  uint8_t *material;              /* index into the material table [count] */
  uint8_t *pinned;                /* 1 = pinned (MPM3 skips G2P) [count] */
  uint8_t *cache_valid;           /* optional [count] (MPM3.cacheValid) */
  void *jp;                       /* optional [count]: snow's plastic volume
                                     ratio MPM3.Jp (1 when NULL) */
  void *dilation;                 /* optional [count]: sand's MPM3.dilation
                                     (0 when NULL) */
  void *damage;                   /* optional [count]: tensile damage in
                                     [0, 1] (the scene's s.damage; 0 when
                                     NULL). A particle at 1 is failed. */
} mpm_particles;

/* FP64 conservation report. Grid totals are those of the most recent P2G. */
typedef struct {
  double particle_mass;      /* sum m_p */
  double grid_mass;          /* sum over nodes (fixed-point total / scale) */
  int64_t particle_mass_q;   /* sum of per-particle fixed-point masses */
  int64_t grid_mass_q;       /* sum of node fixed-point masses; equals
                                particle_mass_q exactly after every P2G */
  double mass_scale;         /* fixed-point units per unit mass (2^k) */
  double momentum[3];        /* sum m_p v_p */
  double grid_momentum[3];   /* sum of node momenta after P2G */
  double angular[3];         /* sum x_p x m_p v_p + m_p D spin(C_p), D = h^2/4
                                (src/simulation/mechanical-diagnostics.js:19-31) */
  double kinetic;            /* sum 1/2 m |v|^2 */
  double affine;             /* sum 1/2 m D |C|^2 */
  double min_J;              /* min det F */
  double max_speed;          /* max |v_p| */
  int64_t nonfinite;         /* particles with a non-finite x, v or F */
} mpm_conservation;

/* Per-kernel timing accumulated by HIP events while profiling is on. */
typedef struct {
  char name[32];
  int64_t launches;
  double total_ms;
} mpm_kernel_time;

/* Create / destroy; replaces `new MPM3(capacity)` + configure(domain). */
int mpm_create(const mpm_config* config, mpm_engine** out);
int mpm_destroy(mpm_engine* e);
const char* mpm_last_error(const mpm_engine* e);

/* Material table (up to 64 entries); replaces the per-part records the
 * reference passes to projectPlastic (src/three/mpm.js:127). Every entry
 * must pass mpm_material_check. */
int mpm_set_materials(mpm_engine* e, const mpm_material* materials, int32_t count);

/* A material of the given model with mu = lambda = 0.4 E and every other
 * field at the reference's defaults: thermal record
 * (src/three/thermal-plasticity.js:3), Johnson-Cook record
 * (src/three/metal-plasticity.js:3), pressure record
 * (src/three/pressure-plasticity.js:2), yield strain .025, friction 35,
 * fracture off with onset .08, failure .16 (src/three/fracture.js:7). Host
 * only. */
int mpm_material_defaults(int32_t model, float E, mpm_material* out);

/* MPM_OK when the record is usable; MPM_ERR_UNSUPPORTED for a model number
 * outside 0-11; else MPM_ERR_INVALID_ARGUMENT (non-finite or non-positive E,
 * negative mu, non-finite lambda; for models 8 and 9
 * the pressure record outside the ranges of pressureParameters or violating
 * dilatancy <= slope and slope * tension <= cohesion,
 * src/three/pressure-plasticity.js:3-10; with fracture on, onset and failure
 * outside .005 <= onset < failure <= 2, src/three/fracture.js:8). Host only. */
int mpm_material_check(const mpm_material* m);

/* Upload particle state (replaces reading s.x.. and MPM3's own arrays at the
 * start of MPM3.step). Resets the fixed-point mass scale and re-bins. */
int mpm_upload(mpm_engine* e, const mpm_particles* p);

/* Download particle state in the caller's original particle order. */
int mpm_download(mpm_engine* e, mpm_particles* p);

/* Advance nsteps steps of size dt; replaces MPM3.step(s, dt)
 * (src/three/mpm.js:56-134). Asynchronous on the engine stream. With the
 * guard on, a dt above cfl_cap, cfl_number h or grad_number (the stable step of any
 * state) gives MPM_ERR_CFL at once and enqueues nothing; the state-dependent
 * checks are made on the device and reported by mpm_synchronize. */
int mpm_step(mpm_engine* e, double dt, int32_t nsteps);

/* CFL-type stable step of the current state; replaces MPM3.stableStep(s)
 * (src/three/mpm.js:43-55). cap <= 0 uses h/8/240, cfl <= 0 uses 1.4/8 (the
 * reference constants in reference units). */
int mpm_stable_dt(mpm_engine* e, double cap, double cfl, double* out);

/* FP64 conservation reductions of the current state (see mpm_conservation);
 * min_J replaces MPM3.minJ(s) (src/three/mpm.js:135). Synchronising. */
int mpm_conservation_report(mpm_engine* e, mpm_conservation* out);

/* ---- split step: the grid between its update and G2P ----
 *
 * The reference lets its callers act on the grid inside MPM3.step: the
 * externalGridContact hook changes node velocities after the grid update
 * (src/three/mpm.js:104) and contactGridDeltas adds a per-part velocity to
 * every node a particle of that part gathers from (:114). A split step is
 * mpm_step_to_grid, any number of grid reads and velocity writes, then
 * mpm_step_from_grid with the same dt. mpm_step, mpm_advance_frame and a
 * second mpm_step_to_grid refuse (MPM_ERR_NOT_READY) while one is pending;
 * mpm_upload drops it. */

/* Stress, P2G and the grid update of one step (src/three/mpm.js:61-102). */
int mpm_step_to_grid(mpm_engine* e, double dt);
/* The grid as the reference's MPM3 holds it: mass [N^3], v [3 N^3]
 * (node-major x, y, z) and fixed [N^3] (1 where a pinned particle holds the
 * node), node index (z N + y) N + x; nodes no particle reached hold 0. Any
 * pointer may be NULL. */
int mpm_grid_download(mpm_engine* e, double* mass, double* v, uint8_t* fixed);
/* Writes node velocities v [3 N^3] back for the nodes particles reached. */
int mpm_grid_upload_velocity(mpm_engine* e, const double* v);
/* G2P (plus per-part deltas), F update, advection and the plastic return of
 * the pending step (src/three/mpm.js:107-131). deltas holds nslots arrays of
 * [3 N^3] velocities; material_slot [nmat] gives each material's slot or -1.
 * nslots = 0 means no deltas. */
int mpm_step_from_grid(mpm_engine* e, double dt, const int32_t* material_slot, int32_t nmat, const double* deltas,
                       int32_t nslots);

/* ---- frame snapshot and wire packet (the GPU service protocol) ----
 *
 * The reference's GPU service sends one binary packet per frame
 * (gpu/server.py:17-20): a u32 little-endian header length, the header (a
 * compact UTF-8 JSON object), spaces up to a multiple of 4 bytes, then 8
 * little-endian float32 per previewed particle: x, y, z, material code, T,
 * pinned, stress, plastic (gpu/solver.py:986-1032). The browser worker
 * checks that the sizes agree (src/three/gpu-worker.js:36-42). */

/* Totals of a frame (the header fields of gpu/solver.py:1024-1030). All
 * particles of the engine are MPM solids, so liquids is 0. */
typedef struct {
  uint32_t count;          /* particles */
  uint32_t render_count;   /* rows: ceil(count / sample_stride) */
  uint32_t sample_stride;  /* max(1, ceil(count / preview_limit)) */
  int64_t solids, liquids;
  int64_t models[12];      /* particles per effective model (a failed particle counts as 3) */
  double mass;             /* sum m */
  double min_temperature, max_temperature; /* 0 and 0 for an empty engine */
  double min_J;            /* min(1, min det F) */
  double max_stress;       /* max of the von Mises diagnostic, from 0 */
  double max_plastic;      /* max plastic strain, from 0 */
  int64_t nonfinite;       /* particles with |x|^2 or |v|^2 not below 1e20 */
} mpm_snapshot_stats;

/* Stride and row count of a preview of `count` particles limited to
 * `preview_limit` rows (the service uses 100000). Host only. */
int mpm_preview_layout(uint32_t count, uint32_t preview_limit, uint32_t* sample_stride, uint32_t* render_count);

/* The frame snapshot of the resident state: one row per block of
 * sample_stride consecutive particle ids, holding the particle the service's
 * hash picks in that block (gpu/solver.py:990-996); data is [render_count *
 * 8] host floats, or NULL for the totals only. codes [ncodes] gives each
 * material's code (11, the MPM solid, shows as 11 + model); NULL gives 11
 * for every material. Replaces Solver.snapshot and Simulation3.snapshot's
 * packing (src/three/simulation.js:390-400). */
int mpm_snapshot(mpm_engine* e, uint32_t preview_limit, const int32_t* codes, int32_t ncodes, float* data,
                 mpm_snapshot_stats* stats);

/* Builds one frame packet into out [capacity]: *length receives the packet
 * size (4 + json_len + padding + rows * 32). With out NULL only *length is
 * set; a capacity below it gives MPM_ERR_CAPACITY. Host only. */
int mpm_frame_packet(const char* json, uint32_t json_len, const float* data, uint64_t rows, uint8_t* out,
                     uint64_t capacity, uint64_t* length);

/* Per-particle von Mises diagnostic (Cauchy von Mises / E), in the caller's
 * particle order; replaces MPM3.refreshDiagnostics (src/three/mpm.js:40-42).
 * stress is [count] FP64. */
int mpm_refresh_diagnostics(mpm_engine* e, double* stress);

/* The Kirchhoff stress the next step applies to each particle, in the
 * caller's particle order, 9 values row-major per particle (MPM3.tau,
 * src/three/mpm.js:33-37): the cached return stress of models >= 2 when
 * valid, else the model's stress of F, times (1 - damage) below failure. */
int mpm_stress_tensors(mpm_engine* e, double* tau);

/* Wait for all work of the engine, then report and clear the guard's
 * findings since the last call: MPM_ERR_INVALID_STATE (8) when some step
 * left an invalid state, else MPM_ERR_CFL (7) when some step started from a
 * state whose stable step is below its dt (8 wins, as in check_state). */
int mpm_synchronize(mpm_engine* e);

/* Number of particles resident on the device. */
int mpm_count(const mpm_engine* e, uint32_t* out);

/* ---- frame wrapper ----
 * One display frame of Simulation3.step for an all-solid scene
 * (src/three/simulation.js:305-315) with its substeps (:330-388): while more
 * than 1e-9 of dt remains and fewer than max_substeps were taken, a substep
 * of h = min(remaining, stable step * time_step_scale) adds the gravity of
 * the selected mode to unpinned particles (src/three/gravity-options.js:7-16),
 * runs the MPM step, clamps unpinned positions to the bounds (pinned ones
 * stay), rebuilds v = clamp((x - x_start) / h, +-velocity_limit) and, with
 * anti_clump, applies regularizeSolids (src/three/solid-sampling.js:5-42) to
 * particles of equal material index (the drop-in keys materials by part).
 * Not covered: fluids, drag, the press, heat exchange. The engine's own
 * gravity should be zero (the wrapper applies gravity itself). */
enum { MPM_GRAVITY_NONE = 0, MPM_GRAVITY_UNIFORM = 1, MPM_GRAVITY_ATTRACT = 2, MPM_GRAVITY_REPEL = 3 };

typedef struct {
  double time_step_scale;      /* settings.timeStepScale, in [0.1, 1] */
  int32_t max_substeps;        /* 48 in the reference */
  double velocity_limit;       /* 160 in the reference */
  double bounds_lo[3], bounds_hi[3]; /* s.boundsLo, s.boundsHi */
  int32_t anti_clump;          /* settings.solidAntiClump */
  double volume_unit;          /* particle volume in the spacing rule's unit per
                                  engine unit (1/64: the reference's s.volume) */
  int32_t gravity_mode;        /* MPM_GRAVITY_* */
  double gravity_strength;     /* settings.gravityStrength */
  double gravity_direction[3]; /* the unit vector of settings.gravityDirection */
  double gravity_center[3];    /* settings.gravityCenter */
  double gravity_softening;    /* settings.gravitySoftening */
} mpm_frame_config;

typedef struct {
  int32_t substeps;
  double advanced;             /* dt minus what remained */
} mpm_frame_report;

/* Fills c with the reference defaults: scale 1, 48 substeps, limit 160,
 * anti-clump on, volume unit 1/64, uniform gravity down, strength 1,
 * softening 8; bounds from the engine's domain. */
int mpm_frame_defaults(const mpm_engine* e, mpm_frame_config* c);
int mpm_advance_frame(mpm_engine* e, const mpm_frame_config* c, double dt, mpm_frame_report* out);

/* ---- measurement and parity hooks (no reference counterpart) ---- */

/* Turn per-kernel HIP-event timing on (1) or off (0); reset clears totals. */
int mpm_profile(mpm_engine* e, int32_t enable, int32_t reset);
/* Copy up to max kernel timing records; *n receives the record count. */
int mpm_kernel_times(mpm_engine* e, mpm_kernel_time* out, int32_t max, int32_t* n);

/* Bin the resident particles as a re-bin does and return the cell key of
 * every particle indexed by its original id (keys_in), the sorted keys
 * (keys_out) and the original ids in sorted order (perm); nothing moves.
 * Any pointer may be NULL. A key is the blocked index of the base cell
 * b = clamp(floor((x - O)/h - 1/2), 0, N - 3) of the FP32 position
 * (src/three/mpm.js:28): cells by 4x4x4 block (blocks in z, y, x order),
 * x fastest inside a block; the stable sort breaks ties by id. */
int mpm_debug_bin(mpm_engine* e, uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm);

/* Stable LSD radix sort of arbitrary 32-bit keys on the device; values are
 * the input positions. Used to test the binning sort on edge cases. */
int mpm_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t key_bits, uint32_t* keys_out, uint32_t* perm_out);

/* mpm_sort_pairs on device memory without any allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops). keys [n], keys_out [n] and
 * perm_out [n] are device pointers (keys may alias keys_out); the sort is
 * enqueued on `stream` (a hipStream_t; NULL is the null stream) and the
 * caller synchronises. Scratch (second buffers, histograms, scan partials)
 * is one static device array per GPU sized for MPM_SORT_NOALLOC_MAX keys;
 * larger n gives MPM_ERR_CAPACITY. Calls that overlap on one GPU share that
 * array and race. Same result as mpm_sort_pairs. */
#define MPM_SORT_NOALLOC_MAX (1u << 20)
int mpm_sort_pairs_noalloc(const uint32_t* keys, uint32_t n, uint32_t key_bits, uint32_t* keys_out,
                           uint32_t* perm_out, void* stream);

/* One P2G transfer of the resident state with step dt, leaving the dense
 * grid: mass [N^3] and momentum [3*N^3] as FP64 copies (node-major,
 * x fastest). Does not advance the state. */
int mpm_debug_p2g(mpm_engine* e, double dt, double* grid_mass, double* grid_momentum, int32_t* N);

/* The engine's plain constitutive update applied on the device to a batch of
 * n deformation gradients F [9n] with plastic strain ep [n] and temperature
 * T [n]: for models >= 2 the plastic return of the step (SVD, radial return,
 * recomposed F, ep += increment, Hencky stress of the returned state); for
 * model 0 F is unchanged and tau is the fixed-corotated stress. tau_out is
 * [6n] (xx, yy, zz, xy, xz, yz). Replaces the reference's per-particle
 * projectPlastic call (src/three/mpm.js:127, src/three/constitutive.js:52-78),
 * and is the label generator of the learned constitutive surrogate. */
int mpm_return_map_batch(const mpm_material* m, uint32_t n, float dt, const float* F, const float* ep,
                         const float* T, float* F_out, float* ep_out, float* tau_out);

/* Step stress of a batch of F without a cached return (updateStress,
 * src/three/mpm.js:33-39): fixed-corotated for model 0, Hencky from F's own
 * decomposition for plastic models. tau_out [6n]. */
int mpm_stress_batch(const mpm_material* m, uint32_t n, const float* F, float* tau_out);

/* mpm_return_map_batch with the history variables in and out: jp [n] (snow's
 * Jp; the pressure models' plastic volume ratio) and dilation [n] (sand).
 * jp_out and dilation_out may alias jp and dilation. */
int mpm_return_map_state_batch(const mpm_material* m, uint32_t n, float dt, const float* F, const float* ep,
                               const float* T, const float* jp, const float* dilation, float* F_out, float* ep_out,
                               float* jp_out, float* dilation_out, float* tau_out);

/* ---- learned constitutive surrogate (no reference counterpart) ----
 *
 * A 7 -> 64 -> 64 -> 64 -> 6 network with hard-swish z min(max(z/6 + 1/2, 0), 1)
 * between layers (MPM_MLP_NPARAM FP32 parameters,
 * packed W1[64x7] b1 W2[64x64] b2 W3[64x64] b3 W4[6x64] b4, weights
 * output-major) maps the trial elastic logarithmic strain eps (xx, yy, zz,
 * xy, xz, yz) and the temperature T to the deviatoric part of the
 * return-mapped Kirchhoff stress, both in the units of the yield point at T:
 * inputs are eps sqrt(6) (mu/E) / y(T) * in_scale and (T - T_center) *
 * T_scale, outputs are in units of out_scale * y(T) * E, with y(T) the
 * E-normalised yield stress (a perfectly plastic J2 return is then the same
 * map at every temperature). It replaces
 * the plastic return of projectPlastic (src/three/constitutive.js:52-78) for
 * rate-independent J2 models without hardening (model 4; model 10 with
 * H = 0), where the return depends on eps and T only. */
#define MPM_MLP_NPARAM 9222

typedef struct {
  uint32_t samples;    /* training samples drawn per run */
  uint32_t batch;      /* minibatch size */
  uint32_t epochs;
  uint32_t seed;
  float lr, lr_final;  /* Adam step, cosine-annealed from lr to lr_final */
  float strain_sigma;  /* eps components ~ N(0, strain_sigma) */
  float T_lo, T_hi;    /* T ~ U(T_lo, T_hi) */
  float dt;            /* step passed to the return map */
  float eps_max;       /* envelope half-width of every eps component; <= 0: 4 sigma */
  float near_yield_fraction; /* share of samples drawn near the yield surface:
                                |dev eps| = 1 + 1.5 u^2 (u uniform) times the
                                yield strain at T, random direction, volumetric
                                part N(0, sigma); the rest is N(0, sigma).
                                Draws inside the yield surface get their
                                deviatoric part stretched to 1 + 1.5 u times
                                the yield strain: all samples are plastic */
} mpm_mlp_train_config;

typedef struct {
  float in_scale, T_center, T_scale;
  float out_scale;     /* output unit in units of y(T) E; the deviatoric
                          stress is output * out_scale * y(T) * E */
  float eps_max;       /* envelope: |eps_ij| <= eps_max */
  float T_lo, T_hi;    /* envelope: T_lo <= T <= T_hi */
  float min_yield_ratio; /* envelope: the network takes trial states with
                            q >= min_yield_ratio * yield (1: every state past
                            yield); the ones between yield and that take
                            the plain return */
  int32_t material;    /* set by mpm_set_surrogate */
  /* the training material, in E-normalised form: a network applies to any
   * material with the same model, mu/E, lambda/E, yield ratio and thermal
   * softening (the return map scales with E) */
  int32_t model;
  float mu_over_E, lambda_over_E, A, Tref, Thot, hotRatio;
} mpm_mlp_norm;

/* Normalisation for material m and a training config: in_scale 1/2, T
 * mapped to [-1, 1], out_scale 1, min_yield_ratio 1, and m's E-normalised
 * fingerprint. */
int mpm_mlp_default_norm(const mpm_material* m, const mpm_mlp_train_config* c, mpm_mlp_norm* out);

/* MPM_OK when material m matches norm's training fingerprint (E-normalised
 * parameters within a relative 1e-5), else MPM_ERR_INVALID_ARGUMENT. Host
 * only. */
int mpm_mlp_check_material(const mpm_material* m, const mpm_mlp_norm* norm);

/* n training samples generated on the device by the plain return map
 * (mpm_return_map_batch's update with F = exp(eps), ep = 0): X [7n] inputs,
 * Y [6n] targets, sample-major. */
int mpm_mlp_samples(const mpm_material* m, const mpm_mlp_train_config* c, const mpm_mlp_norm* norm, uint32_t n,
                    uint32_t seed, float* X, float* Y);

/* Train the network on the device (FP32 forward/backward kernels, Adam) on
 * c->samples samples from mpm_mlp_samples' generator. params receives the
 * MPM_MLP_NPARAM trained parameters; loss_history (may be NULL) the mean
 * squared error of every epoch [c->epochs]. */
int mpm_mlp_train(const mpm_material* m, const mpm_mlp_train_config* c, const mpm_mlp_norm* norm, float* params,
                  float* loss_history);

/* Mean squared error (summed over the 6 outputs, averaged over the n
 * samples) of the network on X [7n], Y [6n] and its gradient with respect to
 * every parameter, by the training kernels. */
int mpm_mlp_loss_grad(const float* params, uint32_t n, const float* X, const float* Y, float* grad, float* loss);

/* Network outputs for X [7n] by the BF16 MFMA inference of the step kernel
 * (BF16 weights and activations, FP32 accumulation). Y is [6n]. */
int mpm_mlp_infer(const float* params, uint32_t n, const float* X, float* Y);

/* The step's surrogate return on a batch of n trial states F [9n] with ep
 * [n] and T [n] of material m: closed-form elastic branch inside the yield
 * surface, network inside the envelope, plain return (as
 * mpm_return_map_batch) otherwise. Outputs as mpm_return_map_batch, plus
 * branch [n]: 0 network, 1 elastic, 2 plain return. m must match norm's
 * fingerprint. */
int mpm_surrogate_return_batch(const mpm_material* m, const float* params, const mpm_mlp_norm* norm, uint32_t n,
                               float dt, const float* F, const float* ep, const float* T, float* F_out, float* ep_out,
                               float* tau_out, int32_t* branch);

/* Use the network for particles of material index `material` whose trial
 * strain and temperature lie inside norm's envelope; the others keep the
 * plain return. The material must match norm's fingerprint (relative
 * tolerance 1e-5), else MPM_ERR_INVALID_ARGUMENT. params == NULL switches
 * the surrogate off, and so does mpm_set_materials. */
int mpm_set_surrogate(mpm_engine* e, int32_t material, const float* params, const mpm_mlp_norm* norm);

/* Particle-steps of the surrogate's material since the last reset that took
 * the network (used), the closed-form elastic branch (elastic) and the plain
 * return outside the envelope (fallback). reset != 0 clears the counters. */
int mpm_surrogate_counts(mpm_engine* e, uint64_t* used, uint64_t* elastic, uint64_t* fallback, int32_t reset);

/* ---- batched scenes (no reference counterpart) ----
 *
 * Many independent scenes advanced together: one particle pool and one
 * concatenated grid on one GPU, one particle launch per size bucket (a
 * workgroup per chunk of 64-256 particles; a single launch over all buckets
 * when they share a chunk size) and one grid launch per step for all scenes,
 * and optionally the whole re-bin period replayed as a captured HIP graph. The reference runs
 * one scene per Worker or Python process (gpu/server.py:34-40); these calls
 * batch what would be many MPM3 instances. Scenes share one material table;
 * each has its own domain (mpm_config; capacity and sort_every are ignored),
 * particle count (fixed at creation) and step size. */
typedef struct mpm_batch mpm_batch;

typedef struct {
  int32_t sort_every;      /* re-bin period in steps (>= 1; default 16) */
  int32_t use_graph;       /* 1: replay whole re-bin periods as HIP graphs */
  int32_t particle_blocks; /* ignored (kept for ABI compatibility): the
                              particle launch runs one workgroup per chunk */
} mpm_batch_options;

/* Launch configuration: scenes are bucketed by grid edge (nodes per axis
 * <= bucket_edge[0], <= bucket_edge[1], larger); each bucket has its own
 * particles per workgroup (64, 128 or 256). */
typedef struct {
  int32_t buckets;
  int32_t bucket_edge[4];
  int32_t bucket_scenes[4];
  int32_t particles_per_group[4];
  int32_t sort_every;
  int32_t particle_blocks;  /* reported: particle workgroups per step (one per
                               chunk); ignored by mpm_batch_set_tuning */
  double ms_per_step;   /* from mpm_batch_tune: whole-step time of the choice */
  double bucket_ms[4];  /* from mpm_batch_tune: particle launch time per step */
} mpm_batch_tuning;

int mpm_batch_create(const mpm_config* scenes, const uint32_t* counts, int32_t nscenes,
                     const mpm_batch_options* options, mpm_batch** out);
int mpm_batch_destroy(mpm_batch* b);
const char* mpm_batch_last_error(const mpm_batch* b);
int mpm_batch_scene_count(const mpm_batch* b, int32_t* scenes, uint32_t* particles);
int mpm_batch_set_materials(mpm_batch* b, const mpm_material* materials, int32_t count);
/* Upload / download one scene; p->count must equal the scene's count. */
int mpm_batch_upload(mpm_batch* b, int32_t scene, const mpm_particles* p);
int mpm_batch_download(mpm_batch* b, int32_t scene, mpm_particles* p);
/* Step size of every scene, dt[nscenes]. */
int mpm_batch_set_dt(mpm_batch* b, const double* dt);
/* Advance every scene nsteps steps (asynchronous on the batch stream). */
int mpm_batch_step(mpm_batch* b, int32_t nsteps);
/* Waits for the batch; reports and clears the guard's findings of every
 * scene as mpm_synchronize does (the message names the first scene). */
int mpm_batch_synchronize(mpm_batch* b);
/* FP64 conservation report of one scene (as mpm_conservation_report). */
int mpm_batch_conservation(mpm_batch* b, int32_t scene, mpm_conservation* out);
/* The batch's binning, as mpm_debug_bin for the whole pool (total_n
 * entries, ids and slots of scene s in [p0_s, p0_s + n_s)): keys are the
 * scene-local blocked cell keys of each scene's grid, and the sort runs per
 * scene, so each scene's ids stay in its range. A test hook with no
 * counterpart in the reference. */
int mpm_batch_debug_bin(mpm_batch* b, uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm);
int mpm_batch_get_tuning(const mpm_batch* b, mpm_batch_tuning* out);
int mpm_batch_set_tuning(mpm_batch* b, const mpm_batch_tuning* t);
/* Autotuner: times particles per workgroup {64, 128, 256} per bucket on
 * trial_steps steps each, then the re-bin period {4, 8, 16, 32}, on the
 * resident scenes (which it advances), and keeps the fastest. */
int mpm_batch_tune(mpm_batch* b, int32_t trial_steps, mpm_batch_tuning* out);

/* ---- 2D elastic MLS-MPM (the reference's SolidMPM) ----
 *
 * The reference's 2D solver for its MPM solid material (src/solid-mpm.js):
 * quadratic MLS-MPM/APIC with the 2D fixed-corotated law (mu = lambda =
 * 0.4 E, E = m / (12.96 V) 16000 stiffness) on a fixed grid of 149 x 149
 * nodes at cell 8 from -592 (src/solid-mpm.js:4). Particle state stays in
 * the caller's order. */
typedef struct mpm2_engine mpm2_engine;

/* Host view of 2D particles (structure of arrays, FP64 or FP32 as dtype). F
 * and C hold 4 values per particle: 00, 01, 10, 11 (SolidMPM's f00..c11). On
 * download NULL pointers are skipped. */
typedef struct {
  uint32_t count;
  int32_t dtype;
  void *x, *y, *vx, *vy;   /* [count] */
  void *F, *C;             /* [4 count] */
  void *mass, *volume;     /* [count] (the reference's s.mass, s.volume) */
  uint8_t* pinned;         /* [count] */
} mpm2_particles;

/* The settings the 2D step reads: settings.solidStiffness, and the wall box
 * that applies when settings.boundary is on or the scene is the softbody
 * demo or the basin: |x| < bound, floor < y < top (src/solid-mpm.js:69-80). */
typedef struct {
  float stiffness;
  int32_t walls;
  float bound, floor, top;
} mpm2_settings;

/* SolidMPM.stats (src/solid-mpm.js:106-115) plus the exact mass totals. */
typedef struct {
  uint32_t count;
  double min_J, max_J;     /* 1 and 1 without particles */
  double strain;           /* largest Green strain norm */
  uint32_t nodes;          /* grid nodes with mass in the last step */
  uint32_t substeps;       /* substeps of the last step call or frame */
  int64_t grid_mass_q;     /* fixed-point node mass of the last step */
  int64_t particle_mass_q; /* fixed-point particle total; equal to grid_mass_q */
} mpm2_stats;

/* new SolidMPM(capacity). */
int mpm2_create(uint32_t capacity, mpm2_engine** out);
int mpm2_destroy(mpm2_engine* e);
const char* mpm2_last_error(const mpm2_engine* e);
/* Settings of a scene: "softbody", "basin" or any other name (NULL) for the
 * 555 box; stiffness 1 and walls on (src/simulation.js:26). Host only. */
int mpm2_settings_defaults(const char* scene, mpm2_settings* out);
int mpm2_upload(mpm2_engine* e, const mpm2_particles* p);
int mpm2_download(mpm2_engine* e, mpm2_particles* p);
/* nsteps SolidMPM.step(s, dt) calls (src/solid-mpm.js:48-92), asynchronous. */
int mpm2_step(mpm2_engine* e, const mpm2_settings* s, double dt, int32_t nsteps);
/* SolidMPM.stableStep(s) (src/solid-mpm.js:38-47). */
int mpm2_stable_dt(mpm2_engine* e, const mpm2_settings* s, double* out);
/* One frame of the 2D driver for a scene of MPM solids only
 * (src/simulation.js:240-300): dt clamped to [0, .05]; substeps of
 * min(remaining, stableStep) while time remains, at most max_substeps (the
 * reference's 96); each adds gravity (0 none, 1 uniform, 2 centre) times 27
 * strength to unpinned velocities, zeroes pinned ones, clamps to +-180 and
 * steps. The loop runs on the device. */
int mpm2_advance(mpm2_engine* e, const mpm2_settings* s, double dt, int32_t gravity_mode, float strength,
                 int32_t max_substeps, int32_t* substeps, double* advanced);
int mpm2_statistics(mpm2_engine* e, mpm2_stats* out);
int mpm2_synchronize(mpm2_engine* e);

#ifdef __cplusplus
}
#endif

#endif /* MPM_ENGINE_H */
