// C ABI of the MLS-MPM engine (include/mpm_engine.h): engine object, device
// buffers, upload/download, the step driver and the measurement hooks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mpm_engine.h"
#include "mpm_host.h"
#include "mpm_mlp_device.h"
#include "mpm_internal.h"

using namespace mpm;

struct PendingTiming {
  std::string name;
  hipEvent_t start, stop;
};

struct mpm_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  mpm_config cfg{};
  GridDims g{};
  int lo_idx[3], hi_idx[3];
  DevState d{};
  uint32_t n = 0;
  std::vector<mpm_material> materials;
  double mass_scale = 1.0;
  int64_t particle_mass_q = 0;
  int steps_since_sort = 0;
  std::string err;
  double* red_host = nullptr;  // pinned
  uint32_t* status_host = nullptr;  // pinned copy of the guard's status word
  uint32_t grid_blocks = 1;    // workgroups of the grid update
  // frame wrapper scratch (allocated on first use)
  float* frame_ox = nullptr;
  float* frame_disp = nullptr;
  uint32_t* frame_cell_start = nullptr;
  unsigned int* frame_max = nullptr;
  // split step (grid hooks): the grid of a step waits for mpm_step_from_grid
  bool split_pending = false;
  double split_dt = 0;
  // profiling
  bool profiling = false;
  std::vector<PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<int64_t, double>> times;
  std::vector<std::string> order;
};

namespace {

int fail(mpm_engine* e, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  return code;
}

#define HIP_OK(e, call)                                                                         \
  do {                                                                                          \
    hipError_t _st = (call);                                                                    \
    if (_st != hipSuccess) return fail((e), MPM_ERR_HIP, "%s: %s", #call, hipGetErrorString(_st)); \
  } while (0)

hipEvent_t take_event(mpm_engine* e) {
  if (!e->event_pool.empty()) {
    hipEvent_t ev = e->event_pool.back();
    e->event_pool.pop_back();
    return ev;
  }
  hipEvent_t ev;
  hipEventCreate(&ev);
  return ev;
}

// Times one launch when profiling is on. Launchers that take LaunchEvents
// have the events stamped by the kernel dispatch itself (stamped = true);
// otherwise the events are recorded on the stream around the launches.
struct Timed {
  mpm_engine* e;
  const char* name;
  bool stamped;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(mpm_engine* e_, const char* n, bool stamped_ = false) : e(e_), name(n), stamped(stamped_) {
    if (e->profiling) {
      a = take_event(e);
      b = take_event(e);
      if (!stamped) hipEventRecord(a, e->stream);
    }
  }
  LaunchEvents ev() const { return LaunchEvents{a, b}; }
  ~Timed() {
    if (e->profiling) {
      if (!stamped) hipEventRecord(b, e->stream);
      e->pending.push_back({name, a, b});
    }
  }
};

void collect_times(mpm_engine* e) {
  for (auto& p : e->pending) {
    hipEventSynchronize(p.stop);
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.start, p.stop);
    auto it = e->times.find(p.name);
    if (it == e->times.end()) {
      e->times[p.name] = {1, ms};
      e->order.push_back(p.name);
    } else {
      it->second.first += 1;
      it->second.second += ms;
    }
    e->event_pool.push_back(p.start);
    e->event_pool.push_back(p.stop);
  }
  e->pending.clear();
}

uint32_t key_bits(int N) {
  const uint64_t cells = blocked_key_count(N);
  uint32_t b = 1;
  while (b < 32 && (1ull << b) < cells) ++b;
  return b;
}

}  // namespace

namespace mpm {
DevMaterial to_dev_material(const mpm_material& m) {
  DevMaterial d{};
  d.model = m.model;
  d.E = m.E;
  d.mu = m.mu;
  d.lambda = m.lambda;
  d.G = m.E > 0 ? m.mu / m.E : 0.f;
  d.lr = m.mu != 0 ? m.lambda / m.mu : 0.f;
  d.th_Tref = m.th_Tref;
  d.th_Thot = m.th_Thot;
  d.th_hotRatio = m.th_hotRatio;
  d.th_A = m.th_A;
  d.th_H = m.th_H;
  d.th_rate = m.th_rate;
  d.th_Q = m.th_Q;
  d.th_n = m.th_n;
  d.th_stress = m.th_stress;
  d.th_kelvin = m.th_kelvin_offset;
  // the correctly rounded reciprocal of the span (0 where there is none)
  const float span = m.th_Thot - m.th_Tref, inv_span = 1.f / span;
  d.th_inv_span = (span > 0.f && std::isfinite(inv_span) && inv_span >= 1e-30f) ? inv_span : 0.f;
  d.pl_A = m.pl_A;
  d.pl_B = m.pl_B;
  d.pl_n = m.pl_n;
  d.pl_C = m.pl_C;
  d.pl_rate0 = m.pl_rate0;
  d.pl_Tref = m.pl_Tref;
  d.pl_Tm = m.pl_Tm;
  d.pl_m = m.pl_m;
  // Snow and sand parameters with the reference's defaults and ranges
  // (src/three/simulation.js:36-40).
  const float ys = m.yield_strain > 0 ? std::min(std::max(m.yield_strain, .005f), .08f) : .025f;
  const float fr = m.friction > 0 ? std::min(std::max(m.friction, 15.f), 50.f) : 35.f;
  const double sn = std::sin(fr * M_PI / 180.0);
  d.sn_ys = ys;
  d.sd_alpha = (float)(std::sqrt(2.0 / 3.0) * 2 * sn / (3 - sn));
  d.pr_c = m.pr_cohesion;
  d.pr_slope = m.pr_slope;
  d.pr_dil = m.pr_dilatancy;
  d.pr_T = m.pr_tension;
  d.pr_alpha = m.pr_alpha;
  d.pr_plateau = m.pr_plateau;
  d.pr_gamma = m.pr_gamma;
  d.pr_densify = m.pr_densify;
  d.pr_beta = m.pr_beta;
  d.pr_epsD = m.pr_epsD;
  d.fr_on = m.fracture != 0;
  d.fr_onset = m.fr_onset;
  d.fr_inv = m.fracture ? (float)(1.0 / ((double)m.fr_failure - m.fr_onset)) : 0.f;
  return d;
}
}  // namespace mpm

extern "C" int mpm_material_defaults(int32_t model, float E, mpm_material* out) {
  if (!out) return MPM_ERR_INVALID_ARGUMENT;
  mpm_material m{};
  m.model = model;
  m.E = E;
  m.mu = m.lambda = .4f * E;
  // src/three/thermal-plasticity.js:3
  m.th_Tref = 20.f;
  m.th_Thot = 180.f;
  m.th_hotRatio = .08f;
  m.th_A = .04f;
  m.th_H = .04f;
  m.th_rate = .0005f;
  m.th_Q = 35000.f;
  m.th_n = 3.f;
  m.th_stress = .04f;
  m.th_kelvin_offset = 273.15f;
  // src/three/metal-plasticity.js:3
  m.pl_A = .025f;
  m.pl_B = .08f;
  m.pl_n = .35f;
  m.pl_C = .025f;
  m.pl_rate0 = 1.f;
  m.pl_Tref = 20.f;
  m.pl_Tm = 1000.f;
  m.pl_m = 1.f;
  m.yield_strain = .025f;
  m.friction = 35.f;
  // src/three/pressure-plasticity.js:2
  m.pr_cohesion = .025f;
  m.pr_slope = .6f;
  m.pr_dilatancy = .15f;
  m.pr_tension = .01f;
  m.pr_alpha = 1.5f;
  m.pr_plateau = .015f;
  m.pr_gamma = .01f;
  m.pr_densify = .02f;
  m.pr_beta = 2.f;
  m.pr_epsD = .7f;
  // src/three/fracture.js:7
  m.fracture = 0;
  m.fr_onset = .08f;
  m.fr_failure = .16f;
  *out = m;
  return MPM_OK;
}

extern "C" int mpm_material_check(const mpm_material* m) {
  if (!m) return MPM_ERR_INVALID_ARGUMENT;
  if (m->model < 0 || m->model > 11) return MPM_ERR_UNSUPPORTED;
  if (!(m->E > 0) || !std::isfinite(m->E) || !(m->mu >= 0) || !std::isfinite(m->mu) || !std::isfinite(m->lambda))
    return MPM_ERR_INVALID_ARGUMENT;
  if (m->model == 8 || m->model == 9) {
    // Ranges and constraints of pressureParameters (src/three/pressure-plasticity.js:3-10),
    // with a relative 1e-6 allowance for the FP32 fields.
    const struct {
      float v;
      double lo, hi;
    } r[] = {{m->pr_cohesion, 0, .5},       {m->pr_slope, 0, 2},           {m->pr_dilatancy, 0, 2},
             {m->pr_tension, 0, .5},        {m->pr_alpha, .05, std::sqrt(4.5)}, {m->pr_plateau, .0001, .5},
             {m->pr_gamma, 0, 1},           {m->pr_densify, .0001, 1},     {m->pr_beta, 1, 5},
             {m->pr_epsD, .05, 3}};
    for (const auto& x : r)
      if (!std::isfinite(x.v) || x.v < x.lo * (1 - 1e-6) || x.v > x.hi * (1 + 1e-6)) return MPM_ERR_INVALID_ARGUMENT;
    if (!((double)m->pr_dilatancy <= (double)m->pr_slope * (1 + 1e-6)) ||
        !((double)m->pr_slope * m->pr_tension <= (double)m->pr_cohesion + 1e-7))
      return MPM_ERR_INVALID_ARGUMENT;
  }
  if (m->fracture) {  // src/three/fracture.js:8
    if (!std::isfinite(m->fr_onset) || !std::isfinite(m->fr_failure) || m->fr_onset < .005f ||
        !(m->fr_failure > m->fr_onset) || m->fr_failure > 2.f)
      return MPM_ERR_INVALID_ARGUMENT;
  }
  return MPM_OK;
}

namespace {

bool supported_model(int m) { return m >= 0 && m <= 11; }

// Bin and sort the resident particles into cell order: d.vals then holds,
// for each new slot, the slot the particle lives in now.
int sort_keys(mpm_engine* e) {
  {
    Timed t(e, "bin_keys");
    launch_bin_keys(e->stream, e->d, e->n, e->g);
  }
  {
    Timed t(e, "radix_sort");
    radix_sort_pairs(e->stream, e->d.keys, e->d.vals, e->d.keys_alt, e->d.vals_alt, e->d.hist, e->d.partials, e->n,
                     key_bits(e->g.N));
  }
  HIP_OK(e, hipGetLastError());
  e->steps_since_sort = 0;
  return MPM_OK;
}

// Bin, sort and permute the resident particles into cell order (a re-bin at
// the start of a call: every plane moves). A re-bin between two fused steps
// sorts only, and the fused step after it re-gathers the particles.
int rebin(mpm_engine* e) {
  if (e->n == 0) return MPM_OK;
  int rc = sort_keys(e);
  if (rc) return rc;
  {
    Timed t(e, "permute");
    launch_permute(e->stream, e->d, e->n, e->d.vals);
  }
  HIP_OK(e, hipGetLastError());
  return MPM_OK;
}

// Step parameters of a step of size dt; MPM_ERR_CFL in *rc when the guard
// finds that no state can take dt.
StepParams params(const mpm_engine* e, double dt, int* rc = nullptr) {
  StepParams p;
  p.dt = (float)dt;
  p.mass_scale = (float)e->mass_scale;
  p.inv_mass_scale = (float)(1.0 / e->mass_scale);
  p.returns = e->d.returns ? 1u : 0u;
  const int r = guard_params(e->cfg, dt, p);
  if (rc) *rc = r;
  return p;
}

// Runs a device kernel over host arrays of n items: copies the inputs in,
// launches, copies the outputs back, frees everything.
template <typename Launch>
int batch_call(uint32_t n, std::vector<std::pair<const float*, size_t>> ins,
                      std::vector<std::pair<float*, size_t>> outs, Launch launch) {
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MPM_ERR_HIP;
  std::vector<float*> din(ins.size(), nullptr), dout(outs.size(), nullptr);
  hipError_t st = hipSuccess;
  for (size_t k = 0; k < ins.size() && st == hipSuccess; ++k) {
    st = hipMalloc((void**)&din[k], sizeof(float) * std::max<size_t>(ins[k].second * n, 1));
    if (st == hipSuccess && n)
      st = hipMemcpyAsync(din[k], ins[k].first, sizeof(float) * ins[k].second * n, hipMemcpyHostToDevice, s);
  }
  for (size_t k = 0; k < outs.size() && st == hipSuccess; ++k)
    st = hipMalloc((void**)&dout[k], sizeof(float) * std::max<size_t>(outs[k].second * n, 1));
  if (st == hipSuccess) {
    launch(s, din, dout);
    st = hipGetLastError();
  }
  for (size_t k = 0; k < outs.size() && st == hipSuccess; ++k)
    if (n) st = hipMemcpyAsync(outs[k].first, dout[k], sizeof(float) * outs[k].second * n, hipMemcpyDeviceToHost, s);
  if (st == hipSuccess) st = hipStreamSynchronize(s);
  for (float* p : din)
    if (p) (void)hipFree(p);
  for (float* p : dout)
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(s);
  return st == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // namespace

extern "C" {

const char* mpm_last_error(const mpm_engine* e) { return e ? e->err.c_str() : "null engine"; }

int mpm_create(const mpm_config* c, mpm_engine** out) {
  if (!c || !out) return MPM_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  GridDims g;
  int lo_idx[3], hi_idx[3];
  const int rc = grid_geometry(c, g, lo_idx, hi_idx);
  if (rc) return rc;
  const uint64_t Nn = (uint64_t)g.N * g.N * g.N;
  auto* e = new mpm_engine();
  e->cfg = *c;
  if (e->cfg.sort_every < 1) e->cfg.sort_every = 1;
  e->g = g;
  for (int a = 0; a < 3; ++a) {
    e->lo_idx[a] = lo_idx[a];
    e->hi_idx[a] = hi_idx[a];
  }

  hipError_t st = hipGetDevice(&e->device);
  if (st == hipSuccess) st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  const uint32_t cap = std::max<uint32_t>(c->capacity, 1);
  DevState& d = e->d;
  d.cap = cap;
  auto alloc = [&](void** p, size_t bytes) {
    if (st == hipSuccess) st = hipMalloc(p, std::max<size_t>(bytes, 256));
  };
  alloc((void**)&d.planes, sizeof(float) * NPLANES * (size_t)cap);
  alloc((void**)&d.planes_alt, sizeof(float) * NPLANES * (size_t)cap);
  alloc((void**)&d.ids, sizeof(uint32_t) * cap);
  alloc((void**)&d.ids_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.meta, sizeof(uint32_t) * cap);
  alloc((void**)&d.meta_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.acc_m, sizeof(unsigned long long) * Nn);
  alloc((void**)&d.acc_p, sizeof(float) * 3 * Nn);
  alloc((void**)&d.vel, sizeof(float4) * Nn);
  alloc((void**)&d.fixed, Nn);
  alloc((void**)&d.pieces, piece_flag_bytes(pieces_of_nodes(Nn)));
  alloc((void**)&d.keys, sizeof(uint32_t) * cap);
  alloc((void**)&d.keys_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.vals, sizeof(uint32_t) * cap);
  alloc((void**)&d.vals_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.hist, sizeof(uint32_t) * radix_hist_words(cap));
  alloc((void**)&d.partials, sizeof(uint32_t) * radix_partial_words(cap));
  const uint32_t npieces = grid_pieces(g);
  alloc((void**)&d.red, sizeof(double) * 4 * SUM_SLOTS);
  alloc((void**)&d.piece_list, sizeof(uint32_t) * npieces);
  alloc((void**)&d.piece_counters, sizeof(unsigned long long) * 3);
  alloc((void**)&d.materials, sizeof(DevMaterial) * 64);
  alloc((void**)&d.status, sizeof(uint32_t));
  if (st == hipSuccess)
    st = hipHostMalloc((void**)&e->red_host, sizeof(double) * 4 * SUM_SLOTS, hipHostMallocDefault);
  if (st == hipSuccess) st = hipHostMalloc((void**)&e->status_host, sizeof(uint32_t), hipHostMallocDefault);
  if (st == hipSuccess) st = hipMemsetAsync(d.status, 0, sizeof(uint32_t), e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.acc_m, 0, sizeof(unsigned long long) * Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.acc_p, 0, sizeof(float) * 3 * Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.vel, 0, sizeof(float4) * Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.fixed, 0, Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.pieces, 0, piece_flag_bytes(pieces_of_nodes(Nn)), e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.red, 0, sizeof(double) * 4 * SUM_SLOTS, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.piece_counters, 0, sizeof(unsigned long long) * 3, e->stream);
  int cus = 256;
  if (st == hipSuccess) st = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device);
  // Grid update workgroups: one per CU below 2^17 pieces (up to about 200^3
  // nodes), rising to four per CU from 2^18 pieces (a 256^3-cell grid has
  // 265k). Workgroups past the step's list exit at once but each still costs
  // a dispatch, which a small grid's short list does not repay.
  const int per_cu = std::max(1, std::min(4, (int)(((size_t)npieces * PIECE_NODES) >> 22)));
  e->grid_blocks = (uint32_t)std::max(1, std::min<int>(per_cu * cus, (int)((npieces + 15) / 16)));
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) {
    mpm_destroy(e);
    return MPM_ERR_HIP;
  }
  // Default material: fixed-corotated, E = 1, mu = lambda = 0.4.
  mpm_material m0{};
  m0.model = 0;
  m0.E = 1.f;
  m0.mu = m0.lambda = .4f;
  mpm_set_materials(e, &m0, 1);
  *out = e;
  return MPM_OK;
}

int mpm_destroy(mpm_engine* e) {
  if (!e) return MPM_OK;
  if (e->stream) hipStreamSynchronize(e->stream);
  DevState& d = e->d;
  void* ptrs[] = {d.planes, d.planes_alt, d.ids,  d.ids_alt,  d.meta,     d.meta_alt, d.acc_m,
                  d.acc_p,  d.vel,        d.fixed, d.pieces,  d.keys,     d.keys_alt, d.vals,
                  d.vals_alt, d.hist,     d.partials, d.red,  d.materials, d.mlp, d.sur_counts,
                  d.piece_list, d.piece_counters, d.status, e->frame_ox, e->frame_disp, e->frame_cell_start,
                  e->frame_max};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (e->red_host) hipHostFree(e->red_host);
  if (e->status_host) hipHostFree(e->status_host);
  for (auto& p : e->pending) {
    hipEventDestroy(p.start);
    hipEventDestroy(p.stop);
  }
  for (auto ev : e->event_pool) hipEventDestroy(ev);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return MPM_OK;
}

int mpm_set_materials(mpm_engine* e, const mpm_material* mats, int32_t count) {
  if (!e || !mats || count < 1 || count > 64) return fail(e, MPM_ERR_INVALID_ARGUMENT, "1..64 materials required");
  std::vector<DevMaterial> dm(count);
  for (int i = 0; i < count; ++i) {
    const int rc = mpm_material_check(&mats[i]);
    if (rc == MPM_ERR_UNSUPPORTED)
      return fail(e, rc, "material %d: model %d is not supported", i, mats[i].model);
    if (rc != MPM_OK)
      return fail(e, rc, "material %d: E, mu, lambda or the model's parameters are out of range", i);
    dm[i] = to_dev_material(mats[i]);
  }
  e->materials.assign(mats, mats + count);
  e->d.ext_laws = false;
  e->d.history = false;
  e->d.returns = false;
  for (int i = 0; i < count; ++i) {
    e->d.ext_laws |= mats[i].model == 8 || mats[i].model == 9 || mats[i].fracture != 0;
    e->d.history |= keeps_history(mats[i]);
    e->d.returns |= returns_in_g2p(mats[i]);
  }
  if (e->d.mlp) {  // a new table switches the surrogate off
    HIP_OK(e, hipStreamSynchronize(e->stream));
    HIP_OK(e, hipFree(e->d.mlp));
    e->d.mlp = nullptr;
  }
  HIP_OK(e, hipMemcpyAsync(e->d.materials, dm.data(), sizeof(DevMaterial) * count, hipMemcpyHostToDevice,
                           e->stream));
  HIP_OK(e, hipStreamSynchronize(e->stream));
  return MPM_OK;
}

int mpm_upload(mpm_engine* e, const mpm_particles* p) {
  if (e) e->split_pending = false;  // a new state drops a pending split step
  if (!e || !p) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = p->count;
  if (n > e->d.cap) return fail(e, MPM_ERR_CAPACITY, "count %u exceeds capacity %u", n, e->d.cap);
  if (!particles_complete(p)) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad dtype or missing particle array");
  const uint32_t cap = e->d.cap;
  std::vector<float> planes((size_t)NPLANES * cap, 0.f);
  std::vector<uint32_t> ids(n), meta(n);
  const std::string why = pack_particles(p, e->materials.size(), planes.data(), cap, 0, 0, ids.data(), meta.data());
  if (!why.empty()) return fail(e, MPM_ERR_INVALID_ARGUMENT, "%s", why.c_str());
  double mmax = 0;
  for (uint32_t i = 0; i < n; ++i) mmax = std::max(mmax, (double)planes[(size_t)P_MASS * cap + i]);
  e->mass_scale = std::ldexp(1.0, mass_exponent(mmax));
  int64_t q = 0;
  const float sf = (float)e->mass_scale;
  for (uint32_t i = 0; i < n; ++i) q += (int64_t)std::nearbyint(planes[(size_t)P_MASS * cap + i] * sf);
  e->particle_mass_q = q;
  e->n = n;
  e->d.narrow = narrow_chunks(n, occupied_cells(planes.data(), cap, n, e->g));
  HIP_OK(e, hipMemcpyAsync(e->d.planes, planes.data(), sizeof(float) * planes.size(), hipMemcpyHostToDevice,
                           e->stream));
  if (n) {
    HIP_OK(e, hipMemcpyAsync(e->d.ids, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, e->stream));
    HIP_OK(e, hipMemcpyAsync(e->d.meta, meta.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, e->stream));
  }
  int rc = rebin(e);
  if (rc) return rc;
  HIP_OK(e, hipStreamSynchronize(e->stream));
  return MPM_OK;
}

int mpm_download(mpm_engine* e, mpm_particles* p) {
  if (!e || !p) return MPM_ERR_INVALID_ARGUMENT;
  if (p->count != e->n) return fail(e, MPM_ERR_INVALID_ARGUMENT, "count %u != resident %u", p->count, e->n);
  const uint32_t n = e->n, cap = e->d.cap;
  if (p->dtype != MPM_F64 && p->dtype != MPM_F32) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad dtype");
  std::vector<float> planes((size_t)NPLANES * cap);
  std::vector<uint32_t> ids(n), meta(n);
  HIP_OK(e, hipMemcpyAsync(planes.data(), e->d.planes, sizeof(float) * planes.size(), hipMemcpyDeviceToHost,
                           e->stream));
  if (n) {
    HIP_OK(e, hipMemcpyAsync(ids.data(), e->d.ids, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(e, hipMemcpyAsync(meta.data(), e->d.meta, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, e->stream));
  }
  HIP_OK(e, hipStreamSynchronize(e->stream));
  unpack_particles(planes.data(), cap, 0, ids.data(), meta.data(), n, 0, p);
  return MPM_OK;
}

int mpm_step(mpm_engine* e, double dt, int32_t nsteps) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  if (!(dt > 0) || !std::isfinite(dt) || nsteps < 0) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad dt or nsteps");
  if (e->split_pending) return fail(e, MPM_ERR_NOT_READY, "a split step waits for mpm_step_from_grid");
  if (nsteps == 0 || e->n == 0) return MPM_OK;
  int guard = MPM_OK;
  const StepParams sp = params(e, dt, &guard);
  if (guard) return fail(e, guard, "MPM CFL limit exceeded: dt %g is above the stable step of any state", dt);
  if (e->steps_since_sort >= e->cfg.sort_every) {
    int rc = rebin(e);
    if (rc) return rc;
  }
  {
    Timed t(e, "p2g", true);
    launch_fused(e->stream, e->d, e->n, e->g, sp, 0, t.ev());
  }
  for (int s = 0; s < nsteps; ++s) {
    {
      Timed t(e, "pieces", true);
      launch_compact(e->stream, e->d, grid_pieces(e->g), 1, t.ev());
    }
    {
      Timed t(e, "grid", true);
      launch_grid(e->stream, e->d, e->g, sp, e->lo_idx, e->hi_idx, e->grid_blocks, t.ev());
    }
    e->steps_since_sort++;
    if (s == nsteps - 1) {
      Timed t(e, "g2p", true);
      launch_fused(e->stream, e->d, e->n, e->g, sp, 2, t.ev());
    } else {
      // A re-bin due here sorts the positions this grid was built from,
      // which the next G2P gathers at; that fused step reads the particles
      // through the new order and writes them in it.
      if (e->steps_since_sort >= e->cfg.sort_every) {
        int rc = sort_keys(e);
        if (rc) return rc;
        Timed t(e, "g2p2g_rebin", true);
        launch_fused(e->stream, e->d, e->n, e->g, sp, 3, t.ev());
        swap_particle_buffers(e->d);
      } else {
        Timed t(e, "g2p2g", true);
        launch_fused(e->stream, e->d, e->n, e->g, sp, 1, t.ev());
      }
    }
  }
  HIP_OK(e, hipGetLastError());
  return MPM_OK;
}

int mpm_step_to_grid(mpm_engine* e, double dt) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  if (!(dt > 0) || !std::isfinite(dt)) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad dt");
  if (e->split_pending) return fail(e, MPM_ERR_NOT_READY, "a split step waits for mpm_step_from_grid");
  if (e->n == 0) return MPM_OK;
  int guard = MPM_OK;
  const StepParams sp = params(e, dt, &guard);
  if (guard) return fail(e, guard, "MPM CFL limit exceeded: dt %g is above the stable step of any state", dt);
  if (e->steps_since_sort >= e->cfg.sort_every) {
    int rc = rebin(e);
    if (rc) return rc;
  }
  {
    Timed t(e, "p2g", true);
    launch_fused(e->stream, e->d, e->n, e->g, sp, 0, t.ev());
  }
  {
    Timed t(e, "pieces", true);
    launch_compact(e->stream, e->d, grid_pieces(e->g), 1, t.ev());
  }
  {
    Timed t(e, "grid", true);
    launch_grid(e->stream, e->d, e->g, sp, e->lo_idx, e->hi_idx, e->grid_blocks, t.ev());
  }
  HIP_OK(e, hipGetLastError());
  e->split_pending = true;
  e->split_dt = dt;
  return MPM_OK;
}

int mpm_grid_download(mpm_engine* e, double* mass, double* v, uint8_t* fixed) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  if (!e->split_pending) return fail(e, MPM_ERR_NOT_READY, "no grid: call mpm_step_to_grid first");
  const size_t Nn = (size_t)e->g.N * e->g.N * e->g.N;
  float *dm = nullptr, *dv = nullptr;
  uint8_t* df = nullptr;
  hipError_t st = hipMallocAsync((void**)&dm, sizeof(float) * Nn, e->stream);
  if (st == hipSuccess) st = hipMallocAsync((void**)&dv, sizeof(float) * 3 * Nn, e->stream);
  if (st == hipSuccess) st = hipMallocAsync((void**)&df, Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(dm, 0, sizeof(float) * Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(dv, 0, sizeof(float) * 3 * Nn, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(df, 0, Nn, e->stream);
  if (st == hipSuccess) {
    launch_grid_export(e->stream, e->d, e->g, dm, dv, df);
    st = hipGetLastError();
  }
  std::vector<float> hm(mass ? Nn : 0), hv(v ? 3 * Nn : 0);
  if (st == hipSuccess && mass) st = hipMemcpyAsync(hm.data(), dm, sizeof(float) * Nn, hipMemcpyDeviceToHost, e->stream);
  if (st == hipSuccess && v) st = hipMemcpyAsync(hv.data(), dv, sizeof(float) * 3 * Nn, hipMemcpyDeviceToHost, e->stream);
  if (st == hipSuccess && fixed) st = hipMemcpyAsync(fixed, df, Nn, hipMemcpyDeviceToHost, e->stream);
  for (void* p : {(void*)dm, (void*)dv, (void*)df})
    if (p) (void)hipFreeAsync(p, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "grid download: %s", hipGetErrorString(st));
  for (size_t k = 0; k < hm.size(); ++k) mass[k] = hm[k];
  for (size_t k = 0; k < hv.size(); ++k) v[k] = hv[k];
  return MPM_OK;
}

int mpm_grid_upload_velocity(mpm_engine* e, const double* v) {
  if (!e || !v) return MPM_ERR_INVALID_ARGUMENT;
  if (!e->split_pending) return fail(e, MPM_ERR_NOT_READY, "no grid: call mpm_step_to_grid first");
  const size_t Nn = (size_t)e->g.N * e->g.N * e->g.N;
  std::vector<float> hv(3 * Nn);
  for (size_t k = 0; k < hv.size(); ++k) hv[k] = (float)v[k];
  float* dv = nullptr;
  hipError_t st = hipMallocAsync((void**)&dv, sizeof(float) * 3 * Nn, e->stream);
  if (st == hipSuccess) st = hipMemcpyAsync(dv, hv.data(), sizeof(float) * 3 * Nn, hipMemcpyHostToDevice, e->stream);
  if (st == hipSuccess) {
    launch_grid_import(e->stream, e->d, e->g, dv);
    st = hipGetLastError();
  }
  if (dv) (void)hipFreeAsync(dv, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);  // hv must outlive the copy
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "grid upload: %s", hipGetErrorString(st));
  return MPM_OK;
}

int mpm_step_from_grid(mpm_engine* e, double dt, const int32_t* material_slot, int32_t nmat, const double* deltas,
                       int32_t nslots) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  if (e->n == 0) return MPM_OK;
  if (!e->split_pending) return fail(e, MPM_ERR_NOT_READY, "no grid: call mpm_step_to_grid first");
  if (dt != e->split_dt) return fail(e, MPM_ERR_INVALID_ARGUMENT, "dt differs from mpm_step_to_grid's");
  if (nslots < 0 || (nslots > 0 && (!deltas || !material_slot || nmat < (int32_t)e->materials.size())))
    return fail(e, MPM_ERR_INVALID_ARGUMENT, "deltas need a slot for every material");
  const size_t Nn = (size_t)e->g.N * e->g.N * e->g.N;
  float4* dd = nullptr;
  int32_t* ds = nullptr;
  std::vector<float4> hd;
  hipError_t st = hipSuccess;
  if (nslots > 0) {
    int32_t table[64];
    for (int k = 0; k < 64; ++k) {
      table[k] = k < nmat ? material_slot[k] : -1;
      if (table[k] >= nslots || table[k] < -1) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad delta slot %d", table[k]);
    }
    hd.resize((size_t)nslots * Nn);
    for (size_t k = 0; k < hd.size(); ++k)
      hd[k] = make_float4((float)deltas[3 * k], (float)deltas[3 * k + 1], (float)deltas[3 * k + 2], 0.f);
    st = hipMallocAsync((void**)&dd, sizeof(float4) * hd.size(), e->stream);
    if (st == hipSuccess) st = hipMallocAsync((void**)&ds, sizeof table, e->stream);
    if (st == hipSuccess)
      st = hipMemcpyAsync(dd, hd.data(), sizeof(float4) * hd.size(), hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) st = hipMemcpyAsync(ds, table, sizeof table, hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);  // host buffers go out of scope
  }
  if (st == hipSuccess) {
    const StepParams sp = params(e, dt);
    e->d.deltas = dd;
    e->d.delta_slot = ds;
    Timed t(e, "g2p", true);
    launch_fused(e->stream, e->d, e->n, e->g, sp, 2, t.ev());
    e->d.deltas = nullptr;
    e->d.delta_slot = nullptr;
    st = hipGetLastError();
  }
  if (dd) (void)hipFreeAsync(dd, e->stream);
  if (ds) (void)hipFreeAsync(ds, e->stream);
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "step from grid: %s", hipGetErrorString(st));
  e->steps_since_sort++;
  e->split_pending = false;
  return MPM_OK;
}

int mpm_synchronize(mpm_engine* e) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  HIP_OK(e, hipMemcpyAsync(e->status_host, e->d.status, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
  launch_zero_u32(e->stream, e->d.status, 1);
  HIP_OK(e, hipGetLastError());
  HIP_OK(e, hipStreamSynchronize(e->stream));
  const uint32_t status = *e->status_host;
  if (status & STATUS_INVALID)
    return fail(e, MPM_ERR_INVALID_STATE,
                "GPU state became invalid: det F outside [1e-6, 1e6], a value that is not a number, |v| above 1e5 "
                "or a particle outside the domain");
  if (status & STATUS_CFL) return fail(e, MPM_ERR_CFL, "MPM CFL limit exceeded; reduce dt");
  return MPM_OK;
}

int mpm_count(const mpm_engine* e, uint32_t* out) {
  if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
  *out = e->n;
  return MPM_OK;
}

int mpm_stable_dt(mpm_engine* e, double cap, double cfl, double* out) {
  if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
  const int nb = particle_reduce_blocks();
  float* dbuf = nullptr;
  HIP_OK(e, hipMallocAsync((void**)&dbuf, sizeof(float) * 2 * nb, e->stream));
  launch_stable(e->stream, e->d, e->n, dbuf);
  std::vector<float> h(2 * nb);
  hipError_t st = hipMemcpyAsync(h.data(), dbuf, sizeof(float) * 2 * nb, hipMemcpyDeviceToHost, e->stream);
  hipFreeAsync(dbuf, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "stable_dt: %s", hipGetErrorString(st));
  double mx = 1, grad = 1;
  for (int b = 0; b < nb; ++b) {
    mx = std::max(mx, (double)h[2 * b]);
    grad = std::max(grad, (double)h[2 * b + 1]);
  }
  const double H = e->cfg.cell_size;
  *out = std::min(std::min(cap > 0 ? cap : H / 8 / 240, (cfl > 0 ? cfl * H : 1.4 * H / 8) / mx), .08 / grad);
  return MPM_OK;
}

int mpm_conservation_report(mpm_engine* e, mpm_conservation* out) {
  if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
  const int K = particle_reduce_k(), nb = particle_reduce_blocks();
  double* dbuf = nullptr;
  HIP_OK(e, hipMallocAsync((void**)&dbuf, sizeof(double) * K * nb, e->stream));
  const float D = (float)(e->cfg.cell_size * e->cfg.cell_size / 4);
  launch_particle_reduce(e->stream, e->d, e->n, (float)e->mass_scale, D, dbuf);
  std::vector<double> h((size_t)K * nb);
  hipError_t st = hipMemcpyAsync(h.data(), dbuf, sizeof(double) * K * nb, hipMemcpyDeviceToHost, e->stream);
  if (st == hipSuccess)
    st = hipMemcpyAsync(e->red_host, e->d.red, sizeof(double) * 4 * SUM_SLOTS, hipMemcpyDeviceToHost, e->stream);
  hipFreeAsync(dbuf, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "conservation: %s", hipGetErrorString(st));
  finish_conservation(h.data(), K, nb, e->red_host, SUM_SLOTS, e->particle_mass_q, e->mass_scale, e->n, out);
  return MPM_OK;
}

int mpm_refresh_diagnostics(mpm_engine* e, double* stress) {
  if (!e || !stress) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = e->n;
  if (!n) return MPM_OK;
  float* dbuf = nullptr;
  HIP_OK(e, hipMallocAsync((void**)&dbuf, sizeof(float) * n, e->stream));
  launch_von_mises(e->stream, e->d, n, dbuf);
  std::vector<float> h(n);
  std::vector<uint32_t> ids(n);
  hipError_t st = hipMemcpyAsync(h.data(), dbuf, sizeof(float) * n, hipMemcpyDeviceToHost, e->stream);
  if (st == hipSuccess) st = hipMemcpyAsync(ids.data(), e->d.ids, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, e->stream);
  hipFreeAsync(dbuf, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "diagnostics: %s", hipGetErrorString(st));
  for (uint32_t s = 0; s < n; ++s) stress[ids[s]] = h[s];
  return MPM_OK;
}

int mpm_stress_tensors(mpm_engine* e, double* tau) {
  if (!e || !tau) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = e->n;
  if (!n) return MPM_OK;
  float* dbuf = nullptr;
  HIP_OK(e, hipMallocAsync((void**)&dbuf, sizeof(float) * 6 * (size_t)n, e->stream));
  launch_von_mises(e->stream, e->d, n, nullptr, dbuf);
  std::vector<float> h(6 * (size_t)n);
  std::vector<uint32_t> ids(n);
  hipError_t st = hipGetLastError();
  if (st == hipSuccess)
    st = hipMemcpyAsync(h.data(), dbuf, sizeof(float) * h.size(), hipMemcpyDeviceToHost, e->stream);
  if (st == hipSuccess) st = hipMemcpyAsync(ids.data(), e->d.ids, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, e->stream);
  hipFreeAsync(dbuf, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "stress tensors: %s", hipGetErrorString(st));
  for (uint32_t s = 0; s < n; ++s) {
    const float* t = &h[6 * (size_t)s];
    double* o = tau + 9 * (size_t)ids[s];
    o[0] = t[0];
    o[4] = t[1];
    o[8] = t[2];
    o[1] = o[3] = t[3];
    o[2] = o[6] = t[4];
    o[5] = o[7] = t[5];
  }
  return MPM_OK;
}

int mpm_preview_layout(uint32_t count, uint32_t limit, uint32_t* stride, uint32_t* rows) {
  if (!stride || !rows || limit == 0) return MPM_ERR_INVALID_ARGUMENT;
  *stride = std::max<uint32_t>(1, (uint32_t)((count + (uint64_t)limit - 1) / limit));
  *rows = (uint32_t)((count + (uint64_t)*stride - 1) / *stride);
  return MPM_OK;
}

int mpm_snapshot(mpm_engine* e, uint32_t limit, const int32_t* codes, int32_t ncodes, float* data,
                 mpm_snapshot_stats* st) {
  if (!e || !st) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = e->n;
  uint32_t stride, rows;
  if (mpm_preview_layout(n, limit, &stride, &rows)) return fail(e, MPM_ERR_INVALID_ARGUMENT, "preview limit 0");
  if (codes && ncodes < (int32_t)e->materials.size())
    return fail(e, MPM_ERR_INVALID_ARGUMENT, "%d material codes for %zu materials", ncodes, e->materials.size());
  *st = mpm_snapshot_stats{};
  st->count = n;
  st->sample_stride = stride;
  st->render_count = rows;
  st->min_J = 1;
  if (!n) return MPM_OK;
  int32_t table[64];
  for (int k = 0; k < 64; ++k) table[k] = codes && k < ncodes ? codes[k] : 11;
  const uint32_t nb = snapshot_blocks(n);
  int32_t* dcodes = nullptr;
  float* ddata = nullptr;
  double* dpart = nullptr;
  unsigned long long* dcounts = nullptr;
  hipError_t err = hipMallocAsync((void**)&dcodes, sizeof table, e->stream);
  if (err == hipSuccess) err = hipMallocAsync((void**)&dpart, sizeof(double) * 6 * nb, e->stream);
  if (err == hipSuccess) err = hipMallocAsync((void**)&dcounts, sizeof(unsigned long long) * 13, e->stream);
  if (err == hipSuccess && data) err = hipMallocAsync((void**)&ddata, sizeof(float) * 8 * (size_t)rows, e->stream);
  if (err == hipSuccess) err = hipMemcpyAsync(dcodes, table, sizeof table, hipMemcpyHostToDevice, e->stream);
  if (err == hipSuccess) err = hipMemsetAsync(dcounts, 0, sizeof(unsigned long long) * 13, e->stream);
  std::vector<double> part(6 * (size_t)nb);
  unsigned long long counts[13];
  if (err == hipSuccess) {
    launch_snapshot(e->stream, e->d, n, dcodes, stride, ddata, dpart, dcounts);
    err = hipGetLastError();
  }
  if (err == hipSuccess)
    err = hipMemcpyAsync(part.data(), dpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipMemcpyAsync(counts, dcounts, sizeof counts, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess && data)
    err = hipMemcpyAsync(data, ddata, sizeof(float) * 8 * (size_t)rows, hipMemcpyDeviceToHost, e->stream);
  for (void* p : {(void*)dcodes, (void*)ddata, (void*)dpart, (void*)dcounts})
    if (p) (void)hipFreeAsync(p, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return fail(e, MPM_ERR_HIP, "snapshot: %s", hipGetErrorString(err));
  st->min_temperature = 1e30;
  st->max_temperature = -1e30;
  for (uint32_t b = 0; b < nb; ++b) {
    const double* r = &part[6 * (size_t)b];
    st->mass += r[0];
    st->min_temperature = std::min(st->min_temperature, r[1]);
    st->max_temperature = std::max(st->max_temperature, r[2]);
    st->min_J = std::min(st->min_J, r[3]);
    st->max_stress = std::max(st->max_stress, r[4]);
    st->max_plastic = std::max(st->max_plastic, r[5]);
  }
  for (int k = 0; k < 12; ++k) {
    st->models[k] = (int64_t)counts[k];
    st->solids += (int64_t)counts[k];
  }
  st->nonfinite = (int64_t)counts[12];
  return MPM_OK;
}

int mpm_frame_packet(const char* json, uint32_t json_len, const float* data, uint64_t rows, uint8_t* out,
                     uint64_t capacity, uint64_t* length) {
  if (!length || (json_len && !json) || (rows && !data)) return MPM_ERR_INVALID_ARGUMENT;
  const uint64_t pad = (4 - json_len % 4) % 4;
  const uint64_t need = 4 + json_len + pad + 32 * rows;
  *length = need;
  if (!out) return MPM_OK;
  if (capacity < need) return MPM_ERR_CAPACITY;
  for (int k = 0; k < 4; ++k) out[k] = (uint8_t)(json_len >> (8 * k));  // little-endian u32
  if (json_len) std::memcpy(out + 4, json, json_len);
  std::memset(out + 4 + json_len, ' ', pad);
  uint8_t* q = out + 4 + json_len + pad;
  for (uint64_t i = 0; i < 8 * rows; ++i) {
    uint32_t bits;
    std::memcpy(&bits, data + i, 4);
    for (int k = 0; k < 4; ++k) q[4 * i + k] = (uint8_t)(bits >> (8 * k));  // little-endian float32
  }
  return MPM_OK;
}

int mpm_profile(mpm_engine* e, int32_t enable, int32_t reset) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  collect_times(e);
  if (reset) {
    e->times.clear();
    e->order.clear();
  }
  e->profiling = enable != 0;
  return MPM_OK;
}

int mpm_kernel_times(mpm_engine* e, mpm_kernel_time* out, int32_t max, int32_t* n) {
  if (!e || !n) return MPM_ERR_INVALID_ARGUMENT;
  collect_times(e);
  int k = 0;
  for (const auto& name : e->order) {
    if (out && k < max) {
      std::memset(&out[k], 0, sizeof out[k]);
      std::snprintf(out[k].name, sizeof out[k].name, "%s", name.c_str());
      out[k].launches = e->times[name].first;
      out[k].total_ms = e->times[name].second;
    }
    ++k;
  }
  *n = k;
  return MPM_OK;
}

int mpm_debug_bin(mpm_engine* e, uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = e->n;
  if (!n) return MPM_OK;
  launch_bin_keys(e->stream, e->d, n, e->g);
  if (keys_in) HIP_OK(e, hipMemcpyAsync(keys_in, e->d.keys, 4ull * n, hipMemcpyDeviceToHost, e->stream));
  std::vector<uint32_t> ids(n);
  HIP_OK(e, hipMemcpyAsync(ids.data(), e->d.ids, 4ull * n, hipMemcpyDeviceToHost, e->stream));
  radix_sort_pairs(e->stream, e->d.keys, e->d.vals, e->d.keys_alt, e->d.vals_alt, e->d.hist, e->d.partials, n,
                   key_bits(e->g.N));
  std::vector<uint32_t> order(n);
  if (keys_out) HIP_OK(e, hipMemcpyAsync(keys_out, e->d.keys, 4ull * n, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(e, hipMemcpyAsync(order.data(), e->d.vals, 4ull * n, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(e, hipGetLastError());
  HIP_OK(e, hipStreamSynchronize(e->stream));
  if (perm)
    for (uint32_t s = 0; s < n; ++s) perm[s] = ids[order[s]];
  return MPM_OK;
}

int mpm_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t bits, uint32_t* keys_out, uint32_t* perm_out) {
  if (n && (!keys || !keys_out || !perm_out)) return MPM_ERR_INVALID_ARGUMENT;
  if (bits > 32) return MPM_ERR_INVALID_ARGUMENT;
  if (n == 0) return MPM_OK;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MPM_ERR_HIP;
  uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *hist = nullptr, *part = nullptr;
  hipError_t st = hipMalloc((void**)&k0, 4ull * n);
  if (st == hipSuccess) st = hipMalloc((void**)&k1, 4ull * n);
  if (st == hipSuccess) st = hipMalloc((void**)&v0, 4ull * n);
  if (st == hipSuccess) st = hipMalloc((void**)&v1, 4ull * n);
  if (st == hipSuccess) st = hipMalloc((void**)&hist, 4 * radix_hist_words(n));
  if (st == hipSuccess) st = hipMalloc((void**)&part, 4 * radix_partial_words(n));
  std::vector<uint32_t> iota(n);
  for (uint32_t i = 0; i < n; ++i) iota[i] = i;
  if (st == hipSuccess) st = hipMemcpyAsync(k0, keys, 4ull * n, hipMemcpyHostToDevice, s);
  if (st == hipSuccess) st = hipMemcpyAsync(v0, iota.data(), 4ull * n, hipMemcpyHostToDevice, s);
  if (st == hipSuccess) {
    radix_sort_pairs(s, k0, v0, k1, v1, hist, part, n, bits);
    st = hipGetLastError();
  }
  if (st == hipSuccess) st = hipMemcpyAsync(keys_out, k0, 4ull * n, hipMemcpyDeviceToHost, s);
  if (st == hipSuccess) st = hipMemcpyAsync(perm_out, v0, 4ull * n, hipMemcpyDeviceToHost, s);
  if (st == hipSuccess) st = hipStreamSynchronize(s);
  for (void* p : {(void*)k0, (void*)k1, (void*)v0, (void*)v1, (void*)hist, (void*)part})
    if (p) hipFree(p);
  hipStreamDestroy(s);
  return st == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

int mpm_return_map_batch(const mpm_material* m, uint32_t n, float dt, const float* F, const float* ep,
                         const float* T, float* F_out, float* ep_out, float* tau_out) {
  if (!m || (n && (!F || !ep || !T || !F_out || !ep_out || !tau_out))) return MPM_ERR_INVALID_ARGUMENT;
  if (const int rc = mpm_material_check(m)) return rc;
  const DevMaterial dm = to_dev_material(*m);
  return batch_call(n, {{F, 9}, {ep, 1}, {T, 1}}, {{F_out, 9}, {ep_out, 1}, {tau_out, 6}},
                    [&](hipStream_t s, std::vector<float*>& in, std::vector<float*>& out) {
                      launch_return_batch(s, dm, n, dt, in[0], in[1], in[2], nullptr, nullptr, out[0], out[1],
                                          nullptr, nullptr, out[2]);
                    });
}

int mpm_return_map_state_batch(const mpm_material* m, uint32_t n, float dt, const float* F, const float* ep,
                               const float* T, const float* jp, const float* dilation, float* F_out, float* ep_out,
                               float* jp_out, float* dilation_out, float* tau_out) {
  if (!m || (n && (!F || !ep || !T || !jp || !dilation || !F_out || !ep_out || !jp_out || !dilation_out || !tau_out)))
    return MPM_ERR_INVALID_ARGUMENT;
  if (const int rc = mpm_material_check(m)) return rc;
  const DevMaterial dm = to_dev_material(*m);
  return batch_call(n, {{F, 9}, {ep, 1}, {T, 1}, {jp, 1}, {dilation, 1}},
                    {{F_out, 9}, {ep_out, 1}, {jp_out, 1}, {dilation_out, 1}, {tau_out, 6}},
                    [&](hipStream_t s, std::vector<float*>& in, std::vector<float*>& out) {
                      launch_return_batch(s, dm, n, dt, in[0], in[1], in[2], in[3], in[4], out[0], out[1], out[2],
                                          out[3], out[4]);
                    });
}

int mpm_stress_batch(const mpm_material* m, uint32_t n, const float* F, float* tau_out) {
  if (!m || (n && (!F || !tau_out))) return MPM_ERR_INVALID_ARGUMENT;
  if (!supported_model(m->model)) return MPM_ERR_UNSUPPORTED;
  const DevMaterial dm = to_dev_material(*m);
  return batch_call(n, {{F, 9}}, {{tau_out, 6}}, [&](hipStream_t s, std::vector<float*>& in, std::vector<float*>& out) {
    launch_stress_batch(s, dm, n, in[0], out[0]);
  });
}

int mpm_debug_p2g(mpm_engine* e, double dt, double* grid_mass, double* grid_momentum, int32_t* Nout) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  const int N = e->g.N;
  if (Nout) *Nout = N;
  if (!grid_mass && !grid_momentum) return MPM_OK;
  const size_t Nn = (size_t)N * N * N;
  StepParams sp = params(e, dt);
  sp.guard = 0;  // a transfer for inspection, not a step
  launch_fused(e->stream, e->d, e->n, e->g, sp, 0);
  std::vector<unsigned long long> m(Nn);
  std::vector<float> p(3 * Nn);
  HIP_OK(e, hipMemcpyAsync(m.data(), e->d.acc_m, 8 * Nn, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(e, hipMemcpyAsync(p.data(), e->d.acc_p, 12 * Nn, hipMemcpyDeviceToHost, e->stream));
  // Leave the accumulators clear for the next step.
  HIP_OK(e, hipMemsetAsync(e->d.acc_m, 0, 8 * Nn, e->stream));
  HIP_OK(e, hipMemsetAsync(e->d.acc_p, 0, 12 * Nn, e->stream));
  HIP_OK(e, hipMemsetAsync(e->d.fixed, 0, Nn, e->stream));
  HIP_OK(e, hipMemsetAsync(e->d.pieces, 0, piece_flag_bytes(pieces_of_nodes(Nn)), e->stream));
  HIP_OK(e, hipGetLastError());
  HIP_OK(e, hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < Nn; ++i) {
    if (grid_mass) grid_mass[i] = (double)(int64_t)m[i] / e->mass_scale;
    if (grid_momentum)
      for (int a = 0; a < 3; ++a) grid_momentum[i * 3 + a] = p[a * Nn + i];
  }
  return MPM_OK;
}

int mpm_set_surrogate(mpm_engine* e, int32_t material, const float* params, const mpm_mlp_norm* norm) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  if (!params) {
    if (e->d.mlp) {
      HIP_OK(e, hipStreamSynchronize(e->stream));
      HIP_OK(e, hipFree(e->d.mlp));
      e->d.mlp = nullptr;
    }
    return MPM_OK;
  }
  if (!norm || material < 0 || material >= (int32_t)e->materials.size())
    return fail(e, MPM_ERR_INVALID_ARGUMENT, "surrogate: material index %d outside the table", material);
  const mpm_material& m = e->materials[material];
  const bool rate_free = (m.model == MPM_MODEL_J2) || (m.model == MPM_MODEL_THERMAL_J2 && m.th_H == 0.f);
  if (!rate_free)
    return fail(e, MPM_ERR_UNSUPPORTED,
                "surrogate: material %d (model %d) is not rate-independent J2 without hardening", material, m.model);
  if (!(norm->in_scale > 0) || !(norm->out_scale > 0) || !(norm->eps_max > 0) || !(norm->T_hi >= norm->T_lo))
    return fail(e, MPM_ERR_INVALID_ARGUMENT, "surrogate: invalid normalisation");
  if (mpm_mlp_check_material(&m, norm) != MPM_OK)
    return fail(e, MPM_ERR_INVALID_ARGUMENT,
                "surrogate: material %d differs from the network's training material (model, mu/E, lambda/E, "
                "yield ratio, thermal softening)", material);
  for (int k = 0; k < MPM_MLP_NPARAM; ++k)
    if (!std::isfinite(params[k])) return fail(e, MPM_ERR_INVALID_ARGUMENT, "surrogate: parameter %d not finite", k);
  MlpDev* host = new MlpDev;
  MlpNorm nm;
  nm.in_scale = norm->in_scale;
  nm.T_center = norm->T_center;
  nm.T_scale = norm->T_scale;
  nm.out_scale = norm->out_scale;
  nm.eps_max = norm->eps_max;
  nm.T_lo = norm->T_lo;
  nm.T_hi = norm->T_hi;
  nm.min_yield_ratio = norm->min_yield_ratio;
  nm.material = material;
  pack_mlp(params, nm, *host);
  hipError_t st = hipSuccess;
  if (!e->d.mlp) st = hipMalloc((void**)&e->d.mlp, sizeof(MlpDev));
  if (st == hipSuccess && !e->d.sur_counts) {
    st = hipMalloc((void**)&e->d.sur_counts, sizeof(unsigned long long) * 4 * SUR_SLOTS);
    if (st == hipSuccess) st = hipMemsetAsync(e->d.sur_counts, 0, sizeof(unsigned long long) * 4 * SUR_SLOTS, e->stream);
  }
  if (st == hipSuccess) st = hipMemcpyAsync(e->d.mlp, host, sizeof(MlpDev), hipMemcpyHostToDevice, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  delete host;
  if (st != hipSuccess) return fail(e, MPM_ERR_HIP, "surrogate upload: %s", hipGetErrorString(st));
  return MPM_OK;
}

int mpm_surrogate_counts(mpm_engine* e, uint64_t* used, uint64_t* elastic, uint64_t* fallback, int32_t reset) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  uint64_t tot[3] = {0, 0, 0};
  if (e->d.sur_counts) {
    std::vector<unsigned long long> h(4 * SUR_SLOTS);
    HIP_OK(e, hipMemcpyAsync(h.data(), e->d.sur_counts, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                             e->stream));
    HIP_OK(e, hipStreamSynchronize(e->stream));
    for (int s = 0; s < SUR_SLOTS; ++s)
      for (int k = 0; k < 3; ++k) tot[k] += h[4 * s + k];
    if (reset) {
      HIP_OK(e, hipMemsetAsync(e->d.sur_counts, 0, sizeof(unsigned long long) * h.size(), e->stream));
      HIP_OK(e, hipStreamSynchronize(e->stream));
    }
  }
  if (used) *used = tot[0];
  if (elastic) *elastic = tot[1];
  if (fallback) *fallback = tot[2];
  return MPM_OK;
}

int mpm_frame_defaults(const mpm_engine* e, mpm_frame_config* c) {
  if (!e || !c) return MPM_ERR_INVALID_ARGUMENT;
  std::memset(c, 0, sizeof *c);
  c->time_step_scale = 1.0;
  c->max_substeps = 48;
  c->velocity_limit = 160.0;
  for (int a = 0; a < 3; ++a) {
    c->bounds_lo[a] = e->cfg.lower[a];
    c->bounds_hi[a] = e->cfg.upper[a];
  }
  c->anti_clump = 1;
  c->volume_unit = 1.0 / 64.0;
  c->gravity_mode = MPM_GRAVITY_UNIFORM;
  c->gravity_strength = 1.0;
  c->gravity_direction[1] = -1.0;
  c->gravity_softening = 8.0;
  return MPM_OK;
}

int mpm_advance_frame(mpm_engine* e, const mpm_frame_config* c, double dt, mpm_frame_report* out) {
  if (e && e->split_pending) return fail(e, MPM_ERR_NOT_READY, "a split step waits for mpm_step_from_grid");
  if (!e || !c) return MPM_ERR_INVALID_ARGUMENT;
  if (!(dt > 0) || !std::isfinite(dt)) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad dt");
  if (!(c->time_step_scale > 0) || c->max_substeps < 1 || !(c->velocity_limit > 0) || c->gravity_mode < 0 ||
      c->gravity_mode > 3 || !(c->volume_unit > 0))
    return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad frame settings");
  for (int a = 0; a < 3; ++a)
    if (!(c->bounds_hi[a] >= c->bounds_lo[a])) return fail(e, MPM_ERR_INVALID_ARGUMENT, "bad bounds");
  // The spacing rule reaches 5.6 in the reference's units; the neighbour
  // search covers it with +-1 base cells for h >= 5.6 and +-2 for h >= 2.8.
  const double h = e->cfg.cell_size;
  const int R = h >= 5.6 ? 1 : 2;
  if (c->anti_clump && h < 2.8)
    return fail(e, MPM_ERR_INVALID_ARGUMENT, "anti-clump spacing rule needs a cell size >= 2.8 (reference units)");
  const uint32_t cap = e->d.cap;
  const size_t Nn = (size_t)blocked_key_count(e->g.N);  // cell-start table, indexed by sort key
  if (!e->frame_ox) HIP_OK(e, hipMalloc((void**)&e->frame_ox, sizeof(float) * 3 * (size_t)cap));
  if (c->anti_clump) {
    if (!e->frame_disp) HIP_OK(e, hipMalloc((void**)&e->frame_disp, sizeof(float) * 3 * (size_t)cap));
    if (!e->frame_cell_start) HIP_OK(e, hipMalloc((void**)&e->frame_cell_start, sizeof(uint32_t) * Nn));
    if (!e->frame_max) HIP_OK(e, hipMalloc((void**)&e->frame_max, sizeof(unsigned int)));
  }
  FrameDev f;
  for (int a = 0; a < 3; ++a) {
    f.lo[a] = (float)c->bounds_lo[a];
    f.hi[a] = (float)c->bounds_hi[a];
    f.gdir[a] = (float)c->gravity_direction[a];
    f.gcenter[a] = (float)c->gravity_center[a];
  }
  f.vlimit = (float)c->velocity_limit;
  f.gmode = c->gravity_mode;
  f.gstrength = (float)c->gravity_strength;
  f.gsoft = (float)c->gravity_softening;
  f.vol_unit = (float)c->volume_unit;
  double remaining = dt;
  int steps = 0;
  while (remaining > 1e-9 && steps < c->max_substeps) {
    double stable = 0;
    int rc = mpm_stable_dt(e, 0, 0, &stable);
    if (rc) return rc;
    const double hs = std::min(remaining, stable * c->time_step_scale);
    launch_frame_pre(e->stream, e->d, e->n, e->frame_ox, f, (float)hs);
    rc = mpm_step(e, hs, 1);
    if (rc) return rc;
    launch_frame_post(e->stream, e->d, e->n, e->frame_ox, f, (float)hs);
    if (c->anti_clump && e->n) {
      rc = rebin(e);
      if (rc) return rc;
      launch_regularize(e->stream, e->d, e->n, e->g, R, f, (float)hs, e->frame_cell_start, e->frame_disp,
                        e->frame_max);
    }
    HIP_OK(e, hipGetLastError());
    remaining -= hs;
    ++steps;
  }
  if (out) {
    out->substeps = steps;
    out->advanced = dt - remaining;
  }
  return MPM_OK;
}

}  // extern "C"
