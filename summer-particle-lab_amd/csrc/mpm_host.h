// Host-side conversion between a caller's particle view (mpm_particles, FP64
// or FP32, array of structures for tensors) and the engine's FP32 planes.
// Shared by the single-scene engine and the batch engine.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>

#include "mpm_internal.h"

namespace mpm {

template <typename T>
inline double rd(const void* p, size_t i) {
  return (double)static_cast<const T*>(p)[i];
}
inline double rdv(const void* p, int dtype, size_t i) {
  return dtype == MPM_F64 ? rd<double>(p, i) : rd<float>(p, i);
}
inline void wrv(void* p, int dtype, size_t i, double v) {
  if (dtype == MPM_F64)
    static_cast<double*>(p)[i] = v;
  else
    static_cast<float*>(p)[i] = (float)v;
}

// Checks a view for upload: dtype and the required arrays.
inline bool particles_complete(const mpm_particles* p) {
  if (p->dtype != MPM_F64 && p->dtype != MPM_F32) return false;
  return !p->count || (p->x && p->y && p->z && p->vx && p->vy && p->vz && p->F && p->C && p->mass && p->volume &&
                       p->temperature && p->plastic && p->material && p->pinned);
}

// Packs p into planes (stride `stride` floats per plane) at slots
// [slot0, slot0 + count), with ids id0 + i and meta material | flags << 8.
// Returns an empty string or the reason the view was rejected.
inline std::string pack_particles(const mpm_particles* p, size_t nmaterials, float* planes, size_t stride,
                                  uint32_t slot0, uint32_t id0, uint32_t* ids, uint32_t* meta) {
  const int dt = p->dtype;
  char buf[160];
  for (uint32_t i = 0; i < p->count; ++i) {
    if (p->material[i] >= nmaterials) {
      snprintf(buf, sizeof buf, "particle %u: material %u out of range", i, p->material[i]);
      return buf;
    }
    const double m = rdv(p->mass, dt, i), vol = rdv(p->volume, dt, i);
    if (!(m > 0) || !(vol > 0)) {
      snprintf(buf, sizeof buf, "particle %u: mass and volume must be > 0", i);
      return buf;
    }
    const size_t s = (size_t)slot0 + i;
    auto set = [&](int k, double v) { planes[(size_t)k * stride + s] = (float)v; };
    set(P_X, rdv(p->x, dt, i));
    set(P_Y, rdv(p->y, dt, i));
    set(P_Z, rdv(p->z, dt, i));
    set(P_VX, rdv(p->vx, dt, i));
    set(P_VY, rdv(p->vy, dt, i));
    set(P_VZ, rdv(p->vz, dt, i));
    for (int k = 0; k < 9; ++k) {
      set(P_C0 + k, rdv(p->C, dt, (size_t)i * 9 + k));
      set(P_F0 + k, rdv(p->F, dt, (size_t)i * 9 + k));
    }
    uint32_t flags = p->pinned[i] ? FLAG_PINNED : 0u;
    if (p->cached_tau && p->cache_valid && p->cache_valid[i]) {
      const size_t b = (size_t)i * 9;
      set(P_T0 + 0, rdv(p->cached_tau, dt, b + 0));
      set(P_T0 + 1, rdv(p->cached_tau, dt, b + 4));
      set(P_T0 + 2, rdv(p->cached_tau, dt, b + 8));
      set(P_T0 + 3, 0.5 * (rdv(p->cached_tau, dt, b + 1) + rdv(p->cached_tau, dt, b + 3)));
      set(P_T0 + 4, 0.5 * (rdv(p->cached_tau, dt, b + 2) + rdv(p->cached_tau, dt, b + 6)));
      set(P_T0 + 5, 0.5 * (rdv(p->cached_tau, dt, b + 5) + rdv(p->cached_tau, dt, b + 7)));
      flags |= FLAG_CACHE;
    } else {
      for (int k = 0; k < 6; ++k) set(P_T0 + k, 0.0);
    }
    set(P_MASS, m);
    set(P_VOL, vol);
    set(P_TEMP, rdv(p->temperature, dt, i));
    set(P_EP, rdv(p->plastic, dt, i));
    set(P_JP, p->jp ? rdv(p->jp, dt, i) : 1.0);
    set(P_DIL, p->dilation ? rdv(p->dilation, dt, i) : 0.0);
    set(P_DMG, p->damage ? std::min(std::max(rdv(p->damage, dt, i), 0.0), 1.0) : 0.0);
    ids[i] = id0 + i;
    meta[i] = (uint32_t)p->material[i] | (flags << 8);
  }
  return std::string();
}

// Fixed-point mass exponent: unit 2^-k with the heaviest particle at most
// 2^24 units, so a particle's 27 integer parts and the sums of up to 64 of
// them inside a wave fit 32 bits; node and grand totals are 64-bit.
inline int mass_exponent(double mmax) {
  int k = 0;
  if (mmax > 0) k = (int)std::floor(24.0 - std::log2(mmax));
  return std::max(-100, std::min(100, k));
}

// Writes the state of `count` slots (planes with stride `stride`, starting at
// slot0) back to p in original particle order: slot s holds original id
// ids[s] - id0.
inline void unpack_particles(const float* planes, size_t stride, uint32_t slot0, const uint32_t* ids,
                             const uint32_t* meta, uint32_t count, uint32_t id0, mpm_particles* p) {
  const int dt = p->dtype;
  for (uint32_t j = 0; j < count; ++j) {
    const size_t s = (size_t)slot0 + j;
    const uint32_t i = ids[j] - id0;
    auto get = [&](int k) { return (double)planes[(size_t)k * stride + s]; };
    if (p->x) wrv(p->x, dt, i, get(P_X));
    if (p->y) wrv(p->y, dt, i, get(P_Y));
    if (p->z) wrv(p->z, dt, i, get(P_Z));
    if (p->vx) wrv(p->vx, dt, i, get(P_VX));
    if (p->vy) wrv(p->vy, dt, i, get(P_VY));
    if (p->vz) wrv(p->vz, dt, i, get(P_VZ));
    for (int k = 0; k < 9; ++k) {
      if (p->F) wrv(p->F, dt, (size_t)i * 9 + k, get(P_F0 + k));
      if (p->C) wrv(p->C, dt, (size_t)i * 9 + k, get(P_C0 + k));
    }
    if (p->mass) wrv(p->mass, dt, i, get(P_MASS));
    if (p->volume) wrv(p->volume, dt, i, get(P_VOL));
    if (p->temperature) wrv(p->temperature, dt, i, get(P_TEMP));
    if (p->plastic) wrv(p->plastic, dt, i, get(P_EP));
    if (p->jp) wrv(p->jp, dt, i, get(P_JP));
    if (p->dilation) wrv(p->dilation, dt, i, get(P_DIL));
    if (p->damage) wrv(p->damage, dt, i, get(P_DMG));
    if (p->cached_tau) {
      const double t[6] = {get(P_T0), get(P_T0 + 1), get(P_T0 + 2), get(P_T0 + 3), get(P_T0 + 4), get(P_T0 + 5)};
      const double full[9] = {t[0], t[3], t[4], t[3], t[1], t[5], t[4], t[5], t[2]};
      for (int k = 0; k < 9; ++k) wrv(p->cached_tau, dt, (size_t)i * 9 + k, full[k]);
    }
    if (p->material) p->material[i] = (uint8_t)(meta[j] & 255u);
    if (p->pinned) p->pinned[i] = (meta[j] >> 8) & FLAG_PINNED ? 1 : 0;
    if (p->cache_valid) p->cache_valid[i] = (meta[j] >> 8) & FLAG_CACHE ? 1 : 0;
  }
}

// Finishes a conservation report from the particle reduction records (nb
// records of K doubles, see k_particle_reduce) and the scene's grid-sum
// slots (nslots records of {fixed-point mass bits, momentum x, y, z}).
inline void finish_conservation(const double* h, int K, int nb, const double* grid, int nslots,
                                int64_t particle_mass_q, double mass_scale, uint32_t n, mpm_conservation* out) {
  std::memset(out, 0, sizeof *out);
  double acc[16] = {0};
  double minJ = 1.0, maxv = 0;
  for (int b = 0; b < nb; ++b) {
    for (int k = 0; k < 11; ++k) acc[k] += h[(size_t)b * K + k];
    minJ = std::min(minJ, h[(size_t)b * K + 11]);
    maxv = std::max(maxv, h[(size_t)b * K + 12]);
  }
  out->particle_mass = acc[0];
  out->particle_mass_q = particle_mass_q;
  int64_t gq = 0;
  double gp[3] = {0, 0, 0};
  for (int b = 0; b < nslots; ++b) {
    int64_t q;
    std::memcpy(&q, &grid[4 * (size_t)b], sizeof q);
    gq += q;
    for (int c = 0; c < 3; ++c) gp[c] += grid[4 * (size_t)b + 1 + c];
  }
  out->grid_mass_q = gq;
  out->mass_scale = mass_scale;
  out->grid_mass = (double)gq / mass_scale;
  for (int k = 0; k < 3; ++k) {
    out->momentum[k] = acc[2 + k];
    out->angular[k] = acc[5 + k];
    out->grid_momentum[k] = gp[k];
  }
  out->kinetic = acc[8];
  out->affine = acc[9];
  out->nonfinite = (int64_t)acc[10];
  out->min_J = n ? minJ : 1.0;
  out->max_speed = maxv;
}

// Number of distinct base cells (the cell of a particle's first stencil node,
// as the kernels compute it) over particles [0, n) of FP32 planes of stride
// cap. With the particle count it gives the mean particles per occupied
// cell, which picks the scatter's chunk width.
inline uint64_t occupied_cells(const float* planes, uint32_t cap, uint32_t n, const GridDims& g) {
  if (!n) return 0;
  const int M = g.N - 2;  // base cells per axis
  std::vector<uint64_t> bits(((size_t)M * M * M + 63) / 64, 0ull);
  uint64_t count = 0;
  for (uint32_t i = 0; i < n; ++i) {
    size_t cell = 0;
    for (int a = 2; a >= 0; --a) {
      const float q = (planes[(size_t)(P_X + a) * cap + i] - g.O[a]) * g.invH;
      int b = (int)std::floor(q - 0.5f);
      b = b < 0 ? 0 : (b > g.N - 3 ? g.N - 3 : b);
      cell = cell * M + (size_t)b;
    }
    const uint64_t bit = 1ull << (cell & 63);
    if (!(bits[cell >> 6] & bit)) {
      bits[cell >> 6] |= bit;
      ++count;
    }
  }
  return count;
}

// Scatter chunk width from the particle and occupied-cell counts: 8-lane
// chunks below a mean of 12 particles per cell (cells of 8 or fewer
// particles), 16 otherwise.
inline bool narrow_chunks(uint64_t particles, uint64_t cells) { return cells && particles < 12 * cells; }

// Guard thresholds of a step of size dt for a scene of configuration c
// (mpm_config's guard fields): a particle violates the stable step when
// cfl h / max(wave, 1) < 0.99999 dt, i.e. wave > cfl h / (0.99999 dt) (the
// floor of 1 cannot save it once that limit is >= 1), or when
// grad / max(sum |C|, 1) < 0.99999 dt (grad = grad_number, the reference's .08
// by default). Returns MPM_ERR_CFL when no state can pass: the cap term, or
// a limit below the floor of 1.
inline int guard_params(const mpm_config& c, double dt, StepParams& sp) {
  sp.guard = c.guard_off ? 0u : 1u;
  sp.wave_limit = sp.grad_limit = 3.0e38f;
  if (!sp.guard) return MPM_OK;
  const double lim = dt * 0.99999;
  const double cfl = c.cfl_number > 0 ? c.cfl_number : 1.4 / 8;
  const double wave = cfl * c.cell_size / lim, grad = (c.grad_number > 0 ? c.grad_number : .08) / lim;
  if ((c.cfl_cap > 0 && c.cfl_cap < lim) || wave < 1 || grad < 1) return MPM_ERR_CFL;
  sp.wave_limit = (float)std::min(wave, 3.0e38);
  sp.grad_limit = (float)std::min(grad, 3.0e38);
  return MPM_OK;
}

// Validates a scene configuration and derives its grid: the origin and node
// rule of src/three/discretization.js:13 with a cubic node block of edge
// max(nodes) (src/three/mpm.js:18), and the wall bands as node-index
// thresholds evaluated in FP64 exactly as the reference compares node
// positions O + i h with lower + h / upper - h. Returns an MPM status.
inline int grid_geometry(const mpm_config* c, GridDims& g, int lo_idx[3], int hi_idx[3]) {
  const double h = c->cell_size;
  if (!(h > 0) || !std::isfinite(h)) return MPM_ERR_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a) {
    if (!std::isfinite(c->lower[a]) || !std::isfinite(c->upper[a]) || c->upper[a] - c->lower[a] < h)
      return MPM_ERR_INVALID_ARGUMENT;
  }
  for (int f = 0; f < 6; ++f)
    if (c->wall_mode[f] < 0 || c->wall_mode[f] > 2) return MPM_ERR_INVALID_ARGUMENT;
  if (c->strict_domain) {  // src/three/discretization.js:8-10
    if (h < 4 || h > 16) return MPM_ERR_INVALID_ARGUMENT;
    for (int a = 0; a < 3; ++a) {
      const double cells = std::ceil(c->upper[a] / h - 1e-10) - std::floor(c->lower[a] / h + 1e-10);
      if (c->lower[a] < -2400 || c->upper[a] > 2400 || cells > 60) return MPM_ERR_INVALID_ARGUMENT;
    }
  }
  int N = 0;
  double origin[3];
  for (int a = 0; a < 3; ++a) {
    origin[a] = std::floor(c->lower[a] / h) * h - 2 * h;
    N = std::max(N, (int)(std::ceil((c->upper[a] - origin[a]) / h) + 3));
  }
  if ((uint64_t)N * N * N > (1ull << 31)) return MPM_ERR_CAPACITY;
  g = GridDims{};
  g.N = N;
  for (int a = 0; a < 3; ++a) {
    g.O[a] = (float)origin[a];
    g.lower[a] = (float)c->lower[a];
    g.upper[a] = (float)c->upper[a];
    g.gravity[a] = (float)c->gravity[a];
    int lo = -1, hi = N;
    for (int i = 0; i < N; ++i) {
      const double q = origin[a] + (double)i * h;
      if (q < c->lower[a] + h) lo = i;
      if (q > c->upper[a] - h && hi == N) hi = i;
    }
    lo_idx[a] = lo;
    hi_idx[a] = hi;
  }
  for (int f = 0; f < 6; ++f) g.wall[f] = c->wall_mode[f];
  g.H = (float)h;
  g.invH = (float)(1.0 / h);
  g.gravity_on_particles = c->gravity_on_particles ? 1 : 0;
  g.mass_eps = (float)c->mass_eps;
  return MPM_OK;
}

}  // namespace mpm
