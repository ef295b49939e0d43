// MLS-MPM step kernels for gfx950 (FP32 state, FP64 conservation sums).
//
// Grid update, binning, permutation, reductions and batch constitutive
// kernels. The particle transfers are in mpm_fused.hip.
//   k_grid  v = p/m, gravity, wall rules (src/three/mpm.js:89-102)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "mpm_internal.h"
#include "mpm_wave.h"

namespace mpm {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ float plane(const float* P, uint32_t cap, int k, uint32_t i) {
  return P[(size_t)k * cap + i];
}
__device__ __forceinline__ void set_plane(float* P, uint32_t cap, int k, uint32_t i, float v) {
  P[(size_t)k * cap + i] = v;
}

// Base node and offset along one axis: q = (x - O) * (1/h),
// b = clamp(floor(q - 1/2), 0, N - 3), t = q - b (src/three/mpm.js:28).
// Written with explicit round-to-nearest operations so the host can reproduce
// the binning key bit for bit.
__device__ __forceinline__ int base_of(float x, float O, float invH, int N, float& t) {
  const float q = __fmul_rn(__fsub_rn(x, O), invH);
  int b = (int)floorf(__fsub_rn(q, 0.5f));
  b = b < 0 ? 0 : (b > N - 3 ? N - 3 : b);
  t = __fsub_rn(q, (float)b);
  return b;
}

// Cell key of every slot (blocked_cell_key), written at the slot's original
// particle id so the stable sort breaks ties by ascending id (the reference's
// P2G visiting order within a cell, src/three/mpm.js:66); the value is the
// slot the particle lives in.
__global__ __launch_bounds__(BLOCK) void k_bin_keys(const float* __restrict__ P, const uint32_t* __restrict__ ids,
                                                    uint32_t cap, uint32_t n, GridDims g,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  float t;
  const int bx = base_of(plane(P, cap, P_X, i), g.O[0], g.invH, g.N, t);
  const int by = base_of(plane(P, cap, P_Y, i), g.O[1], g.invH, g.N, t);
  const int bz = base_of(plane(P, cap, P_Z, i), g.O[2], g.invH, g.N, t);
  const uint32_t id = ids[i];
  keys[id] = blocked_cell_key(bx, by, bz, g.N);
  vals[id] = i;
}

// Batched keys: chunk -> scene; the key is the base cell's blocked key in
// its scene. The sort runs per scene (a scene's slots and ids are its own
// range), so each scene's cells come in 4x4x4 blocks, as the single-scene
// binning does.
__global__ __launch_bounds__(BLOCK) void k_bin_keys_batch(const float* __restrict__ P,
                                                          const uint32_t* __restrict__ ids, uint32_t cap,
                                                          BatchArgs b, uint32_t chunk_size,
                                                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t c = blockIdx.x / (chunk_size / BLOCK);
  if (c >= b.nchunks) return;
  const uint2 ch = b.chunks[c];
  const SceneDesc* sd = b.scenes + ch.x;
  const uint32_t i = ch.y + (blockIdx.x % (chunk_size / BLOCK)) * BLOCK + threadIdx.x;
  if (i >= sd->p0 + sd->n || i >= ch.y + chunk_size) return;
  const GridDims& g = sd->g;
  float t;
  const int bx = base_of(plane(P, cap, P_X, i), g.O[0], g.invH, g.N, t);
  const int by = base_of(plane(P, cap, P_Y, i), g.O[1], g.invH, g.N, t);
  const int bz = base_of(plane(P, cap, P_Z, i), g.O[2], g.invH, g.N, t);
  const uint32_t id = ids[i];
  keys[id] = blocked_cell_key(bx, by, bz, g.N);  // sorted per scene
  vals[id] = i;
}

// Gather every plane of the particle state into sorted slot order.
__global__ __launch_bounds__(BLOCK) void k_permute(const float* __restrict__ P, float* __restrict__ Q,
                                                   const uint32_t* __restrict__ ids, uint32_t* __restrict__ ids_out,
                                                   const uint32_t* __restrict__ meta,
                                                   uint32_t* __restrict__ meta_out, uint32_t cap, uint32_t n,
                                                   const uint32_t* __restrict__ order) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t src = order[i];
#pragma unroll 4
  for (int k = 0; k < NPLANES; ++k) Q[(size_t)k * cap + i] = P[(size_t)k * cap + src];
  ids_out[i] = ids[src];
  meta_out[i] = meta[src];
}

// Kirchhoff stress of this step before the damage factor (6-vector), for
// the particle's effective model.
__device__ __forceinline__ void step_stress(const M3& F, int model, const DevMaterial& m, uint32_t flags,
                                            const float* __restrict__ P, uint32_t cap, uint32_t i, float tau[6]) {
  if (model >= 2 && (flags & FLAG_CACHE)) {
#pragma unroll
    for (int k = 0; k < 6; ++k) tau[k] = plane(P, cap, P_T0 + k, i);
  } else {
    elastic_tau(F, model, m, model == 2 ? plane(P, cap, P_JP, i) : 1.f, tau);
  }
}

// Grid update over the active pieces only. P2G flags every piece of 16
// consecutive nodes it adds mass to; k_compact_pieces appends the flagged pieces to a circular
// list (and clears the flags), and k_grid_list updates the pieces appended
// since the previous update: v = p/m, gravity, wall rules
// (src/three/mpm.js:89-102), clear of the accumulators. Grid sums
// (fixed-point mass, momentum) go to SUM_SLOTS slots per scene with atomics.
//
// counters (64-bit): [0] running append position, [1] first position of this
// step, [2] end position the previous update reached. Positions only grow,
// so no counter is ever reset; list slots are positions modulo the list
// capacity (at most every piece is appended in one step).

// Compaction of the piece flags, G flags per lane. G = 1: a ballot per wave.
// G = 16, for large grids: one 16-byte load per lane (flags are 0 or 1, so a
// word's popcount is its number of set flags) and a wave scan; the flag array
// is zero-padded to whole 16-byte groups. Either way a workgroup of 16 waves
// sums its waves' counts in LDS and takes its list positions with one counter
// atomic, so the single append counter sees one atomic per 1024 (G = 1) or
// 16384 (G = 16) flags. Block 0 also starts the step's list at the previous
// update's end and clears this step's grid sums.
constexpr uint32_t COMPACT_BLOCK = 1024;

template <int G>
__global__ __launch_bounds__(COMPACT_BLOCK) void k_compact_pieces(uint8_t* __restrict__ pieces, uint32_t npieces,
                                                                  uint32_t* __restrict__ list,
                                                                  unsigned long long* __restrict__ counters,
                                                                  double* __restrict__ sums, uint32_t nsums) {
  static_assert(G == 1 || G == 16, "1 or 16 flags per lane");
  __shared__ uint32_t s_off[COMPACT_BLOCK / 64];
  __shared__ unsigned long long s_base;
  if (blockIdx.x == 0) {
    for (uint32_t k = threadIdx.x; k < nsums; k += COMPACT_BLOCK) sums[k] = 0.0;
    if (threadIdx.x == 0) counters[1] = counters[2];
  }
  const uint32_t i = blockIdx.x * COMPACT_BLOCK + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  // This lane's flag count, the count of the lanes before it in the wave and
  // the wave's total.
  uint32_t cnt, before, total;
  uint4 f = make_uint4(0u, 0u, 0u, 0u);
  if constexpr (G == 1) {
    cnt = 0;
    if (i < npieces && pieces[i] != 0) {
      cnt = 1;
      pieces[i] = 0;
    }
    const uint64_t m = __ballot(cnt != 0);
    before = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    total = (uint32_t)__popcll(m);
  } else {
    uint4* groups = reinterpret_cast<uint4*>(pieces);  // group i holds flags 16i .. 16i+15
    if (i < (npieces + 15u) / 16u) f = groups[i];
    cnt = __popc(f.x) + __popc(f.y) + __popc(f.z) + __popc(f.w);
    uint32_t inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(inc, d, 64);
      if ((int)lane >= d) inc += t;
    }
    before = inc - cnt;
    total = __shfl(inc, 63, 64);
    if (cnt) groups[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  if (lane == 0) s_off[wave] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < COMPACT_BLOCK / 64; ++w) {
      const uint32_t x = s_off[w];
      s_off[w] = t;
      t += x;
    }
    s_base = t ? atomicAdd(&counters[0], (unsigned long long)t) : 0ull;
  }
  __syncthreads();
  if (!cnt) return;
  uint32_t slot = (uint32_t)((s_base + s_off[wave] + before) % npieces);
  if constexpr (G == 1) {
    list[slot] = i;
  } else {
    const uint32_t w[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if ((w[k >> 2] >> (8 * (k & 3))) & 0xffu) {
        list[slot] = i * 16u + (uint32_t)k;
        if (++slot == npieces) slot = 0;
      }
    }
  }
}

// One node of a listed piece (node index within its scene's grid).
// Accumulated state of one node, loaded ahead of its update.
struct NodeIn {
  unsigned long long M;
  float px, py, pz;
  bool pin;
};

__device__ __forceinline__ NodeIn load_node(size_t node, const unsigned long long* __restrict__ acc_m,
                                            const float* __restrict__ acc_p, size_t pstride,
                                            const uint8_t* __restrict__ fixed) {
  NodeIn a;
  a.M = acc_m[node];
  a.pin = fixed[node] != 0;
  a.px = acc_p[node];
  a.py = acc_p[pstride + node];
  a.pz = acc_p[2 * pstride + node];
  return a;
}

// n / d and n % d for d >= 1 and n / d < 2^22: the float quotient (three
// roundings, relative error below 3 2^-24) is within one of the true one,
// and one correction step makes it exact.
__device__ __forceinline__ uint32_t div_small(uint32_t n, uint32_t d, float inv_d, uint32_t& r) {
  int q = (int)((float)n * inv_d);
  int rem = (int)(n - (uint32_t)q * d);
  if (rem < 0) {
    --q;
    rem += (int)d;
  } else if (rem >= (int)d) {
    ++q;
    rem -= (int)d;
  }
  r = (uint32_t)rem;
  return (uint32_t)q;
}

__device__ __forceinline__ void update_node(size_t node, const NodeIn& a, unsigned long long* __restrict__ acc_m,
                                            float* __restrict__ acc_p, size_t pstride, uint8_t* __restrict__ fixed,
                                            float4* __restrict__ vel, const GridDims& g, const StepParams& sp,
                                            const int lo_idx[3], const int hi_idx[3], long long& Mq, double& p0,
                                            double& p1, double& p2) {
  Mq += (long long)a.M;
  p0 += a.px;
  p1 += a.py;
  p2 += a.pz;
  const float m = (float)a.M * sp.inv_mass_scale;
  float v[3] = {a.px, a.py, a.pz};
  if (m > g.mass_eps) {
    const float im = 1.f / m;
    const float gd = g.gravity_on_particles ? 0.f : sp.dt;
    v[0] = a.px * im + g.gravity[0] * gd;
    v[1] = a.py * im + g.gravity[1] * gd;
    v[2] = a.pz * im + g.gravity[2] * gd;
    // a scene's N^3 nodes fit 32 bits, and node / N < N^2 < 2^22 (N < 2048)
    const uint32_t N = (uint32_t)g.N, nd = (uint32_t)node;
    const float invN = 1.f / (float)N;
    uint32_t ix, iy, iz;
    const uint32_t row = div_small(nd, N, invN, ix);
    iz = div_small(row, N, invN, iy);
    const int idx[3] = {(int)ix, (int)iy, (int)iz};
    bool sticky = a.pin;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if ((g.wall[2 * k] == MPM_WALL_STICKY && idx[k] <= lo_idx[k]) ||
          (g.wall[2 * k + 1] == MPM_WALL_STICKY && idx[k] >= hi_idx[k]))
        sticky = true;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const bool lo_hit = idx[k] <= lo_idx[k] && (g.wall[2 * k] == MPM_WALL_SLIP || v[k] < 0.f);
      const bool hi_hit = idx[k] >= hi_idx[k] && (g.wall[2 * k + 1] == MPM_WALL_SLIP || v[k] > 0.f);
      if (sticky || lo_hit || hi_hit) v[k] = 0.f;
    }
  }
  // w carries the node mass, negative at a node a pinned particle holds (the
  // transfers read only xyz; grid views read the flag back).
  vel[node] = make_float4(v[0], v[1], v[2], a.pin ? -m : m);
  acc_m[node] = 0ull;
  acc_p[node] = 0.f;
  acc_p[pstride + node] = 0.f;
  acc_p[2 * pstride + node] = 0.f;
  fixed[node] = 0;
}

__device__ __forceinline__ void grid_node(size_t node, size_t Nn, unsigned long long* __restrict__ acc_m,
                                          float* __restrict__ acc_p, size_t pstride, uint8_t* __restrict__ fixed,
                                          float4* __restrict__ vel, const GridDims& g, const StepParams& sp,
                                          const int lo_idx[3], const int hi_idx[3], long long& Mq, double& p0,
                                          double& p1, double& p2) {
  if (node >= Nn) return;
  const NodeIn a = load_node(node, acc_m, acc_p, pstride, fixed);
  update_node(node, a, acc_m, acc_p, pstride, fixed, vel, g, sp, lo_idx, hi_idx, Mq, p0, p1, p2);
}

// Dense view of this step's grid (after k_grid_list): mass, velocity and the
// pinned flag of every node of the listed pieces. The outputs hold zeros
// elsewhere (the caller clears them). Used by callers that act on the grid
// between the update and G2P (src/three/mpm.js:104).
__global__ __launch_bounds__(BLOCK) void k_grid_export(const float4* __restrict__ vel,
                                                       const uint32_t* __restrict__ list, uint32_t npieces,
                                                       const unsigned long long* __restrict__ counters, size_t Nn,
                                                       float* __restrict__ mass, float* __restrict__ v,
                                                       uint8_t* __restrict__ fixed) {
  const unsigned long long start = counters[1], end = counters[2];
  const uint32_t W = gridDim.x * (BLOCK / PIECE_NODES);  // one piece per quarter-wave
  for (uint64_t e = blockIdx.x * (BLOCK / PIECE_NODES) + (threadIdx.x >> PIECE_SHIFT); e < end - start; e += W) {
    const size_t node = (size_t)list[(start + e) % npieces] * PIECE_NODES + (threadIdx.x & (PIECE_NODES - 1));
    if (node >= Nn) continue;
    const float4 u = vel[node];
    mass[node] = fabsf(u.w);
    fixed[node] = signbit(u.w) ? 1 : 0;
    v[3 * node] = u.x;
    v[3 * node + 1] = u.y;
    v[3 * node + 2] = u.z;
  }
}

// Velocities written back into the listed pieces (xyz; the mass and pinned
// flag are kept).
__global__ __launch_bounds__(BLOCK) void k_grid_import(float4* __restrict__ vel, const uint32_t* __restrict__ list,
                                                       uint32_t npieces,
                                                       const unsigned long long* __restrict__ counters, size_t Nn,
                                                       const float* __restrict__ v) {
  const unsigned long long start = counters[1], end = counters[2];
  const uint32_t W = gridDim.x * (BLOCK / PIECE_NODES);  // one piece per quarter-wave
  for (uint64_t e = blockIdx.x * (BLOCK / PIECE_NODES) + (threadIdx.x >> PIECE_SHIFT); e < end - start; e += W) {
    const size_t node = (size_t)list[(start + e) % npieces] * PIECE_NODES + (threadIdx.x & (PIECE_NODES - 1));
    if (node >= Nn) continue;
    float4 u = vel[node];
    u.x = v[3 * node];
    u.y = v[3 * node + 1];
    u.z = v[3 * node + 2];
    vel[node] = u;
  }
}

// Adds a wave's grid sums of one scene to that scene's slot.
__device__ __forceinline__ void flush_sums(double* __restrict__ sums, uint32_t scene, uint32_t slot, long long Mq,
                                           double p0, double p1, double p2) {
  double* s = sums + ((size_t)scene * SUM_SLOTS + slot) * 4;
  atomicAdd(reinterpret_cast<unsigned long long*>(s), (unsigned long long)Mq);
  atomicAdd(s + 1, p0);
  atomicAdd(s + 2, p1);
  atomicAdd(s + 3, p2);
}

// Sums of the lanes in `need` (a wave-uniform mask) flushed to their scenes'
// slots, one flush per distinct scene among them.
__device__ __forceinline__ void flush_lanes(double* __restrict__ sums, uint32_t slot, uint64_t need, uint32_t cur,
                                            long long Mq, double p0, double p1, double p2, uint32_t lane) {
  while (need) {
    const uint32_t sc = __shfl(cur, (int)__builtin_ctzll(need), 64);
    const bool in = ((need >> lane) & 1ull) && cur == sc;
    const long long m = wave_sum(in ? Mq : 0ll);
    const double a = wave_sum(in ? p0 : 0.0), b = wave_sum(in ? p1 : 0.0), c = wave_sum(in ? p2 : 0.0);
    if (lane == 0) flush_sums(sums, sc, slot, m, a, b, c);
    need &= ~__ballot(in);
  }
}

// Workgroup b updates a contiguous range of the step's list entries, at
// least four per wave: each quarter-wave updates one piece of 16 nodes, the
// waves taking every fourth group of four entries. Workgroups past the end
// exit at once. Neighbouring entries mostly belong to one scene, so each lane
// keeps running sums for its current scene; when a lane's scene changes the
// wave flushes the old sums of the changing lanes, grouped by scene. At the
// end the lanes of a wave combine sums of equal scenes, and the waves of a
// workgroup combine the last group of each before the atomics.
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_grid_list(unsigned long long* __restrict__ acc_m, float* __restrict__ acc_p,
                                                     uint8_t* __restrict__ fixed, float4* __restrict__ vel,
                                                     GridDims g1, StepParams sp1, int lo0, int lo1, int lo2, int hi0,
                                                     int hi1, int hi2, const SceneDesc* __restrict__ scenes,
                                                     const uint32_t* __restrict__ block_scene, uint32_t node_stride,
                                                     const uint32_t* __restrict__ list, uint32_t npieces,
                                                     unsigned long long* __restrict__ counters,
                                                     double* __restrict__ sums) {
  __shared__ double s_sum[BLOCK / 64][4];
  __shared__ uint32_t s_scene[BLOCK / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned long long start = counters[1], end = counters[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[2] = end;
  const uint32_t count = (uint32_t)(end - start);
  constexpr uint32_t WAVES = BLOCK / 64, QUARTERS = 64 / PIECE_NODES;
  const uint32_t per = max(QUARTERS * WAVES, (count + gridDim.x - 1) / gridDim.x);
  const uint32_t e0 = blockIdx.x * per;
  if (e0 >= count) return;  // the same for the whole workgroup
  const uint32_t e1 = min(count, e0 + per);
  uint32_t cur = 0xffffffffu;
  long long Mq = 0;
  double p0 = 0, p1 = 0, p2 = 0;
  // List slot of position start + e: the running position modulo the list
  // capacity, reduced once (e < count <= npieces).
  const uint32_t s0 = (uint32_t)(start % npieces);
  auto slot = [&](uint32_t e) {
    const uint32_t k = s0 + e;
    return k >= npieces ? k - npieces : k;
  };
  const uint32_t q = lane / PIECE_NODES, sub = lane & (PIECE_NODES - 1);
  if constexpr (BATCH) {
    for (uint32_t e4 = e0 + QUARTERS * wave; e4 < e1; e4 += QUARTERS * WAVES) {
      const uint32_t e = e4 + q;
      const bool valid = e < e1;
      uint32_t gp = 0, scene = cur;
      if (valid) {
        gp = list[slot(e)];  // piece index in the (concatenated) grid
        scene = block_scene[gp >> (8 - PIECE_SHIFT)];  // 256-node grid blocks
      }
      const uint64_t change = __ballot(scene != cur);
      if (change) {
        flush_lanes(sums, blockIdx.x % SUM_SLOTS, change & __ballot(cur != 0xffffffffu), cur, Mq, p0, p1, p2, lane);
        if (scene != cur) {
          cur = scene;
          Mq = 0;
          p0 = p1 = p2 = 0;
        }
      }
      // Valid lanes are a prefix of the wave, so lane 0's scene is theirs
      // when they share one (the common case): its descriptor then comes
      // through scalar registers instead of a vector load per lane.
      const uint32_t s0 = __builtin_amdgcn_readfirstlane(scene);
      if (__all(!valid || scene == s0)) {
        if (valid) {
          const SceneDesc* sd = scenes + s0;
          const GridDims g = to_scalar(sd->g);
          const StepParams sp = to_scalar(sd->sp);
          const uint32_t node0 = __builtin_amdgcn_readfirstlane(sd->node0);
          int lo[3], hi[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            lo[k] = __builtin_amdgcn_readfirstlane(sd->lo_idx[k]);
            hi[k] = __builtin_amdgcn_readfirstlane(sd->hi_idx[k]);
          }
          const size_t Nn = (size_t)g.N * g.N * g.N;
          grid_node((size_t)(gp - (node0 >> PIECE_SHIFT)) * PIECE_NODES + sub, Nn, acc_m + node0, acc_p + node0,
                    node_stride, fixed + node0, vel + node0, g, sp, lo, hi, Mq, p0, p1, p2);
        }
      } else if (valid) {
        const SceneDesc* sd = scenes + scene;
        const uint32_t node0 = sd->node0;
        const size_t Nn = (size_t)sd->g.N * sd->g.N * sd->g.N;
        grid_node((size_t)(gp - (node0 >> PIECE_SHIFT)) * PIECE_NODES + sub, Nn, acc_m + node0, acc_p + node0,
                  node_stride, fixed + node0, vel + node0, sd->g, sd->sp, sd->lo_idx, sd->hi_idx, Mq, p0, p1, p2);
      }
    }
  } else {
    cur = 0;
    const int lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    const size_t Nn = (size_t)g1.N * g1.N * g1.N;
    for (uint32_t e4 = e0 + QUARTERS * wave; e4 < e1; e4 += QUARTERS * WAVES) {
      const uint32_t e = e4 + q;
      if (e < e1)
        grid_node((size_t)list[slot(e)] * PIECE_NODES + sub, Nn, acc_m, acc_p, Nn, fixed, vel, g1, sp1, lo, hi, Mq,
                  p0, p1, p2);
    }
  }
  // Lanes group by scene; every group but the last flushes at once, the last
  // goes to the workgroup combine below.
  uint32_t wscene = 0xffffffffu;
  long long wm = 0;
  double wa = 0, wb = 0, wc = 0;
  uint64_t need = __ballot(cur != 0xffffffffu);
  while (need) {
    const uint32_t sc = __shfl(cur, (int)__builtin_ctzll(need), 64);
    const bool in = cur == sc;
    const long long m = wave_sum(in ? Mq : 0ll);
    const double a = wave_sum(in ? p0 : 0.0), b = wave_sum(in ? p1 : 0.0), c = wave_sum(in ? p2 : 0.0);
    need &= ~__ballot(in);
    if (need) {
      if (lane == 0) flush_sums(sums, sc, blockIdx.x % SUM_SLOTS, m, a, b, c);
    } else {
      wscene = sc;
      wm = m;
      wa = a;
      wb = b;
      wc = c;
    }
  }
  if (lane == 0) {
    s_scene[wave] = wscene;
    s_sum[wave][0] = __longlong_as_double(wm);
    s_sum[wave][1] = wa;
    s_sum[wave][2] = wb;
    s_sum[wave][3] = wc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Waves of one workgroup hold neighbouring list entries: combine runs of
    // equal scenes, one flush per run.
    uint32_t sc = 0xffffffffu;
    long long m = 0;
    double a = 0, b = 0, c = 0;
    for (uint32_t w = 0; w < BLOCK / 64; ++w) {
      if (s_scene[w] == 0xffffffffu) continue;
      if (s_scene[w] != sc) {
        if (sc != 0xffffffffu) flush_sums(sums, sc, blockIdx.x % SUM_SLOTS, m, a, b, c);
        sc = s_scene[w];
        m = 0;
        a = b = c = 0;
      }
      m += __double_as_longlong(s_sum[w][0]);
      a += s_sum[w][1];
      b += s_sum[w][2];
      c += s_sum[w][3];
    }
    if (sc != 0xffffffffu) flush_sums(sums, sc, blockIdx.x % SUM_SLOTS, m, a, b, c);
  }
}

// ------------------------------------------------------------ frame wrapper
// Device side of Simulation3.step's all-solid substep around MPM3.step
// (src/three/simulation.js:330-388) and of regularizeSolids
// (src/three/solid-sampling.js:5-42).

// Particle gravity of the gravity modes (src/three/gravity-options.js:7-16).
__device__ __forceinline__ void frame_gravity(const FrameDev& f, float x, float y, float z, float a[3]) {
  a[0] = a[1] = a[2] = 0.f;
  if (f.gmode == 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) a[k] = f.gdir[k] * f.gstrength * 24.f;
  } else if (f.gmode == 2 || f.gmode == 3) {
    const float d[3] = {f.gcenter[0] - x, f.gcenter[1] - y, f.gcenter[2] - z};
    const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + f.gsoft * f.gsoft;
    const float sc = (f.gmode == 3 ? -24000.f : 24000.f) * f.gstrength / (r2 * sqrtf(r2));
#pragma unroll
    for (int k = 0; k < 3; ++k) a[k] = d[k] * sc;
  }
}

// Before a substep: remember x; unpinned particles take gravity into v.
__global__ __launch_bounds__(BLOCK) void k_frame_pre(float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                     uint32_t cap, uint32_t n, float* __restrict__ ox, FrameDev f,
                                                     float dt) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const float x = plane(P, cap, P_X, i), y = plane(P, cap, P_Y, i), z = plane(P, cap, P_Z, i);
  ox[i] = x;
  ox[(size_t)cap + i] = y;
  ox[2 * (size_t)cap + i] = z;
  if ((meta[i] >> 8) & FLAG_PINNED) return;
  float a[3];
  frame_gravity(f, x, y, z, a);
#pragma unroll
  for (int k = 0; k < 3; ++k) set_plane(P, cap, P_VX + k, i, plane(P, cap, P_VX + k, i) + a[k] * dt);
}

// After the step: unpinned x clamped to the bounds, pinned x back to its
// anchor; v = clamp((x - x_start) / dt, +-limit).
__global__ __launch_bounds__(BLOCK) void k_frame_post(float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                      uint32_t cap, uint32_t n, const float* __restrict__ ox,
                                                      FrameDev f, float dt) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const bool pinned = (meta[i] >> 8) & FLAG_PINNED;
  const float inv = 1.f / dt;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float o = ox[(size_t)k * cap + i];
    const float x = pinned ? o : fminf(fmaxf(plane(P, cap, P_X + k, i), f.lo[k]), f.hi[k]);
    set_plane(P, cap, P_X + k, i, x);
    set_plane(P, cap, P_VX + k, i, fminf(fmaxf((x - o) * inv, -f.vlimit), f.vlimit));
  }
}

// First slot of every base cell in the sorted slot order, indexed by the
// cell's sort key (entries of cells with no particle are left stale; readers
// check them).
__global__ __launch_bounds__(BLOCK) void k_cell_starts(const uint32_t* __restrict__ keys, uint32_t n,
                                                       uint32_t* __restrict__ start) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) start[k] = i;
}

__device__ __forceinline__ float frame_spacing(const float* __restrict__ P, uint32_t cap, uint32_t i, float vol_unit) {
  float F[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = plane(P, cap, P_F0 + k, i);
  const float J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                  F[2] * (F[3] * F[7] - F[4] * F[6]);
  return 2.8f * cbrtf(fminf(fmaxf(plane(P, cap, P_VOL, i) * vol_unit * J, 0.015625f), 8.f));
}

// regularizeSolids as a gather: particle i sums the pushes of every particle
// j of the same material closer than the mean of their spacings, from the
// base cells within +-R of its own; a coincident pair uses the reference's
// hashed direction, oriented from the lower to the higher original index.
// d receives the displacement, maxbits the largest |d| (float bits).
__global__ __launch_bounds__(BLOCK) void k_regularize(const float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                      const uint32_t* __restrict__ ids,
                                                      const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ cstart, uint32_t cap, uint32_t n,
                                                      int N, int R, FrameDev f, float gain, float* __restrict__ d,
                                                      unsigned int* __restrict__ maxbits) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  float len = 0.f;
  if (i < n) {
    const uint32_t mi = meta[i];
    const float imi = ((mi >> 8) & FLAG_PINNED) ? 0.f : 1.f / plane(P, cap, P_MASS, i);
    const float xi = plane(P, cap, P_X, i), yi = plane(P, cap, P_Y, i), zi = plane(P, cap, P_Z, i);
    const float ri = frame_spacing(P, cap, i, f.vol_unit);
    const uint32_t oi = ids[i];
    int bx, by, bz;
    blocked_cell_decode(keys[i], N, bx, by, bz);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    for (int oz = -R; oz <= R; ++oz)
      for (int oy = -R; oy <= R; ++oy)
        for (int oxo = -R; oxo <= R; ++oxo) {
          const int cx = bx + oxo, cy = by + oy, cz = bz + oz;
          if (cx < 0 || cy < 0 || cz < 0 || cx > N - 3 || cy > N - 3 || cz > N - 3) continue;
          const uint32_t cc = blocked_cell_key(cx, cy, cz, N);
          uint32_t j = cstart[cc];
          if (j >= n || keys[j] != cc || (j > 0 && keys[j - 1] == cc)) continue;
          for (; j < n && keys[j] == cc; ++j) {
            if (j == i || (meta[j] & 255u) != (mi & 255u)) continue;
            const float imj = ((meta[j] >> 8) & FLAG_PINNED) ? 0.f : 1.f / plane(P, cap, P_MASS, j);
            const float total = imi + imj;
            if (total == 0.f) continue;
            float ex = xi - plane(P, cap, P_X, j), ey = yi - plane(P, cap, P_Y, j), ez = zi - plane(P, cap, P_Z, j);
            float r = sqrtf(ex * ex + ey * ey + ez * ez);
            const float target = 0.5f * (ri + frame_spacing(P, cap, j, f.vol_unit));
            if (r >= target) continue;
            float q;
            if (r < 1e-8f) {
              const uint32_t oj = ids[j];
              const uint32_t lo = oi < oj ? oi : oj, hi = oi < oj ? oj : oi;
              const uint32_t hash = ((lo + 1u) * 73856093u) ^ ((hi + 1u) * 19349663u);
              const float sgn = oi < oj ? 1.f : -1.f;
              ex = sgn * ((float)(hash & 1023u) / 511.5f - 1.f);
              ey = sgn * ((float)((hash >> 10) & 1023u) / 511.5f - 1.f);
              ez = sgn * ((float)((hash >> 20) & 1023u) / 511.5f - 1.f);
              r = sqrtf(ex * ex + ey * ey + ez * ez);
              q = target * gain / (r * total);
            } else {
              q = (target - r) * gain / (r * total);
            }
            dx += ex * q * imi;
            dy += ey * q * imi;
            dz += ez * q * imi;
          }
        }
    d[i] = dx;
    d[(size_t)cap + i] = dy;
    d[2 * (size_t)cap + i] = dz;
    len = sqrtf(dx * dx + dy * dy + dz * dz);
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) len = fmaxf(len, __shfl_down(len, s, 64));
  if ((threadIdx.x & 63u) == 0 && len > 0.f) atomicMax(maxbits, __float_as_uint(len));
}

// Applies the displacements scaled so the largest is at most `budget`.
__global__ __launch_bounds__(BLOCK) void k_regularize_apply(float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                            uint32_t cap, uint32_t n, const float* __restrict__ d,
                                                            const unsigned int* __restrict__ maxbits, FrameDev f,
                                                            float budget) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n || ((meta[i] >> 8) & FLAG_PINNED)) return;
  const float scale = fmaxf(1.f, __uint_as_float(*maxbits) / budget);
#pragma unroll
  for (int k = 0; k < 3; ++k)
    set_plane(P, cap, P_X + k, i,
              fminf(fmaxf(plane(P, cap, P_X + k, i) + d[(size_t)k * cap + i] / scale, f.lo[k]), f.hi[k]));
}

__global__ void k_zero_u32(unsigned int* p) { *p = 0u; }

__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Particle-side FP64 sums, one record of RED_K doubles per block (finished on
// the host): mass, fixed-point mass, momentum, angular momentum with the APIC
// spin term, kinetic and affine energy, min J, max |v|, non-finite count.
constexpr int RED_K = 14;
constexpr int RED_BLOCKS = 1024;

__device__ __forceinline__ double block_sum(double v, double* s) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s[wave] = v;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < BLOCK / 64; ++w) t += s[w];
  return t;
}
__device__ __forceinline__ double block_min(double v, double* s) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = fmin(v, __shfl_down(v, d, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s[wave] = v;
  __syncthreads();
  double t = s[0];
  for (int w = 1; w < BLOCK / 64; ++w) t = fmin(t, s[w]);
  return t;
}

__global__ __launch_bounds__(BLOCK) void k_particle_reduce(const float* __restrict__ P, uint32_t cap, uint32_t n,
                                                           float mass_scale, float D, double* __restrict__ out) {
  __shared__ double s[BLOCK / 64];
  double acc[RED_K];
#pragma unroll
  for (int k = 0; k < RED_K; ++k) acc[k] = 0;
  double minJ = 1e300, maxv = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    const double m = plane(P, cap, P_MASS, i);
    const double x[3] = {plane(P, cap, P_X, i), plane(P, cap, P_Y, i), plane(P, cap, P_Z, i)};
    const double v[3] = {plane(P, cap, P_VX, i), plane(P, cap, P_VY, i), plane(P, cap, P_VZ, i)};
    float C[9];
    M3 F;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      C[k] = plane(P, cap, P_C0 + k, i);
      F[k] = plane(P, cap, P_F0 + k, i);
    }
    acc[0] += m;
    acc[1] += (double)rintf((float)m * mass_scale);
    const double P3[3] = {m * v[0], m * v[1], m * v[2]};
    acc[2] += P3[0];
    acc[3] += P3[1];
    acc[4] += P3[2];
    acc[5] += x[1] * P3[2] - x[2] * P3[1] + m * D * (C[7] - C[5]);
    acc[6] += x[2] * P3[0] - x[0] * P3[2] + m * D * (C[2] - C[6]);
    acc[7] += x[0] * P3[1] - x[1] * P3[0] + m * D * (C[3] - C[1]);
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    acc[8] += 0.5 * m * v2;
    double cc = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) cc += (double)C[k] * C[k];
    acc[9] += 0.5 * m * D * cc;
    bool finite = isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]) && isfinite(v2);
#pragma unroll
    for (int k = 0; k < 9; ++k) finite = finite && isfinite(F[k]);
    acc[10] += finite ? 0.0 : 1.0;
    minJ = fmin(minJ, (double)det3(F));
    maxv = fmax(maxv, sqrt(v2));
  }
  double r[RED_K];
#pragma unroll
  for (int k = 0; k < 11; ++k) r[k] = block_sum(acc[k], s);
  r[11] = block_min(minJ, s);
  r[12] = -block_min(-maxv, s);
  r[13] = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < RED_K; ++k) out[blockIdx.x * RED_K + k] = r[k];
  }
}

// Largest wave speed |v| + c max(1, |F|_row |F|_col) and largest sum |C|
// (src/three/mpm.js:43-55); per block {max wave, max grad}.
__global__ __launch_bounds__(BLOCK) void k_stable(const float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                  uint32_t cap, uint32_t n, const DevMaterial* __restrict__ mats,
                                                  float* __restrict__ out) {
  __shared__ double s[BLOCK / 64];
  double mx = 1.0, grad = 1.0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    const DevMaterial mat = mats[meta[i] & 255u];
    float row = 0.f, col = 0.f, gsum = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float r = 0.f, c = 0.f;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        r += fabsf(plane(P, cap, P_F0 + a * 3 + b, i));
        c += fabsf(plane(P, cap, P_F0 + b * 3 + a, i));
        gsum += fabsf(plane(P, cap, P_C0 + a * 3 + b, i));
      }
      row = fmaxf(row, r);
      col = fmaxf(col, c);
    }
    const float rho = plane(P, cap, P_MASS, i) / plane(P, cap, P_VOL, i);
    const int model = effective_model(mat, mat.fr_on ? plane(P, cap, P_DMG, i) : 0.f);
    const float hard = model == 2 ? snow_hardening(plane(P, cap, P_JP, i)) : 1.f;
    const float c2 = (mat.lambda + 2.f * mat.mu) * hard / rho;
    // Neo-Hookean and the pressure models stiffen in compression:
    // 1 / max(.05, J), with J det F Jp for the pressure models.
    float comp = 1.f;
    if (model == 1 || model == 8 || model == 9) {
      M3 F;
#pragma unroll
      for (int k = 0; k < 9; ++k) F[k] = plane(P, cap, P_F0 + k, i);
      comp = 1.f / fmaxf(.05f, det3(F) * (model == 1 ? 1.f : plane(P, cap, P_JP, i)));
    }
    const float vx = plane(P, cap, P_VX, i), vy = plane(P, cap, P_VY, i), vz = plane(P, cap, P_VZ, i);
    const float speed = sqrtf(vx * vx + vy * vy + vz * vz);
    mx = fmax(mx, (double)(speed + sqrtf(c2) * fmaxf(fmaxf(1.f, row * col), comp)));
    grad = fmax(grad, (double)gsum);
  }
  const double a = -block_min(-mx, s);
  const double b = -block_min(-grad, s);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 2] = (float)a;
    out[blockIdx.x * 2 + 1] = (float)b;
  }
}

// Von Mises of this step's Kirchhoff stress over J (J Jp for the pressure
// models) and E: the diagnostic MPM3.stress (src/three/mpm.js:38). F is the
// slot's deformation gradient, `model` its effective model.
__device__ __forceinline__ float diag_stress(const float* __restrict__ P, uint32_t cap, uint32_t i, uint32_t flags,
                                             const DevMaterial& mat, int model, const M3& F) {
  float t[6];
  step_stress(F, model, mat, flags, P, cap, i, t);
  const float vm = sqrtf(fmaxf(0.f, 0.5f * ((t[0] - t[1]) * (t[0] - t[1]) + (t[1] - t[2]) * (t[1] - t[2]) +
                                            (t[2] - t[0]) * (t[2] - t[0])) +
                                         3.f * (t[3] * t[3] + t[4] * t[4] + t[5] * t[5])));
  const float J = det3(F) * ((model == 8 || model == 9) ? plane(P, cap, P_JP, i) : 1.f);
  return vm / fmaxf(1e-6f, J) / mat.E;
}

// The von Mises diagnostic per slot (out, may be null) and, when tau_out is
// set, the Kirchhoff stress the next step applies: the step stress times
// (1 - damage) below failure (src/three/mpm.js:36-37), 6 values per slot.
__global__ __launch_bounds__(BLOCK) void k_von_mises(const float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                     uint32_t cap, uint32_t n, const DevMaterial* __restrict__ mats,
                                                     float* __restrict__ out, float* __restrict__ tau_out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t mt = meta[i];
  const DevMaterial mat = mats[mt & 255u];
  M3 F;
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = plane(P, cap, P_F0 + k, i);
  const float dmg = mat.fr_on ? plane(P, cap, P_DMG, i) : 0.f;
  const int model = effective_model(mat, dmg);
  if (out) out[i] = diag_stress(P, cap, i, mt >> 8, mat, model, F);
  if (tau_out) {
    float t[6];
    step_stress(F, model, mat, mt >> 8, P, cap, i, t);
    const float f = dmg < 1.f ? 1.f - dmg : 1.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) tau_out[(size_t)i * 6 + k] = t[k] * f;
  }
}

// Preview pick of one particle per block of `stride` consecutive original
// ids: the PCG RXS-M-XS output of one LCG step of the block index, modulo the
// block's length (gpu/solver.py:990-996).
__device__ __forceinline__ uint32_t preview_pick(uint32_t row, uint32_t stride, uint32_t n) {
  uint32_t h = row * 747796405u + 2891336453u;
  h = ((h >> ((h >> 28) + 4u)) ^ h) * 277803737u;
  h = (h >> 22) ^ h;
  const uint32_t span = min(stride, n - row * stride);
  return h % span;
}

constexpr int SNAP_K = 6;  // mass, min T, max T, min J, max stress, max plastic

// Frame snapshot (gpu/solver.py:986-1016): every slot whose original id is
// its block's pick writes its row (x, y, z, material code, T, pinned,
// diagnostic stress, plastic strain) -- one writer per row; every slot adds
// to the block totals (FP64) and to the per-model and non-finite counters.
__global__ __launch_bounds__(BLOCK) void k_snapshot(const float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                    const uint32_t* __restrict__ ids, uint32_t cap, uint32_t n,
                                                    const DevMaterial* __restrict__ mats,
                                                    const int32_t* __restrict__ codes, uint32_t stride,
                                                    float* __restrict__ data, double* __restrict__ partial,
                                                    unsigned long long* __restrict__ counts) {
  __shared__ double s[BLOCK / 64];
  __shared__ unsigned int s_counts[13];
  if (threadIdx.x < 13) s_counts[threadIdx.x] = 0u;
  __syncthreads();
  double mass = 0, tmin = 1e30, tmax = -1e30, jmin = 1, smax = 0, pmax = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    const uint32_t mt = meta[i];
    const uint32_t mi = mt & 255u;
    const DevMaterial mat = mats[mi];
    const int model = effective_model(mat, mat.fr_on ? plane(P, cap, P_DMG, i) : 0.f);
    M3 F;
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = plane(P, cap, P_F0 + k, i);
    const float stress = diag_stress(P, cap, i, mt >> 8, mat, model, F);
    const float x = plane(P, cap, P_X, i), y = plane(P, cap, P_Y, i), z = plane(P, cap, P_Z, i);
    const float vx = plane(P, cap, P_VX, i), vy = plane(P, cap, P_VY, i), vz = plane(P, cap, P_VZ, i);
    const float T = plane(P, cap, P_TEMP, i), ep = plane(P, cap, P_EP, i);
    const uint32_t id = ids[i];
    const uint32_t row = id / stride;
    if (data && id - row * stride == preview_pick(row, stride, n)) {
      const int32_t c = codes[mi];
      float* d = data + 8 * (size_t)row;
      d[0] = x;
      d[1] = y;
      d[2] = z;
      d[3] = (float)(c == 11 ? 11 + model : c);
      d[4] = T;
      d[5] = ((mt >> 8) & FLAG_PINNED) ? 1.f : 0.f;
      d[6] = stress;
      d[7] = ep;
    }
    const double x2 = (double)x * x + (double)y * y + (double)z * z;
    const double v2 = (double)vx * vx + (double)vy * vy + (double)vz * vz;
    if (!(x2 < 1e20 && v2 < 1e20)) atomicAdd(&s_counts[12], 1u);
    atomicAdd(&s_counts[model], 1u);
    mass += plane(P, cap, P_MASS, i);
    tmin = fmin(tmin, (double)T);
    tmax = fmax(tmax, (double)T);
    jmin = fmin(jmin, (double)det3(F));
    smax = fmax(smax, (double)stress);
    pmax = fmax(pmax, (double)ep);
  }
  double r[SNAP_K];
  r[0] = block_sum(mass, s);
  r[1] = block_min(tmin, s);
  r[2] = -block_min(-tmax, s);
  r[3] = block_min(jmin, s);
  r[4] = -block_min(-smax, s);
  r[5] = -block_min(-pmax, s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < SNAP_K; ++k) partial[blockIdx.x * SNAP_K + k] = r[k];
  }
  __syncthreads();
  if (threadIdx.x < 13 && s_counts[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)s_counts[threadIdx.x]);
}

// Plain constitutive update of a batch of deformation gradients: the return
// map the step runs after G2P (models >= 2) or, for models 0 and 1, the
// elastic stress; tau out as the symmetric 6-vector. jp / dil (Jp and
// dilation in and out) may be null: Jp = 1, no dilation.
__global__ __launch_bounds__(BLOCK) void k_return_batch(DevMaterial mat, uint32_t n, float dt,
                                                        const float* __restrict__ F, const float* __restrict__ ep,
                                                        const float* __restrict__ T, const float* jp_in,
                                                        const float* dil_in, float* __restrict__ F_out,
                                                        float* __restrict__ ep_out, float* jp_out, float* dil_out,
                                                        float* __restrict__ tau_out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  M3 f;
#pragma unroll
  for (int k = 0; k < 9; ++k) f[k] = F[(size_t)i * 9 + k];
  float tau[6];
  float e = ep[i];
  float jp = jp_in ? jp_in[i] : 1.f, dil = dil_in ? dil_in[i] : 0.f;
  if (mat.model >= 2) {
    e += plastic_project(f, mat.model, e, jp, dil, dt, T[i], mat, tau);
  } else {
    elastic_tau(f, mat, 1.f, tau);
  }
  if (jp_out) jp_out[i] = jp;
  if (dil_out) dil_out[i] = dil;
#pragma unroll
  for (int k = 0; k < 9; ++k) F_out[(size_t)i * 9 + k] = f[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) tau_out[(size_t)i * 6 + k] = tau[k];
  ep_out[i] = e;
}

// Step stress of a batch of F without a cached return: fixed-corotated for
// model 0, Hencky from F's own decomposition otherwise.
__global__ __launch_bounds__(BLOCK) void k_stress_batch(DevMaterial mat, uint32_t n, const float* __restrict__ F,
                                                        float* __restrict__ tau_out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  M3 f;
#pragma unroll
  for (int k = 0; k < 9; ++k) f[k] = F[(size_t)i * 9 + k];
  float tau[6];
  elastic_tau(f, mat, 1.f, tau);
#pragma unroll
  for (int k = 0; k < 6; ++k) tau_out[(size_t)i * 6 + k] = tau[k];
}

inline unsigned blocks(size_t n, int b = BLOCK) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int particle_reduce_k() { return RED_K; }
uint32_t grid_pieces(const GridDims& g) { return pieces_of_nodes((size_t)g.N * g.N * g.N); }
int particle_reduce_blocks() { return RED_BLOCKS; }

void launch_bin_keys(hipStream_t s, const DevState& d, uint32_t n, const GridDims& g) {
  if (!n) return;
  hipLaunchKernelGGL(k_bin_keys, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.ids, d.cap, n, g, d.keys, d.vals);
}

void launch_permute(hipStream_t s, DevState& d, uint32_t n, const uint32_t* order) {
  if (!n) return;
  hipLaunchKernelGGL(k_permute, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.planes_alt, d.ids, d.ids_alt, d.meta,
                     d.meta_alt, d.cap, n, order);
  std::swap(d.planes, d.planes_alt);
  std::swap(d.ids, d.ids_alt);
  std::swap(d.meta, d.meta_alt);
}

void launch_compact(hipStream_t s, const DevState& d, uint32_t npieces, uint32_t nscenes, LaunchEvents ev) {
  // A ballot per 64 flags up to 2^17 pieces (at most 128 counter atomics);
  // 16 flags per lane above (a 128^3-node grid has 2^17 pieces).
  constexpr uint32_t CB = COMPACT_BLOCK;
  if (npieces <= (1u << 17)) {
    hipExtLaunchKernelGGL(k_compact_pieces<1>, dim3(std::max<uint32_t>(1, (npieces + CB - 1) / CB)), dim3(CB), 0, s,
                          ev.start, ev.stop, 0, d.pieces, npieces, d.piece_list, d.piece_counters, d.red,
                          nscenes * SUM_SLOTS * 4);
  } else {
    const uint32_t groups = (npieces + 15) / 16;
    hipExtLaunchKernelGGL(k_compact_pieces<16>, dim3((groups + CB - 1) / CB), dim3(CB), 0, s, ev.start, ev.stop, 0,
                          d.pieces, npieces, d.piece_list, d.piece_counters, d.red, nscenes * SUM_SLOTS * 4);
  }
}

void launch_grid(hipStream_t s, const DevState& d, const GridDims& g, const StepParams& p, const int lo_idx[3],
                 const int hi_idx[3], uint32_t blocks, LaunchEvents ev) {
  hipExtLaunchKernelGGL(k_grid_list<false>, dim3(blocks), dim3(BLOCK), 0, s, ev.start, ev.stop, 0, d.acc_m, d.acc_p,
                        d.fixed, d.vel, g, p, lo_idx[0], lo_idx[1], lo_idx[2], hi_idx[0], hi_idx[1], hi_idx[2],
                        (const SceneDesc*)nullptr, (const uint32_t*)nullptr, 0u, d.piece_list, grid_pieces(g),
                        d.piece_counters, d.red);
}

void launch_bin_keys_batch(hipStream_t s, const DevState& d, const BatchArgs& b) {
  // One 256-thread block per 256-slot binning chunk.
  constexpr uint32_t chunk = 256;
  hipLaunchKernelGGL(k_bin_keys_batch, dim3(b.nchunks * (chunk / BLOCK)), dim3(BLOCK), 0, s, d.planes, d.ids, d.cap,
                     b, chunk, d.keys, d.vals);
}

void launch_grid_batch(hipStream_t s, const DevState& d, const SceneDesc* scenes, const uint32_t* block_scene,
                       uint32_t node_stride, uint32_t blocks, LaunchEvents ev) {
  hipExtLaunchKernelGGL(k_grid_list<true>, dim3(blocks), dim3(BLOCK), 0, s, ev.start, ev.stop, 0, d.acc_m, d.acc_p,
                        d.fixed, d.vel, GridDims{}, StepParams{}, 0, 0, 0, 0, 0, 0, scenes, block_scene, node_stride,
                        d.piece_list, pieces_of_nodes(node_stride), d.piece_counters, d.red);
}

void launch_frame_pre(hipStream_t s, const DevState& d, uint32_t n, float* ox, const FrameDev& f, float dt) {
  if (n) hipLaunchKernelGGL(k_frame_pre, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, ox, f, dt);
}

void launch_frame_post(hipStream_t s, const DevState& d, uint32_t n, const float* ox, const FrameDev& f, float dt) {
  if (n) hipLaunchKernelGGL(k_frame_post, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, ox, f, dt);
}

void launch_regularize(hipStream_t s, const DevState& d, uint32_t n, const GridDims& g, int R, const FrameDev& f,
                       float dt, uint32_t* cell_start, float* disp, unsigned int* maxbits) {
  if (!n) return;
  hipLaunchKernelGGL(k_cell_starts, dim3(blocks(n)), dim3(BLOCK), 0, s, d.keys, n, cell_start);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, maxbits);
  const float gain = 1.f - std::exp(-40.f * dt);
  hipLaunchKernelGGL(k_regularize, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.ids, d.keys, cell_start,
                     d.cap, n, g.N, R, f, gain, disp, maxbits);
  hipLaunchKernelGGL(k_regularize_apply, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, disp,
                     maxbits, f, 8.f * dt);
}

void launch_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks(n)), dim3(BLOCK), 0, s, p, n, v);
}

void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_zero_words, dim3(blocks(n)), dim3(BLOCK), 0, s, p, n);
}

void launch_particle_reduce(hipStream_t s, const DevState& d, uint32_t n, float mass_scale, float D, double* out) {
  hipLaunchKernelGGL(k_particle_reduce, dim3(RED_BLOCKS), dim3(BLOCK), 0, s, d.planes, d.cap, n, mass_scale, D, out);
}

void launch_stable(hipStream_t s, const DevState& d, uint32_t n, float* out) {
  hipLaunchKernelGGL(k_stable, dim3(RED_BLOCKS), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, d.materials, out);
}

void launch_return_batch(hipStream_t s, const DevMaterial& m, uint32_t n, float dt, const float* F, const float* ep,
                         const float* T, const float* jp, const float* dil, float* F_out, float* ep_out,
                         float* jp_out, float* dil_out, float* tau_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_return_batch, dim3(blocks(n)), dim3(BLOCK), 0, s, m, n, dt, F, ep, T, jp, dil, F_out, ep_out,
                     jp_out, dil_out, tau_out);
}

void launch_stress_batch(hipStream_t s, const DevMaterial& m, uint32_t n, const float* F, float* tau_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_stress_batch, dim3(blocks(n)), dim3(BLOCK), 0, s, m, n, F, tau_out);
}

uint32_t snapshot_blocks(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(RED_BLOCKS, blocks(n))); }

void launch_snapshot(hipStream_t s, const DevState& d, uint32_t n, const int32_t* codes, uint32_t stride, float* data,
                     double* partial, unsigned long long* counts) {
  hipLaunchKernelGGL(k_snapshot, dim3(snapshot_blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.ids, d.cap, n,
                     d.materials, codes, stride, data, partial, counts);
}

void launch_grid_export(hipStream_t s, const DevState& d, const GridDims& g, float* mass, float* v, uint8_t* fixed) {
  const size_t Nn = (size_t)g.N * g.N * g.N;
  hipLaunchKernelGGL(k_grid_export, dim3(1024), dim3(BLOCK), 0, s, d.vel, d.piece_list, grid_pieces(g),
                     d.piece_counters, Nn, mass, v, fixed);
}

void launch_grid_import(hipStream_t s, const DevState& d, const GridDims& g, const float* v) {
  const size_t Nn = (size_t)g.N * g.N * g.N;
  hipLaunchKernelGGL(k_grid_import, dim3(1024), dim3(BLOCK), 0, s, d.vel, d.piece_list, grid_pieces(g),
                     d.piece_counters, Nn, v);
}

void launch_von_mises(hipStream_t s, const DevState& d, uint32_t n, float* out, float* tau_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_von_mises, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, d.materials, out,
                     tau_out);
}

}  // namespace mpm
