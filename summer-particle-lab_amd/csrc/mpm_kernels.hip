// MLS-MPM step kernels for gfx950 (FP32 state, FP64 conservation sums).
//
// Grid update, binning, permutation, reductions and batch constitutive
// kernels. The particle transfers are in mpm_fused.hip.
//   k_grid  v = p/m, gravity, wall rules (src/three/mpm.js:89-102)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_internal.h"

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

// Cell key of every slot, written at the slot's original particle id so the
// stable sort breaks ties by ascending id (the reference's P2G visiting order,
// src/three/mpm.js:66); the value is the slot the particle lives in.
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
  keys[id] = ((uint32_t)bz * (uint32_t)g.N + (uint32_t)by) * (uint32_t)g.N + (uint32_t)bx;
  vals[id] = i;
}

// Batched keys: chunk -> scene; the key is the base cell's node index in the
// concatenated grid (node0 + cell), so one sort orders scenes by node0 and
// cells within each scene.
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
  keys[id] = sd->node0 + ((uint32_t)bz * (uint32_t)g.N + (uint32_t)by) * (uint32_t)g.N + (uint32_t)bx;
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

// Kirchhoff stress used by this step's P2G (6-vector).
__device__ __forceinline__ void step_stress(const M3& F, const DevMaterial& m, uint32_t flags,
                                            const float* __restrict__ P, uint32_t cap, uint32_t i, float tau[6]) {
  if (m.model < 2) {
    fixed_corotated_tau(F, m.mu, m.lambda, tau);
  } else if (flags & FLAG_CACHE) {
#pragma unroll
    for (int k = 0; k < 6; ++k) tau[k] = plane(P, cap, P_T0 + k, i);
  } else {
    hencky_tau(F, m, tau);
  }
}

// One wave per 64-node piece of the dense grid; waves of inactive pieces do
// no memory work. Each block writes one record of FP64 grid sums (fixed-point
// mass, momentum) so the sums need no same-address atomics. `blk` is the
// block's index within its scene's grid, `out` its record; the momentum
// planes are `pstride` floats apart.
__device__ __forceinline__ void grid_block(uint32_t blk, uint32_t out, unsigned long long* __restrict__ acc_m,
                                           float* __restrict__ acc_p, size_t pstride, uint8_t* __restrict__ fixed,
                                           uint8_t* __restrict__ pieces, float4* __restrict__ vel, const GridDims& g,
                                           const StepParams& sp, const int lo_idx[3], const int hi_idx[3],
                                           double* __restrict__ partials) {
  __shared__ double s_red[BLOCK / 64][4];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const size_t piece = (size_t)blk * (BLOCK / 64) + wave;
  const size_t Nn = (size_t)g.N * g.N * g.N;
  const size_t npieces = (Nn + 63) / 64;
  const bool active = piece < npieces && pieces[piece] != 0;
  const size_t node = piece * 64 + lane;
  long long Mq = 0;
  double p0 = 0, p1 = 0, p2 = 0;
  if (active && node < Nn) {
    const unsigned long long M = acc_m[node];
    Mq = (long long)M;
    const float px = acc_p[node], py = acc_p[pstride + node], pz = acc_p[2 * pstride + node];
    p0 = px;
    p1 = py;
    p2 = pz;
    const float m = (float)M * sp.inv_mass_scale;
    float v[3] = {px, py, pz};
    if (m > g.mass_eps) {
      const float im = 1.f / m;
      const float gd = g.gravity_on_particles ? 0.f : sp.dt;
      v[0] = px * im + g.gravity[0] * gd;
      v[1] = py * im + g.gravity[1] * gd;
      v[2] = pz * im + g.gravity[2] * gd;
      const int N = g.N;
      const int ix = (int)(node % N), iy = (int)((node / N) % N), iz = (int)(node / ((size_t)N * N));
      const int idx[3] = {ix, iy, iz};
      bool sticky = fixed[node] != 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if ((g.wall[2 * a] == MPM_WALL_STICKY && idx[a] <= lo_idx[a]) ||
            (g.wall[2 * a + 1] == MPM_WALL_STICKY && idx[a] >= hi_idx[a]))
          sticky = true;
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const bool lo_hit = idx[a] <= lo_idx[a] && (g.wall[2 * a] == MPM_WALL_SLIP || v[a] < 0.f);
        const bool hi_hit = idx[a] >= hi_idx[a] && (g.wall[2 * a + 1] == MPM_WALL_SLIP || v[a] > 0.f);
        if (sticky || lo_hit || hi_hit) v[a] = 0.f;
      }
    }
    vel[node] = make_float4(v[0], v[1], v[2], m);
    acc_m[node] = 0ull;
    acc_p[node] = 0.f;
    acc_p[pstride + node] = 0.f;
    acc_p[2 * pstride + node] = 0.f;
    fixed[node] = 0;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    Mq += __shfl_down(Mq, d, 64);
    p0 += __shfl_down(p0, d, 64);
    p1 += __shfl_down(p1, d, 64);
    p2 += __shfl_down(p2, d, 64);
  }
  if (lane == 0) {
    s_red[wave][0] = __longlong_as_double(Mq);
    s_red[wave][1] = p0;
    s_red[wave][2] = p1;
    s_red[wave][3] = p2;
    if (active) pieces[piece] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long m = 0;
    double q[3] = {0, 0, 0};
    for (int w = 0; w < BLOCK / 64; ++w) {
      m += __double_as_longlong(s_red[w][0]);
      for (int c = 0; c < 3; ++c) q[c] += s_red[w][1 + c];
    }
    partials[(size_t)out * 4 + 0] = __longlong_as_double(m);
    for (int c = 0; c < 3; ++c) partials[(size_t)out * 4 + 1 + c] = q[c];
  }
}

__global__ __launch_bounds__(BLOCK) void k_grid(unsigned long long* __restrict__ acc_m, float* __restrict__ acc_p,
                                                uint8_t* __restrict__ fixed, uint8_t* __restrict__ pieces,
                                                float4* __restrict__ vel, GridDims g, StepParams sp,
                                                int lo_idx0, int lo_idx1, int lo_idx2, int hi_idx0, int hi_idx1,
                                                int hi_idx2, double* __restrict__ partials) {
  const int lo_idx[3] = {lo_idx0, lo_idx1, lo_idx2};
  const int hi_idx[3] = {hi_idx0, hi_idx1, hi_idx2};
  grid_block(blockIdx.x, blockIdx.x, acc_m, acc_p, (size_t)g.N * g.N * g.N, fixed, pieces, vel, g, sp, lo_idx, hi_idx,
             partials);
}

// Grid update of many scenes in one launch: block -> scene from a table; the
// scenes' grids are concatenated (each starts on a block boundary).
__global__ __launch_bounds__(BLOCK) void k_grid_batch(unsigned long long* __restrict__ acc_m,
                                                      float* __restrict__ acc_p, uint8_t* __restrict__ fixed,
                                                      uint8_t* __restrict__ pieces, float4* __restrict__ vel,
                                                      const SceneDesc* __restrict__ scenes,
                                                      const uint32_t* __restrict__ block_scene, uint32_t node_stride,
                                                      double* __restrict__ partials) {
  const SceneDesc* sd = scenes + block_scene[blockIdx.x];
  const uint32_t node0 = sd->node0;
  grid_block(blockIdx.x - sd->block0, blockIdx.x, acc_m + node0, acc_p + node0, node_stride, fixed + node0,
             pieces + (node0 >> 6), vel + node0, sd->g, sd->sp, sd->lo_idx, sd->hi_idx, partials);
}

__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
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
    const float c2 = (mat.lambda + 2.f * mat.mu) / rho;
    const float vx = plane(P, cap, P_VX, i), vy = plane(P, cap, P_VY, i), vz = plane(P, cap, P_VZ, i);
    const float speed = sqrtf(vx * vx + vy * vy + vz * vz);
    mx = fmax(mx, (double)(speed + sqrtf(c2) * fmaxf(1.f, row * col)));
    grad = fmax(grad, (double)gsum);
  }
  const double a = -block_min(-mx, s);
  const double b = -block_min(-grad, s);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 2] = (float)a;
    out[blockIdx.x * 2 + 1] = (float)b;
  }
}

// Von Mises of this step's Kirchhoff stress over J and E
// (src/three/mpm.js:38), per slot.
__global__ __launch_bounds__(BLOCK) void k_von_mises(const float* __restrict__ P, const uint32_t* __restrict__ meta,
                                                     uint32_t cap, uint32_t n, const DevMaterial* __restrict__ mats,
                                                     float* __restrict__ out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t mt = meta[i];
  const DevMaterial mat = mats[mt & 255u];
  M3 F;
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = plane(P, cap, P_F0 + k, i);
  float t[6];
  step_stress(F, mat, mt >> 8, P, cap, i, t);
  const float vm = sqrtf(fmaxf(0.f, 0.5f * ((t[0] - t[1]) * (t[0] - t[1]) + (t[1] - t[2]) * (t[1] - t[2]) +
                                            (t[2] - t[0]) * (t[2] - t[0])) +
                                         3.f * (t[3] * t[3] + t[4] * t[4] + t[5] * t[5])));
  out[i] = vm / fmaxf(1e-6f, det3(F)) / mat.E;
}

// Plain constitutive update of a batch of deformation gradients: the return
// map the step runs after G2P (models >= 2) or, for model 0, the
// fixed-corotated stress; tau out as the symmetric 6-vector.
__global__ __launch_bounds__(BLOCK) void k_return_batch(DevMaterial mat, uint32_t n, float dt,
                                                        const float* __restrict__ F, const float* __restrict__ ep,
                                                        const float* __restrict__ T, float* __restrict__ F_out,
                                                        float* __restrict__ ep_out, float* __restrict__ tau_out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  M3 f;
#pragma unroll
  for (int k = 0; k < 9; ++k) f[k] = F[(size_t)i * 9 + k];
  float tau[6];
  float e = ep[i];
  if (mat.model >= 2) {
    e += plastic_project(f, e, dt, T[i], mat, tau);
  } else {
    fixed_corotated_tau(f, mat.mu, mat.lambda, tau);
  }
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
  if (mat.model < 2)
    fixed_corotated_tau(f, mat.mu, mat.lambda, tau);
  else
    hencky_tau(f, mat, tau);
#pragma unroll
  for (int k = 0; k < 6; ++k) tau_out[(size_t)i * 6 + k] = tau[k];
}

inline unsigned blocks(size_t n, int b = BLOCK) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int particle_reduce_k() { return RED_K; }
uint32_t grid_blocks(const GridDims& g) {
  const size_t Nn = (size_t)g.N * g.N * g.N;
  return (uint32_t)(((Nn + 63) / 64 + BLOCK / 64 - 1) / (BLOCK / 64));
}
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

void launch_grid(hipStream_t s, const DevState& d, const GridDims& g, const StepParams& p, const int lo_idx[3],
                 const int hi_idx[3], LaunchEvents ev) {
  hipExtLaunchKernelGGL(k_grid, dim3(grid_blocks(g)), dim3(BLOCK), 0, s, ev.start, ev.stop, 0, d.acc_m, d.acc_p,
                        d.fixed, d.pieces, d.vel, g, p, lo_idx[0], lo_idx[1], lo_idx[2], hi_idx[0], hi_idx[1],
                        hi_idx[2], d.red);
}

void launch_bin_keys_batch(hipStream_t s, const DevState& d, const BatchArgs& b) {
  // Chunks of the batch binning pass are BLOCK-aligned slices of the fused
  // chunks: chunk_size is the largest fused chunk (256 slots).
  constexpr uint32_t chunk = 256;
  hipLaunchKernelGGL(k_bin_keys_batch, dim3(b.nchunks * (chunk / BLOCK)), dim3(BLOCK), 0, s, d.planes, d.ids, d.cap,
                     b, chunk, d.keys, d.vals);
}

void launch_grid_batch(hipStream_t s, const DevState& d, const SceneDesc* scenes, const uint32_t* block_scene,
                       uint32_t nblocks, uint32_t node_stride, LaunchEvents ev) {
  hipExtLaunchKernelGGL(k_grid_batch, dim3(nblocks), dim3(BLOCK), 0, s, ev.start, ev.stop, 0, d.acc_m, d.acc_p,
                        d.fixed, d.pieces, d.vel, scenes, block_scene, node_stride, d.red);
}

void launch_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks(n)), dim3(BLOCK), 0, s, p, n, v);
}

void launch_particle_reduce(hipStream_t s, const DevState& d, uint32_t n, float mass_scale, float D, double* out) {
  hipLaunchKernelGGL(k_particle_reduce, dim3(RED_BLOCKS), dim3(BLOCK), 0, s, d.planes, d.cap, n, mass_scale, D, out);
}

void launch_stable(hipStream_t s, const DevState& d, uint32_t n, float* out) {
  hipLaunchKernelGGL(k_stable, dim3(RED_BLOCKS), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, d.materials, out);
}

void launch_return_batch(hipStream_t s, const DevMaterial& m, uint32_t n, float dt, const float* F, const float* ep,
                         const float* T, float* F_out, float* ep_out, float* tau_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_return_batch, dim3(blocks(n)), dim3(BLOCK), 0, s, m, n, dt, F, ep, T, F_out, ep_out, tau_out);
}

void launch_stress_batch(hipStream_t s, const DevMaterial& m, uint32_t n, const float* F, float* tau_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_stress_batch, dim3(blocks(n)), dim3(BLOCK), 0, s, m, n, F, tau_out);
}

void launch_von_mises(hipStream_t s, const DevState& d, uint32_t n, float* out) {
  if (!n) return;
  hipLaunchKernelGGL(k_von_mises, dim3(blocks(n)), dim3(BLOCK), 0, s, d.planes, d.meta, d.cap, n, d.materials, out);
}

}  // namespace mpm
