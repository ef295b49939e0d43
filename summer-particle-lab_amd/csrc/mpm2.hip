// 2D elastic MLS-MPM for gfx950: the reference's SolidMPM (src/solid-mpm.js)
// on a fixed 149 x 149 node grid (cell 8, origin -592), with its frame
// driver's all-solid substep loop (src/simulation.js:240-300) run on the
// device.
//
// Particles stay in the caller's order in FP32 planes (the reference stores
// x and v as float32 too). P2G adds fixed-point node masses (exact totals)
// and FP32 momenta with global atomics: the grid is 22,201 nodes, small
// enough to stay in L2. The frame loop keeps the remaining time, the
// substep size and the count in device memory, so a frame's substeps are
// enqueued without a host round trip; the host reads the count once per
// chunk of substeps.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpm_engine.h"
#include "mpm_wave.h"

namespace {

using mpm::wave_reduce;
using mpm::wave_sum;

constexpr int BLOCK = 256;
constexpr int N2 = 149;  // nodes per axis (src/solid-mpm.js:4)
constexpr float DX2 = 8.f, ORIGIN2 = -592.f, AREA2 = 12.96f;
constexpr int NODES2 = N2 * N2;
constexpr int RED2_BLOCKS = 256;
constexpr int CHUNK = 8;  // substeps enqueued per host check in a frame

enum Plane2 : int { Q_X = 0, Q_Y, Q_VX, Q_VY, Q_F0, Q_C0 = Q_F0 + 4, Q_MASS = Q_C0 + 4, Q_VOL, NPLANES2 };

struct Params2 {
  float stiffness;
  int walls;
  float bound, floor, top;
  float mass_scale, inv_mass_scale;
};

// Frame loop state in device memory.
struct Loop2 {
  double remaining;   // time left in the frame
  double h;           // this substep (0: nothing left to do)
  int32_t steps, max_steps;
  uint32_t wave_bits, grad_bits;  // stable-step maxima of the current state
};

__device__ __forceinline__ float ld2(const float* P, uint32_t cap, int k, uint32_t i) { return P[(size_t)k * cap + i]; }
__device__ __forceinline__ void st2(float* P, uint32_t cap, int k, uint32_t i, float v) { P[(size_t)k * cap + i] = v; }

// Quadratic B-spline stencil of one axis (src/solid-mpm.js:31-37).
__device__ __forceinline__ int stencil_axis(float x, float& f, float w[3]) {
  const float p = (x - ORIGIN2) * (1.f / DX2);
  int b = (int)floorf(p - .5f);
  b = b < 0 ? 0 : (b > N2 - 3 ? N2 - 3 : b);
  f = p - (float)b;
  const float a = 1.5f - f, c = f - 1.f, d = f - .5f;
  w[0] = .5f * a * a;
  w[1] = .75f - c * c;
  w[2] = .5f * d * d;
  return b;
}

// 2D fixed-corotated Kirchhoff stress with the closed-form polar rotation
// (src/solid-mpm.js:10-18); t = (t00, t01, t10, t11).
__device__ __forceinline__ void solid_stress(float a, float b, float c, float d, float mu, float lambda, float t[4]) {
  const float J = a * d - b * c, q = sqrtf((a + d) * (a + d) + (c - b) * (c - b));
  const float iq = q > 1e-12f ? 1.f / q : 0.f;
  const float r = q > 1e-12f ? (a + d) * iq : 1.f, s = q > 1e-12f ? (c - b) * iq : 0.f;
  const float vol = lambda * (J - 1.f) * J;
  t[0] = 2.f * mu * ((a - r) * a + (b + s) * b) + vol;
  t[1] = 2.f * mu * ((a - r) * c + (b + s) * d);
  t[2] = 2.f * mu * ((c - s) * a + (d - r) * b);
  t[3] = 2.f * mu * ((c - s) * c + (d - r) * d) + vol;
}

// Substep size: the fixed dt, or the frame loop's (0 ends the kernel).
__device__ __forceinline__ float step_dt(float dt, const Loop2* loop) { return loop ? (float)loop->h : dt; }

// P2G (src/solid-mpm.js:51-66): mass as signed fixed-point parts whose
// per-particle sum is exactly the particle's quantum (the centre node takes
// the remainder; a particle outside the grid has clamped stencils whose
// weights can be negative, and two's-complement sums stay exact), momentum
// w (m v + A d) with A = m C - dt V0 4/DX^2 tau, and the pinned weight w m.
// Before the scatter of a frame-loop substep (pre_mode >= 0), the driver's
// particle update of that substep (src/simulation.js:271-290): pinned
// particles stop; the others take gravity (mode 1 uniform (0, -1), 2 towards
// the centre -(x, y) / sqrt(x^2 + y^2 + 400)) times 27 strength over the
// substep and are clamped to +-180 per axis. Returns the particle's velocity.
__device__ __forceinline__ float2 frame_pre(float* __restrict__ P, uint32_t cap, uint32_t p, bool pin, int pre_mode,
                                           float strength, float h) {
  float vx = ld2(P, cap, Q_VX, p), vy = ld2(P, cap, Q_VY, p);
  if (pre_mode < 0) return make_float2(vx, vy);
  if (pin) {
    vx = vy = 0.f;
  } else {
    float gx = 0.f, gy = 0.f;
    if (pre_mode == 2) {
      const float x = ld2(P, cap, Q_X, p), y = ld2(P, cap, Q_Y, p);
      const float ir = 1.f / sqrtf(x * x + y * y + 400.f);
      gx = -x * ir;
      gy = -y * ir;
    } else if (pre_mode == 1) {
      gy = -1.f;
    }
    const float force = 27.f * strength;
    vx = fminf(fmaxf(vx + gx * force * h, -180.f), 180.f);
    vy = fminf(fmaxf(vy + gy * force * h, -180.f), 180.f);
  }
  st2(P, cap, Q_VX, p, vx);
  st2(P, cap, Q_VY, p, vy);
  return make_float2(vx, vy);
}

__global__ __launch_bounds__(BLOCK) void k2_p2g(float* __restrict__ P, const uint8_t* __restrict__ pinned,
                                                uint32_t cap, uint32_t n, float dt, const Loop2* loop, Params2 prm,
                                                int pre_mode, float strength,
                                                unsigned long long* __restrict__ acc_m, float* __restrict__ acc_u,
                                                float* __restrict__ acc_v, float* __restrict__ acc_fix) {
  const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
  const float h = step_dt(dt, loop);
  if (p >= n || h == 0.f) return;
  const bool pin = pinned[p] != 0;
  const float2 vel = frame_pre(P, cap, p, pin, pre_mode, strength, h);
  float fx, fy, wx[3], wy[3];
  const int bx = stencil_axis(ld2(P, cap, Q_X, p), fx, wx), by = stencil_axis(ld2(P, cap, Q_Y, p), fy, wy);
  const float m = ld2(P, cap, Q_MASS, p), V0 = AREA2 * ld2(P, cap, Q_VOL, p);
  const float E = m / V0 * 16000.f * prm.stiffness;
  float t[4];
  solid_stress(ld2(P, cap, Q_F0, p), ld2(P, cap, Q_F0 + 1, p), ld2(P, cap, Q_F0 + 2, p), ld2(P, cap, Q_F0 + 3, p),
               .4f * E, .4f * E, t);
  const float factor = -h * V0 * 4.f / (DX2 * DX2);
  const float a00 = m * ld2(P, cap, Q_C0, p) + factor * t[0], a01 = m * ld2(P, cap, Q_C0 + 1, p) + factor * t[1];
  const float a10 = m * ld2(P, cap, Q_C0 + 2, p) + factor * t[2], a11 = m * ld2(P, cap, Q_C0 + 3, p) + factor * t[3];
  const float mvx = m * vel.x, mvy = m * vel.y;
  const int32_t Q = (int32_t)rintf(m * prm.mass_scale);
  int32_t given = 0;
  int32_t part[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    part[k] = (int32_t)rintf(wx[k % 3] * wy[k / 3] * (float)Q);
    if (k != 4) given += part[k];
  }
  part[4] = Q - given;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const uint32_t node = (uint32_t)(by + j) * N2 + (uint32_t)(bx + i);
      const float w = wx[i] * wy[j], dx = ((float)i - fx) * DX2, dy = ((float)j - fy) * DX2;
      atomicAdd(&acc_m[node], (unsigned long long)(long long)part[j * 3 + i]);
      atomicAdd(&acc_u[node], w * (mvx + a00 * dx + a01 * dy));
      atomicAdd(&acc_v[node], w * (mvy + a10 * dx + a11 * dy));
      if (pin) atomicAdd(&acc_fix[node], w * m);
    }
}

// Grid velocities (src/solid-mpm.js:67-80): p / m where m > 1e-12 (counted as
// active), zero at pinned nodes, walls stop outward motion; clears the
// accumulators. FP64 sums of the fixed-point mass go to `mass_q`.
__global__ __launch_bounds__(BLOCK) void k2_grid(unsigned long long* __restrict__ acc_m, float* __restrict__ acc_u,
                                                 float* __restrict__ acc_v, float* __restrict__ acc_fix,
                                                 float2* __restrict__ vel, const Loop2* loop, Params2 prm,
                                                 unsigned int* __restrict__ active,
                                                 unsigned long long* __restrict__ mass_q) {
  const uint32_t node = blockIdx.x * BLOCK + threadIdx.x;
  if (loop && loop->h == 0.0) return;
  bool act = false;
  unsigned long long Mq = 0;
  if (node < (uint32_t)NODES2) {
    Mq = acc_m[node];
    const float m = (float)(long long)Mq * prm.inv_mass_scale;
    float u = acc_u[node], v = acc_v[node];
    if (m > 1e-12f) {
      act = true;
      const float im = 1.f / m;
      u *= im;
      v *= im;
      if (acc_fix[node] > 1e-9f) {
        u = v = 0.f;
      } else if (prm.walls) {
        const float x = ORIGIN2 + (float)(node % N2) * DX2, y = ORIGIN2 + (float)(node / N2) * DX2;
        if ((x < -prm.bound + DX2 && u < 0.f) || (x > prm.bound - DX2 && u > 0.f)) u = 0.f;
        if ((y < prm.floor + DX2 && v < 0.f) || (y > prm.top - DX2 && v > 0.f)) v = 0.f;
      }
    }
    vel[node] = make_float2(u, v);
    acc_m[node] = 0ull;
    acc_u[node] = 0.f;
    acc_v[node] = 0.f;
    acc_fix[node] = 0.f;
  }
  const uint64_t b = __ballot(act);
  const unsigned long long s = wave_sum(Mq);
  if ((threadIdx.x & 63u) == 0) {
    if (b) atomicAdd(active, (unsigned int)__popcll(b));
    if (s) atomicAdd(mass_q, s);
  }
}

// G2P (src/solid-mpm.js:81-92): APIC C, F <- (I + dt C) F, v, x += dt v,
// positions clamped 0.1 inside the walls. Pinned particles keep their state.
__global__ __launch_bounds__(BLOCK) void k2_g2p(float* __restrict__ P, const uint8_t* __restrict__ pinned,
                                                uint32_t cap, uint32_t n, float dt, const Loop2* loop, Params2 prm,
                                                const float2* __restrict__ vel) {
  const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
  const float h = step_dt(dt, loop);
  if (p >= n || h == 0.f || pinned[p]) return;
  float fx, fy, wx[3], wy[3];
  const float x = ld2(P, cap, Q_X, p), y = ld2(P, cap, Q_Y, p);
  const int bx = stencil_axis(x, fx, wx), by = stencil_axis(y, fy, wy);
  float vx = 0.f, vy = 0.f, a = 0.f, b = 0.f, c = 0.f, d = 0.f;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float2 u = vel[(uint32_t)(by + j) * N2 + (uint32_t)(bx + i)];
      const float w = wx[i] * wy[j], dx = ((float)i - fx) * (4.f / DX2), dy = ((float)j - fy) * (4.f / DX2);
      vx += w * u.x;
      vy += w * u.y;
      a += w * u.x * dx;
      b += w * u.x * dy;
      c += w * u.y * dx;
      d += w * u.y * dy;
    }
  st2(P, cap, Q_C0, p, a);
  st2(P, cap, Q_C0 + 1, p, b);
  st2(P, cap, Q_C0 + 2, p, c);
  st2(P, cap, Q_C0 + 3, p, d);
  const float f = ld2(P, cap, Q_F0, p), g = ld2(P, cap, Q_F0 + 1, p), e = ld2(P, cap, Q_F0 + 2, p),
              k = ld2(P, cap, Q_F0 + 3, p);
  st2(P, cap, Q_F0, p, (1.f + h * a) * f + h * b * e);
  st2(P, cap, Q_F0 + 1, p, (1.f + h * a) * g + h * b * k);
  st2(P, cap, Q_F0 + 2, p, h * c * f + (1.f + h * d) * e);
  st2(P, cap, Q_F0 + 3, p, h * c * g + (1.f + h * d) * k);
  st2(P, cap, Q_VX, p, vx);
  st2(P, cap, Q_VY, p, vy);
  float nx = x + h * vx, ny = y + h * vy;
  if (prm.walls) {
    nx = fminf(fmaxf(nx, -prm.bound + .1f), prm.bound - .1f);
    ny = fminf(fmaxf(ny, prm.floor + .1f), prm.top - .1f);
  }
  st2(P, cap, Q_X, p, nx);
  st2(P, cap, Q_Y, p, ny);
}

// Stable-step maxima (src/solid-mpm.js:38-47): wave = max(1, |v| +
// sqrt(19200 stiffness) max(1, |F|_F)^2) and gradient = max |C|_1, as the
// bits of non-negative floats (their order is the float order).
__global__ __launch_bounds__(BLOCK) void k2_stable(const float* __restrict__ P, uint32_t cap, uint32_t n,
                                                   Params2 prm, Loop2* loop, unsigned int* __restrict__ bits) {
  if (loop && (loop->remaining <= 1e-9 || loop->steps >= loop->max_steps)) return;
  const float sound = sqrtf(19200.f * prm.stiffness);
  float wave = 1.f, grad = 0.f;
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
    const float f0 = ld2(P, cap, Q_F0, p), f1 = ld2(P, cap, Q_F0 + 1, p), f2 = ld2(P, cap, Q_F0 + 2, p),
                f3 = ld2(P, cap, Q_F0 + 3, p);
    const float stretch = fmaxf(1.f, sqrtf(f0 * f0 + f1 * f1 + f2 * f2 + f3 * f3));
    const float vx = ld2(P, cap, Q_VX, p), vy = ld2(P, cap, Q_VY, p);
    wave = fmaxf(wave, sqrtf(vx * vx + vy * vy) + sound * stretch * stretch);
    grad = fmaxf(grad, fabsf(ld2(P, cap, Q_C0, p)) + fabsf(ld2(P, cap, Q_C0 + 1, p)) + fabsf(ld2(P, cap, Q_C0 + 2, p)) +
                           fabsf(ld2(P, cap, Q_C0 + 3, p)));
  }
  const auto fmax_fold = [](float a, float b) { return fmaxf(a, b); };
  wave = wave_reduce(wave, 0.f, fmax_fold);  // both maxima are >= 0
  grad = wave_reduce(grad, 0.f, fmax_fold);
  if ((threadIdx.x & 63u) == 0) {
    atomicMax(&bits[0], __float_as_uint(wave));
    atomicMax(&bits[1], __float_as_uint(grad));
  }
}

__device__ __forceinline__ double stable_of(unsigned int wave_bits, unsigned int grad_bits) {
  const double wave = fmax(1.0, (double)__uint_as_float(wave_bits));
  const double grad = (double)__uint_as_float(grad_bits);
  return fmin(fmin(1.0 / 240, .18 * DX2 / wave), .1 / fmax(1.0, grad));
}

// One thread: this substep's size from the maxima, then the maxima reset for
// the next substep; a substep that runs restarts the grid counters, so they
// describe the frame's last substep.
__global__ void k2_begin(Loop2* loop, unsigned int* active, unsigned long long* grid_mass_q) {
  const bool more = loop->remaining > 1e-9 && loop->steps < loop->max_steps;
  const double h = more ? fmin(loop->remaining, stable_of(loop->wave_bits, loop->grad_bits)) : 0.0;
  loop->h = h;
  if (more) {
    loop->remaining -= h;
    loop->steps += 1;
    *active = 0u;
    *grid_mass_q = 0ull;
  }
  loop->wave_bits = 0u;
  loop->grad_bits = 0u;
}


// Deformation statistics (src/solid-mpm.js:106-115): per-block min/max det F
// and max Green strain norm; particle fixed-point mass total.
__global__ __launch_bounds__(BLOCK) void k2_stats(const float* __restrict__ P, uint32_t cap, uint32_t n,
                                                  float mass_scale, double* __restrict__ out,
                                                  unsigned long long* __restrict__ mass_q) {
  __shared__ double s[3][BLOCK / 64];
  double jmin = INFINITY, jmax = -INFINITY, strain = 0;
  unsigned long long q = 0;
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
    const float a = ld2(P, cap, Q_F0, p), b = ld2(P, cap, Q_F0 + 1, p), c = ld2(P, cap, Q_F0 + 2, p),
                d = ld2(P, cap, Q_F0 + 3, p);
    const double J = (double)a * d - (double)b * c;
    jmin = fmin(jmin, J);
    jmax = fmax(jmax, J);
    const double e0 = (double)a * a + (double)c * c - 1, e1 = (double)b * b + (double)d * d - 1,
                 e2 = 1.4142135623730951 * ((double)a * b + (double)c * d);
    strain = fmax(strain, .5 * sqrt(e0 * e0 + e1 * e1 + e2 * e2));
    q += (unsigned long long)(long long)(int32_t)rintf(ld2(P, cap, Q_MASS, p) * mass_scale);
  }
  jmin = wave_reduce(jmin, (double)INFINITY, [](double a, double b) { return fmin(a, b); });
  jmax = wave_reduce(jmax, -(double)INFINITY, [](double a, double b) { return fmax(a, b); });
  strain = wave_reduce(strain, 0.0, [](double a, double b) { return fmax(a, b); });
  q = wave_sum(q);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s[0][wave] = jmin;
    s[1][wave] = jmax;
    s[2][wave] = strain;
    if (q) atomicAdd(mass_q, q);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; ++w) {
      s[0][0] = fmin(s[0][0], s[0][w]);
      s[1][0] = fmax(s[1][0], s[1][w]);
      s[2][0] = fmax(s[2][0], s[2][w]);
    }
    out[blockIdx.x * 3 + 0] = s[0][0];
    out[blockIdx.x * 3 + 1] = s[1][0];
    out[blockIdx.x * 3 + 2] = s[2][0];
  }
}

inline unsigned blocks2(size_t n) { return (unsigned)std::max<size_t>(1, (n + BLOCK - 1) / BLOCK); }

}  // namespace

struct mpm2_engine {
  uint32_t cap = 0, n = 0;
  hipStream_t stream = nullptr;
  float* planes = nullptr;
  uint8_t* pinned = nullptr;
  unsigned long long* acc_m = nullptr;
  float *acc_u = nullptr, *acc_v = nullptr, *acc_fix = nullptr;
  float2* vel = nullptr;
  unsigned int* counters = nullptr;       // [0] active nodes, [1..2] stable-step maxima bits
  unsigned long long* mass_q = nullptr;   // [0] grid fixed-point mass of the last step, [1] particle total
  Loop2* loop = nullptr;                  // device frame loop state
  Loop2* loop_host = nullptr;             // pinned mirror
  double* red = nullptr;                  // [RED2_BLOCKS * 3]
  float mass_scale = 1.f;
  uint32_t substeps = 0, active_nodes = 0;
  std::string err;
};

namespace {

int fail2(mpm2_engine* e, int code, const char* fmt, ...) {
  if (e) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    e->err = buf;
  }
  return code;
}

#define HIP2(e, call)                                                                  \
  do {                                                                                 \
    hipError_t st_ = (call);                                                           \
    if (st_ != hipSuccess) return fail2(e, MPM_ERR_HIP, "%s: %s", #call, hipGetErrorString(st_)); \
  } while (0)

bool settings_ok(const mpm2_settings* s) {
  return s && std::isfinite(s->stiffness) && s->stiffness > 0 && std::isfinite(s->bound) && std::isfinite(s->floor) &&
         std::isfinite(s->top) && (!s->walls || (s->bound > DX2 && s->top > s->floor + DX2));
}

Params2 params_of(const mpm2_engine* e, const mpm2_settings* s) {
  Params2 p;
  p.stiffness = s->stiffness;
  p.walls = s->walls != 0;
  p.bound = s->bound;
  p.floor = s->floor;
  p.top = s->top;
  p.mass_scale = e->mass_scale;
  p.inv_mass_scale = 1.f / e->mass_scale;
  return p;
}

double rd2(const void* p, int32_t dt, size_t i) {
  return dt == MPM_F64 ? static_cast<const double*>(p)[i] : static_cast<const float*>(p)[i];
}
void wr2(void* p, int32_t dt, size_t i, double v) {
  if (dt == MPM_F64)
    static_cast<double*>(p)[i] = v;
  else
    static_cast<float*>(p)[i] = (float)v;
}

// One MPM step with a fixed dt, or one substep of the frame loop (loop set).
// One step: P2G (after the frame loop's particle update when pre_mode >= 0),
// grid, G2P.
void enqueue_step(mpm2_engine* e, const Params2& prm, float dt, const Loop2* loop, int pre_mode = -1,
                  float strength = 0.f) {
  const uint32_t n = e->n;
  hipLaunchKernelGGL(k2_p2g, dim3(blocks2(n)), dim3(BLOCK), 0, e->stream, e->planes, e->pinned, e->cap, n, dt, loop, prm,
                     pre_mode, strength, e->acc_m, e->acc_u, e->acc_v, e->acc_fix);
  hipLaunchKernelGGL(k2_grid, dim3(blocks2(NODES2)), dim3(BLOCK), 0, e->stream, e->acc_m, e->acc_u, e->acc_v,
                     e->acc_fix, e->vel, loop, prm, e->counters, e->mass_q);
  hipLaunchKernelGGL(k2_g2p, dim3(blocks2(n)), dim3(BLOCK), 0, e->stream, e->planes, e->pinned, e->cap, n, dt, loop, prm,
                     e->vel);
}

}  // namespace

extern "C" {

int mpm2_create(uint32_t capacity, mpm2_engine** out) {
  if (!out || capacity == 0) return MPM_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  mpm2_engine* e = new mpm2_engine();
  e->cap = capacity;
  hipError_t st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  auto alloc = [&](void** p, size_t bytes) {
    if (st == hipSuccess) st = hipMalloc(p, bytes);
  };
  alloc((void**)&e->planes, sizeof(float) * NPLANES2 * (size_t)capacity);
  alloc((void**)&e->pinned, capacity);
  alloc((void**)&e->acc_m, sizeof(unsigned long long) * NODES2);
  alloc((void**)&e->acc_u, sizeof(float) * NODES2);
  alloc((void**)&e->acc_v, sizeof(float) * NODES2);
  alloc((void**)&e->acc_fix, sizeof(float) * NODES2);
  alloc((void**)&e->vel, sizeof(float2) * NODES2);
  alloc((void**)&e->counters, sizeof(unsigned int) * 4);
  alloc((void**)&e->mass_q, sizeof(unsigned long long) * 2);
  alloc((void**)&e->loop, sizeof(Loop2));
  alloc((void**)&e->red, sizeof(double) * 3 * RED2_BLOCKS);
  if (st == hipSuccess) st = hipHostMalloc((void**)&e->loop_host, sizeof(Loop2), hipHostMallocDefault);
  if (st == hipSuccess) st = hipMemsetAsync(e->acc_m, 0, sizeof(unsigned long long) * NODES2, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(e->acc_u, 0, sizeof(float) * NODES2, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(e->acc_v, 0, sizeof(float) * NODES2, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(e->acc_fix, 0, sizeof(float) * NODES2, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(e->counters, 0, sizeof(unsigned int) * 4, e->stream);
  if (st == hipSuccess) st = hipMemsetAsync(e->mass_q, 0, sizeof(unsigned long long) * 2, e->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
  if (st != hipSuccess) {
    mpm2_destroy(e);
    return MPM_ERR_HIP;
  }
  *out = e;
  return MPM_OK;
}

int mpm2_destroy(mpm2_engine* e) {
  if (!e) return MPM_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (void* p : {(void*)e->planes, (void*)e->pinned, (void*)e->acc_m, (void*)e->acc_u, (void*)e->acc_v,
                  (void*)e->acc_fix, (void*)e->vel, (void*)e->counters, (void*)e->mass_q, (void*)e->loop,
                  (void*)e->red})
    if (p) (void)hipFree(p);
  if (e->loop_host) (void)hipHostFree(e->loop_host);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return MPM_OK;
}

const char* mpm2_last_error(const mpm2_engine* e) { return e ? e->err.c_str() : "null engine"; }

int mpm2_settings_defaults(const char* scene, mpm2_settings* out) {
  if (!out) return MPM_ERR_INVALID_ARGUMENT;
  // src/simulation.js:26 (solidStiffness 1, boundary on) and the wall boxes of
  // src/solid-mpm.js:69-70.
  *out = mpm2_settings{1.f, 1, 555.f, -555.f, 555.f};
  if (scene && !std::strcmp(scene, "softbody")) {
    out->bound = 170.f;
    out->floor = -145.f;
    out->top = 145.f;
  } else if (scene && !std::strcmp(scene, "basin")) {
    out->bound = 245.f;
    out->floor = -175.f;
    out->top = 245.f;
  }
  return MPM_OK;
}

int mpm2_upload(mpm2_engine* e, const mpm2_particles* p) {
  if (!e || !p) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = p->count;
  if (n > e->cap) return fail2(e, MPM_ERR_CAPACITY, "count %u exceeds capacity %u", n, e->cap);
  if (n && (!p->x || !p->y || !p->vx || !p->vy || !p->F || !p->C || !p->mass || !p->volume || !p->pinned ||
            (p->dtype != MPM_F64 && p->dtype != MPM_F32)))
    return fail2(e, MPM_ERR_INVALID_ARGUMENT, "missing particle array or bad dtype");
  double mmax = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double m = rd2(p->mass, p->dtype, i), v = rd2(p->volume, p->dtype, i);
    if (!(m > 0) || !(v > 0) || !std::isfinite(m) || !std::isfinite(v))
      return fail2(e, MPM_ERR_INVALID_ARGUMENT, "particle %u: mass and volume must be positive", i);
    mmax = std::max(mmax, m);
  }
  // Fixed-point mass: the largest particle is at most 2^24 units, so every
  // per-particle part and every node total of up to 2^39 particles fit 64 bits.
  e->mass_scale = n ? (float)std::ldexp(1.0, (int)std::floor(24 - std::log2(mmax))) : 1.f;
  std::vector<float> h((size_t)NPLANES2 * e->cap, 0.f);
  std::vector<uint8_t> pin(e->cap, 0);
  auto set = [&](int k, uint32_t i, double v) { h[(size_t)k * e->cap + i] = (float)v; };
  for (uint32_t i = 0; i < n; ++i) {
    set(Q_X, i, rd2(p->x, p->dtype, i));
    set(Q_Y, i, rd2(p->y, p->dtype, i));
    set(Q_VX, i, rd2(p->vx, p->dtype, i));
    set(Q_VY, i, rd2(p->vy, p->dtype, i));
    for (int k = 0; k < 4; ++k) {
      set(Q_F0 + k, i, rd2(p->F, p->dtype, 4 * (size_t)i + k));
      set(Q_C0 + k, i, rd2(p->C, p->dtype, 4 * (size_t)i + k));
    }
    set(Q_MASS, i, rd2(p->mass, p->dtype, i));
    set(Q_VOL, i, rd2(p->volume, p->dtype, i));
    pin[i] = p->pinned[i] ? 1 : 0;
  }
  HIP2(e, hipMemcpyAsync(e->planes, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, e->stream));
  HIP2(e, hipMemcpyAsync(e->pinned, pin.data(), e->cap, hipMemcpyHostToDevice, e->stream));
  HIP2(e, hipStreamSynchronize(e->stream));
  e->n = n;
  return MPM_OK;
}

int mpm2_download(mpm2_engine* e, mpm2_particles* p) {
  if (!e || !p) return MPM_ERR_INVALID_ARGUMENT;
  if (p->count != e->n) return fail2(e, MPM_ERR_INVALID_ARGUMENT, "count %u != resident %u", p->count, e->n);
  if (p->dtype != MPM_F64 && p->dtype != MPM_F32) return fail2(e, MPM_ERR_INVALID_ARGUMENT, "bad dtype");
  const uint32_t n = e->n;
  std::vector<float> h((size_t)NPLANES2 * e->cap);
  HIP2(e, hipMemcpyAsync(h.data(), e->planes, sizeof(float) * h.size(), hipMemcpyDeviceToHost, e->stream));
  HIP2(e, hipStreamSynchronize(e->stream));
  auto get = [&](int k, uint32_t i) { return (double)h[(size_t)k * e->cap + i]; };
  for (uint32_t i = 0; i < n; ++i) {
    if (p->x) wr2(p->x, p->dtype, i, get(Q_X, i));
    if (p->y) wr2(p->y, p->dtype, i, get(Q_Y, i));
    if (p->vx) wr2(p->vx, p->dtype, i, get(Q_VX, i));
    if (p->vy) wr2(p->vy, p->dtype, i, get(Q_VY, i));
    for (int k = 0; k < 4; ++k) {
      if (p->F) wr2(p->F, p->dtype, 4 * (size_t)i + k, get(Q_F0 + k, i));
      if (p->C) wr2(p->C, p->dtype, 4 * (size_t)i + k, get(Q_C0 + k, i));
    }
    if (p->mass) wr2(p->mass, p->dtype, i, get(Q_MASS, i));
    if (p->volume) wr2(p->volume, p->dtype, i, get(Q_VOL, i));
  }
  return MPM_OK;
}

int mpm2_step(mpm2_engine* e, const mpm2_settings* s, double dt, int32_t nsteps) {
  if (!e || !settings_ok(s) || !(dt > 0) || !std::isfinite(dt) || nsteps < 0)
    return fail2(e, MPM_ERR_INVALID_ARGUMENT, "bad settings, dt or step count");
  if (!e->n) return MPM_OK;
  const Params2 prm = params_of(e, s);
  HIP2(e, hipMemsetAsync(e->counters, 0, sizeof(unsigned int), e->stream));
  HIP2(e, hipMemsetAsync(e->mass_q, 0, sizeof(unsigned long long), e->stream));
  for (int k = 0; k < nsteps; ++k) {
    if (k == nsteps - 1 && k > 0) {  // the last step's active-node count and mass
      HIP2(e, hipMemsetAsync(e->counters, 0, sizeof(unsigned int), e->stream));
      HIP2(e, hipMemsetAsync(e->mass_q, 0, sizeof(unsigned long long), e->stream));
    }
    enqueue_step(e, prm, (float)dt, nullptr);
  }
  HIP2(e, hipGetLastError());
  e->substeps = (uint32_t)nsteps;
  return MPM_OK;
}

int mpm2_stable_dt(mpm2_engine* e, const mpm2_settings* s, double* out) {
  if (!e || !settings_ok(s) || !out) return fail2(e, MPM_ERR_INVALID_ARGUMENT, "bad settings or output");
  unsigned int bits[2] = {0, 0};
  if (e->n) {
    HIP2(e, hipMemsetAsync(e->counters + 1, 0, sizeof(unsigned int) * 2, e->stream));
    hipLaunchKernelGGL(k2_stable, dim3(std::min<unsigned>(RED2_BLOCKS, blocks2(e->n))), dim3(BLOCK), 0, e->stream,
                       e->planes, e->cap, e->n, params_of(e, s), nullptr, e->counters + 1);
    HIP2(e, hipMemcpyAsync(bits, e->counters + 1, sizeof bits, hipMemcpyDeviceToHost, e->stream));
    HIP2(e, hipStreamSynchronize(e->stream));
  }
  float w, g;
  std::memcpy(&w, &bits[0], 4);
  std::memcpy(&g, &bits[1], 4);
  *out = std::min(std::min(1.0 / 240, .18 * DX2 / std::max(1.0, (double)w)), .1 / std::max(1.0, (double)g));
  return MPM_OK;
}

int mpm2_advance(mpm2_engine* e, const mpm2_settings* s, double dt, int32_t gravity_mode, float strength,
                 int32_t max_substeps, int32_t* substeps, double* advanced) {
  if (!e || !settings_ok(s) || !std::isfinite(dt) || gravity_mode < 0 || gravity_mode > 2 ||
      !std::isfinite(strength) || max_substeps < 1)
    return fail2(e, MPM_ERR_INVALID_ARGUMENT, "bad settings, dt, gravity mode or substep limit");
  dt = std::min(std::max(dt, 0.0), .05);  // src/simulation.js:241
  Loop2 init{};
  init.remaining = e->n ? dt : 0.0;  // nothing to step
  init.max_steps = max_substeps;
  const Params2 prm = params_of(e, s);
  if (e->n) {
    HIP2(e, hipMemcpyAsync(e->loop, &init, sizeof init, hipMemcpyHostToDevice, e->stream));
    HIP2(e, hipMemsetAsync(e->counters, 0, sizeof(unsigned int), e->stream));
    HIP2(e, hipMemsetAsync(e->mass_q, 0, sizeof(unsigned long long), e->stream));
    const unsigned sblocks = std::min<unsigned>(RED2_BLOCKS, blocks2(e->n));
    // The first round enqueues as many substeps as the previous frame took,
    // plus one, so a steady run of frames costs one host round trip and few
    // launches that find the loop finished; later rounds enqueue CHUNK.
    int round = e->substeps ? std::min(CHUNK, (int)e->substeps + 1) : CHUNK;
    for (;;) {
      for (int k = 0; k < round; ++k) {
        hipLaunchKernelGGL(k2_stable, dim3(sblocks), dim3(BLOCK), 0, e->stream, e->planes, e->cap, e->n, prm, e->loop,
                           &e->loop->wave_bits);
        hipLaunchKernelGGL(k2_begin, dim3(1), dim3(1), 0, e->stream, e->loop, e->counters, e->mass_q);
        enqueue_step(e, prm, 0.f, e->loop, gravity_mode, strength);
      }
      HIP2(e, hipGetLastError());
      HIP2(e, hipMemcpyAsync(e->loop_host, e->loop, sizeof(Loop2), hipMemcpyDeviceToHost, e->stream));
      HIP2(e, hipStreamSynchronize(e->stream));
      if (!(e->loop_host->remaining > 1e-9) || e->loop_host->steps >= max_substeps) break;
      round = CHUNK;
    }
    init = *e->loop_host;
  }
  e->substeps = (uint32_t)init.steps;
  if (substeps) *substeps = init.steps;
  if (advanced) *advanced = dt - init.remaining;
  return MPM_OK;
}

int mpm2_statistics(mpm2_engine* e, mpm2_stats* out) {
  if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
  *out = mpm2_stats{};
  out->count = e->n;
  out->min_J = out->max_J = 1;
  out->substeps = e->substeps;
  unsigned long long q[2] = {0, 0};
  unsigned int active = 0;
  if (e->n) {
    const unsigned nb = std::min<unsigned>(RED2_BLOCKS, blocks2(e->n));
    HIP2(e, hipMemsetAsync(e->mass_q + 1, 0, sizeof(unsigned long long), e->stream));
    hipLaunchKernelGGL(k2_stats, dim3(nb), dim3(BLOCK), 0, e->stream, e->planes, e->cap, e->n, e->mass_scale, e->red,
                       e->mass_q + 1);
    std::vector<double> r(3 * (size_t)nb);
    HIP2(e, hipMemcpyAsync(r.data(), e->red, sizeof(double) * r.size(), hipMemcpyDeviceToHost, e->stream));
    HIP2(e, hipMemcpyAsync(q, e->mass_q, sizeof q, hipMemcpyDeviceToHost, e->stream));
    HIP2(e, hipMemcpyAsync(&active, e->counters, sizeof active, hipMemcpyDeviceToHost, e->stream));
    HIP2(e, hipStreamSynchronize(e->stream));
    out->min_J = INFINITY;
    out->max_J = -INFINITY;
    for (unsigned b = 0; b < nb; ++b) {
      out->min_J = std::min(out->min_J, r[3 * b]);
      out->max_J = std::max(out->max_J, r[3 * b + 1]);
      out->strain = std::max(out->strain, r[3 * b + 2]);
    }
  }
  out->nodes = active;
  out->grid_mass_q = (int64_t)q[0];
  out->particle_mass_q = (int64_t)q[1];
  return MPM_OK;
}

int mpm2_synchronize(mpm2_engine* e) {
  if (!e) return MPM_ERR_INVALID_ARGUMENT;
  HIP2(e, hipStreamSynchronize(e->stream));
  return MPM_OK;
}

}  // extern "C"
