// Learned constitutive surrogate: training-sample generation with the plain
// return map, FP32 training kernels (dense forward, backward, Adam), and the
// packing of trained weights into BF16 MFMA tiles for the step kernel.
// No counterpart in the reference.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mpm_internal.h"
#include "mpm_mlp_device.h"

namespace mpm {
namespace {

constexpr int TB = 256;

// Seeded 32-bit hash mapped to [0, 1).
__device__ __forceinline__ float uhash(uint32_t seed, uint32_t i, uint32_t c) {
  uint32_t h = seed + (i + 1u) * 0x9e3779b9u + (c + 1u) * 0x85ebca6bu;
  h = (h ^ (h >> 16)) * 0x7feb352du;
  h = (h ^ (h >> 15)) * 0x846ca68bu;
  h ^= h >> 16;
  return (float)(h >> 8) * (1.f / 16777216.f);
}

__device__ __forceinline__ float gauss(uint32_t seed, uint32_t i, uint32_t c) {
  const float u1 = fmaxf(uhash(seed, i, 2 * c), 1e-7f);
  const float u2 = uhash(seed, i, 2 * c + 1);
  return sqrtf(-2.f * logf(u1)) * cosf(6.283185307f * u2);
}

// Symmetric 6-vector (xx, yy, zz, xy, xz, yz) helpers.
__device__ __forceinline__ void sym_eig(const float e[6], float V[9], float lam[3]) {
  jacobi_eig3(e[0], e[1], e[2], e[3], e[4], e[5], V, lam);
}

// One training sample: T uniform in [T_lo, T_hi]; log strain eps ~ N(0,
// sigma) per component, or (for a share near of the samples) a deviatoric
// part of norm 1 + 1.5 u^2 (u uniform: denser at the yield surface, where the
// trial states of a flowing scene sit) times the yield strain at T in a
// random direction plus a volumetric part N(0, sigma). Every sample lies on the
// plastic side of the yield surface: a N(0, sigma) draw inside it has its
// deviatoric part stretched to 1 + 1.5 u times the yield strain. The network
// serves only trial states past yield (the elastic ones take the closed
// form), and without the elastic side the map it fits has no kink at the
// yield surface, where the states just past yield crowd. F = exp(eps) (the
// trial state with no rotation); the plain return map of the step gives the
// Kirchhoff stress; the target is its deviatoric part in units of
// out_scale * y(T) * E, the input the strain in units of the yield strain at
// T (eps sqrt(6) G / y(T)) times in_scale. Inputs and targets feature-major.
__global__ __launch_bounds__(TB) void k_mlp_samples(uint32_t seed, uint32_t n, float sigma, float T_lo, float T_hi,
                                                    float dt, float near, DevMaterial mat, MlpNorm nrm,
                                                    float* __restrict__ X, float* __restrict__ Y) {
  const uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  const float T = T_lo + (T_hi - T_lo) * uhash(seed, i, 100);
  float e[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) e[c] = gauss(seed, i, c);
  // yield strain at T: |dev eps| at which q = sqrt(6) G |dev eps| reaches y
  const float ey = surrogate_yield(T, mat) / (2.4494897427831781f * mat.G);
  if (uhash(seed, i, 101) < near) {
    const float m = (e[0] + e[1] + e[2]) * (1.f / 3.f);
    e[0] -= m;
    e[1] -= m;
    e[2] -= m;
    const float len = sym_norm(e);
    const float u = uhash(seed, i, 102);
    const float v = sigma * gauss(seed, i, 7);
    // J = exp(3 v) divides q: the yield strain of this volume is ey J
    const float r = ey * expf(3.f * v) * (1.f + 1.5f * u * u) / fmaxf(len, 1e-20f);
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] = e[c] * r + (c < 3 ? v : 0.f);
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] *= sigma;
    const float m = (e[0] + e[1] + e[2]) * (1.f / 3.f);
    float d[6] = {e[0] - m, e[1] - m, e[2] - m, e[3], e[4], e[5]};
    const float len = sym_norm(d);
    // J = exp(tr eps) divides q: the yield strain of this volume is ey J
    const float eyJ = ey * expf(3.f * m);
    if (len < eyJ) {
      const float r = eyJ * (1.f + 1.5f * uhash(seed, i, 103)) / fmaxf(len, 1e-20f);
#pragma unroll
      for (int c = 0; c < 6; ++c) e[c] = d[c] * r + (c < 3 ? m : 0.f);
    }
  }
  float V[9], lam[3];
  sym_eig(e, V, lam);
  float s[3] = {expf(lam[0]), expf(lam[1]), expf(lam[2])};
  M3 F;
  compose(V, s, V, F);
  float tau[6];
  plastic_project(F, 0.f, dt, T, mat, tau);
  const float p = (tau[0] + tau[1] + tau[2]) * (1.f / 3.f);
  const float dev[6] = {tau[0] - p, tau[1] - p, tau[2] - p, tau[3], tau[4], tau[5]};
  const float ys = surrogate_yield(T, mat);
  const float kin = nrm.in_scale * 2.4494897427831781f * mat.G / ys;
#pragma unroll
  for (int c = 0; c < 6; ++c) X[(size_t)c * n + i] = e[c] * kin;
  X[(size_t)6 * n + i] = (T - nrm.T_center) * nrm.T_scale;
  const float inv = 1.f / (nrm.out_scale * ys * mat.E);
#pragma unroll
  for (int c = 0; c < 6; ++c) Y[(size_t)c * n + i] = dev[c] * inv;
}

__device__ __forceinline__ float hswish_d(float z) { return z < -3.f ? 0.f : (z > 3.f ? 1.f : z * (1.f / 3.f) + 0.5f); }

// z = W in + b, h = act(z); in is [K][ld_in], z/h are [M][B].
template <bool ACT>
__global__ __launch_bounds__(TB) void k_dense_fwd(const float* __restrict__ in, int ld_in,
                                                  const float* __restrict__ W, const float* __restrict__ b,
                                                  float* __restrict__ z, float* __restrict__ h, int M, int K,
                                                  int B) {
  const int m = blockIdx.y;
  const int col = blockIdx.x * TB + threadIdx.x;
  if (col >= B) return;
  float acc = b[m];
  for (int k = 0; k < K; ++k) acc += W[m * K + k] * in[(size_t)k * ld_in + col];
  if (z) z[(size_t)m * B + col] = acc;
  h[(size_t)m * B + col] = ACT ? hswish_f(acc) : acc;
}

// din = (W^T dz) * act'(zprev); dz is [M][B], din/zprev are [K][B].
__global__ __launch_bounds__(TB) void k_dense_bwd_in(const float* __restrict__ dz, const float* __restrict__ W,
                                                     const float* __restrict__ zprev, float* __restrict__ din, int M,
                                                     int K, int B) {
  const int k = blockIdx.y;
  const int col = blockIdx.x * TB + threadIdx.x;
  if (col >= B) return;
  float acc = 0.f;
  for (int m = 0; m < M; ++m) acc += W[m * K + k] * dz[(size_t)m * B + col];
  din[(size_t)k * B + col] = acc * hswish_d(zprev[(size_t)k * B + col]);
}

// dW[m][k] = sum_b dz[m][b] in[k][b]; db[m] = sum_b dz[m][b] (block (m, k),
// k == K is the bias).
__global__ __launch_bounds__(TB) void k_dense_bwd_w(const float* __restrict__ dz, const float* __restrict__ in,
                                                    int ld_in, float* __restrict__ dW, float* __restrict__ db,
                                                    int M, int K, int B) {
  __shared__ float s[TB / 64];
  const int m = blockIdx.x, k = blockIdx.y;
  float acc = 0.f;
  for (int col = threadIdx.x; col < B; col += TB)
    acc += dz[(size_t)m * B + col] * (k < K ? in[(size_t)k * ld_in + col] : 1.f);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < TB / 64; ++w) t += s[w];
    if (k < K)
      dW[m * K + k] = t;
    else
      db[m] = t;
  }
}

// Mean squared error over the batch (summed over outputs); dy = 2 (y - t) / B.
__global__ __launch_bounds__(TB) void k_mse(const float* __restrict__ y, const float* __restrict__ t, int ld_t,
                                            float* __restrict__ dy, int B, double* __restrict__ loss) {
  __shared__ double s[TB / 64];
  const int col = blockIdx.x * TB + threadIdx.x;
  double acc = 0;
  if (col < B) {
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) {
      const float e = y[(size_t)o * B + col] - t[(size_t)o * ld_t + col];
      dy[(size_t)o * B + col] = 2.f * e / (float)B;
      acc += (double)e * e;
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tt = 0;
    for (int w = 0; w < TB / 64; ++w) tt += s[w];
    atomicAdd(loss, tt / B);
  }
}

__global__ __launch_bounds__(TB) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                             float* __restrict__ m1, float* __restrict__ m2, int n, float lr,
                                             float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float v = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = v;
  P[i] -= lr * (a * c1) / (sqrtf(v * c2) + eps);
}

__global__ void k_zero_d(double* p) { *p = 0.0; }

// Device inference of the packed network on n samples (feature-major X) for
// testing the MFMA path against the FP32/FP64 forward: one wave per 64
// samples.
__global__ __launch_bounds__(64) void k_mlp_infer(const MlpDev* __restrict__ M, uint32_t n,
                                                  const float* __restrict__ X, float* __restrict__ Y) {
  __shared__ __align__(16) unsigned char s_act[MLP_WAVE_LDS];
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  const bool use = i < n;
  float x[7], y[6];
#pragma unroll
  for (int c = 0; c < 7; ++c) x[c] = use ? X[(size_t)c * n + i] : 0.f;
  mlp_forward_wave(M, use, x, y, s_act);
  if (use) {
#pragma unroll
    for (int c = 0; c < 6; ++c) Y[(size_t)c * n + i] = y[c];
  }
}

// The step kernel's surrogate return applied to a batch of trial states
// (one wave per 64 states); fallback lanes take the plain return.
__global__ __launch_bounds__(64) void k_sur_batch(const MlpDev* __restrict__ M, DevMaterial mat, uint32_t n, float dt,
                                                  const float* __restrict__ F, const float* __restrict__ ep,
                                                  const float* __restrict__ T, float* __restrict__ Fo,
                                                  float* __restrict__ epo, float* __restrict__ tauo,
                                                  int32_t* __restrict__ branch) {
  __shared__ __align__(16) unsigned char s_act[MLP_WAVE_LDS];
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  const bool valid = i < n;
  const uint32_t ic = valid ? i : 0u;
  M3 f;
#pragma unroll
  for (int k = 0; k < 9; ++k) f[k] = F[(size_t)ic * 9 + k];
  float e = ep[ic], tau[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float t = T[ic];
  int br = surrogate_return(M, valid, f, e, t, mat, tau, s_act);
  if (br == SUR_FALLBACK) e += plastic_project(f, e, dt, t, mat, tau);
  if (valid) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Fo[(size_t)i * 9 + k] = f[k];
    epo[i] = e;
#pragma unroll
    for (int k = 0; k < 6; ++k) tauo[(size_t)i * 6 + k] = tau[k];
    branch[i] = br;
  }
}

// Buffers of one training run.
struct Trainer {
  int B;
  float *P = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;
  float *z1, *h1, *z2, *h2, *z3, *h3, *y, *dy, *d3, *d2, *d1;
  double* loss;
};

void forward(hipStream_t s, Trainer& t, const float* X, int ldx, int B) {
  const dim3 gh((B + TB - 1) / TB, MLP_HID), go((B + TB - 1) / TB, MLP_OUT);
  hipLaunchKernelGGL(k_dense_fwd<true>, gh, dim3(TB), 0, s, X, ldx, t.P + MLP_W1, t.P + MLP_B1, t.z1, t.h1, MLP_HID,
                     MLP_IN, B);
  hipLaunchKernelGGL(k_dense_fwd<true>, gh, dim3(TB), 0, s, t.h1, B, t.P + MLP_W2, t.P + MLP_B2, t.z2, t.h2, MLP_HID,
                     MLP_HID, B);
  hipLaunchKernelGGL(k_dense_fwd<true>, gh, dim3(TB), 0, s, t.h2, B, t.P + MLP_W3, t.P + MLP_B3, t.z3, t.h3, MLP_HID,
                     MLP_HID, B);
  hipLaunchKernelGGL(k_dense_fwd<false>, go, dim3(TB), 0, s, t.h3, B, t.P + MLP_W4, t.P + MLP_B4, (float*)nullptr,
                     t.y, MLP_OUT, MLP_HID, B);
}

void backward(hipStream_t s, Trainer& t, const float* X, int ldx, const float* Y, int ldy, int B) {
  const int nb = (B + TB - 1) / TB;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(TB), 0, s, t.y, Y, ldy, t.dy, B, t.loss);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(MLP_OUT, MLP_HID + 1), dim3(TB), 0, s, t.dy, t.h3, B, t.G + MLP_W4,
                     t.G + MLP_B4, MLP_OUT, MLP_HID, B);
  hipLaunchKernelGGL(k_dense_bwd_in, dim3(nb, MLP_HID), dim3(TB), 0, s, t.dy, t.P + MLP_W4, t.z3, t.d3, MLP_OUT,
                     MLP_HID, B);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(MLP_HID, MLP_HID + 1), dim3(TB), 0, s, t.d3, t.h2, B, t.G + MLP_W3,
                     t.G + MLP_B3, MLP_HID, MLP_HID, B);
  hipLaunchKernelGGL(k_dense_bwd_in, dim3(nb, MLP_HID), dim3(TB), 0, s, t.d3, t.P + MLP_W3, t.z2, t.d2, MLP_HID,
                     MLP_HID, B);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(MLP_HID, MLP_HID + 1), dim3(TB), 0, s, t.d2, t.h1, B, t.G + MLP_W2,
                     t.G + MLP_B2, MLP_HID, MLP_HID, B);
  hipLaunchKernelGGL(k_dense_bwd_in, dim3(nb, MLP_HID), dim3(TB), 0, s, t.d2, t.P + MLP_W2, t.z1, t.d1, MLP_HID,
                     MLP_HID, B);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(MLP_HID, MLP_IN + 1), dim3(TB), 0, s, t.d1, X, ldx, t.G + MLP_W1,
                     t.G + MLP_B1, MLP_HID, MLP_IN, B);
}

hipError_t alloc_trainer(Trainer& t, int B) {
  t.B = B;
  hipError_t st = hipSuccess;
  auto a = [&](float** p, size_t n) {
    if (st == hipSuccess) st = hipMalloc((void**)p, sizeof(float) * std::max<size_t>(n, 1));
  };
  a(&t.P, MLP_NPARAM);
  a(&t.G, MLP_NPARAM);
  a(&t.m1, MLP_NPARAM);
  a(&t.m2, MLP_NPARAM);
  const size_t hb = (size_t)MLP_HID * B;
  a(&t.z1, hb);
  a(&t.h1, hb);
  a(&t.z2, hb);
  a(&t.h2, hb);
  a(&t.z3, hb);
  a(&t.h3, hb);
  a(&t.y, (size_t)MLP_OUT * B);
  a(&t.dy, (size_t)MLP_OUT * B);
  a(&t.d3, hb);
  a(&t.d2, hb);
  a(&t.d1, hb);
  if (st == hipSuccess) st = hipMalloc((void**)&t.loss, sizeof(double));
  return st;
}

void free_trainer(Trainer& t) {
  for (float* p : {t.P, t.G, t.m1, t.m2, t.z1, t.h1, t.z2, t.h2, t.z3, t.h3, t.y, t.dy, t.d3, t.d2, t.d1})
    if (p) (void)hipFree(p);
  if (t.loss) (void)hipFree(t.loss);
}

// Deterministic He-uniform initialisation from a seeded host hash.
void init_params(std::vector<float>& P, uint32_t seed) {
  P.assign(MLP_NPARAM, 0.f);
  auto rnd = [&](uint32_t i) {
    uint32_t h = seed * 0x9e3779b9u + (i + 1u) * 0x85ebca6bu;
    h = (h ^ (h >> 16)) * 0x7feb352du;
    h = (h ^ (h >> 15)) * 0x846ca68bu;
    h ^= h >> 16;
    return (float)(h >> 8) * (1.f / 16777216.f) * 2.f - 1.f;
  };
  const int offs[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
  const int fan_in[4] = {MLP_IN, MLP_HID, MLP_HID, MLP_HID};
  const int outs[4] = {MLP_HID, MLP_HID, MLP_HID, MLP_OUT};
  for (int l = 0; l < 4; ++l) {
    const float lim = std::sqrt(6.f / fan_in[l]) * (l == 3 ? 0.5f : 1.f);
    for (int q = 0; q < outs[l] * fan_in[l]; ++q) P[offs[l] + q] = lim * rnd((uint32_t)(offs[l] + q));
  }
}

}  // namespace

// Weight tiles in MFMA A-operand lane order: tile (layer, mt, ks), lane l
// holds W[16 mt + l%16][32 ks + 8 (l/16) + v], v = 0..7 (zero outside).
void pack_mlp(const float* P, const MlpNorm& norm, MlpDev& out) {
  std::memset(&out, 0, sizeof out);
  const int woff[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
  const int boff[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};
  const int K[4] = {MLP_IN, MLP_HID, MLP_HID, MLP_HID};
  const int Mo[4] = {MLP_HID, MLP_HID, MLP_HID, MLP_OUT};
  const int mts[4] = {4, 4, 4, 1};
  const int kss[4] = {1, 2, 2, 2};
  int tile = 0;
  for (int l = 0; l < 4; ++l) {
    for (int mt = 0; mt < mts[l]; ++mt)
      for (int ks = 0; ks < kss[l]; ++ks, ++tile)
        for (int lane = 0; lane < 64; ++lane)
          for (int v = 0; v < 8; ++v) {
            const int row = 16 * mt + lane % 16, col = 32 * ks + 8 * (lane / 16) + v;
            const float w = (row < Mo[l] && col < K[l]) ? P[woff[l] + row * K[l] + col] : 0.f;
            out.tiles[tile][lane][v] = (__bf16)w;
          }
    for (int r = 0; r < Mo[l]; ++r) out.bias[l * MLP_HID + r] = P[boff[l] + r];
  }
  out.norm = norm;
}

}  // namespace mpm

using namespace mpm;

namespace {
MlpNorm to_norm(const mpm_mlp_norm& n) {
  MlpNorm m;
  m.in_scale = n.in_scale;
  m.T_center = n.T_center;
  m.T_scale = n.T_scale;
  m.out_scale = n.out_scale;
  m.eps_max = n.eps_max;
  m.T_lo = n.T_lo;
  m.T_hi = n.T_hi;
  m.min_yield_ratio = n.min_yield_ratio;
  m.material = n.material;
  return m;
}

bool train_args_ok(const mpm_mlp_train_config* c) {
  return c && c->samples >= 64 && c->batch >= 64 && c->batch <= c->samples && c->epochs >= 1 && c->lr > 0 &&
         c->strain_sigma > 0 && c->T_hi > c->T_lo && c->near_yield_fraction >= 0 && c->near_yield_fraction <= 1;
}
}  // namespace

extern "C" {

int mpm_mlp_default_norm(const mpm_material* m, const mpm_mlp_train_config* c, mpm_mlp_norm* out) {
  if (!m || !c || !out || !(c->strain_sigma > 0) || !(c->T_hi > c->T_lo)) return MPM_ERR_INVALID_ARGUMENT;
  out->in_scale = 0.5f;
  out->T_center = 0.5f * (c->T_lo + c->T_hi);
  out->T_scale = 2.f / (c->T_hi - c->T_lo);
  out->out_scale = 1.f;
  out->eps_max = c->eps_max > 0 ? c->eps_max : 4.f * c->strain_sigma;
  out->T_lo = c->T_lo;
  out->T_hi = c->T_hi;
  out->min_yield_ratio = 1.f;
  out->material = -1;
  out->model = m->model;
  out->mu_over_E = m->mu / m->E;
  out->lambda_over_E = m->lambda / m->E;
  out->A = (m->model >= 10) ? m->th_A : m->pl_A;
  out->Tref = m->th_Tref;
  out->Thot = m->th_Thot;
  out->hotRatio = m->th_hotRatio;
  return MPM_OK;
}

// 0 when material m matches the training fingerprint of norm (E-normalised
// parameters within a relative 1e-5), else MPM_ERR_INVALID_ARGUMENT.
int mpm_mlp_check_material(const mpm_material* m, const mpm_mlp_norm* n) {
  if (!m || !n) return MPM_ERR_INVALID_ARGUMENT;
  auto close = [](float a, float b) { return std::fabs(a - b) <= 1e-5f * std::max(std::fabs(a), std::fabs(b)); };
  const float A = (m->model >= 10) ? m->th_A : m->pl_A;
  bool ok = m->model == n->model && m->E > 0 && close(m->mu / m->E, n->mu_over_E) &&
            close(m->lambda / m->E, n->lambda_over_E) && close(A, n->A);
  if (m->model == MPM_MODEL_THERMAL_J2)
    ok = ok && close(m->th_Tref, n->Tref) && close(m->th_Thot, n->Thot) && close(m->th_hotRatio, n->hotRatio);
  return ok ? MPM_OK : MPM_ERR_INVALID_ARGUMENT;
}

int mpm_mlp_samples(const mpm_material* m, const mpm_mlp_train_config* c, const mpm_mlp_norm* norm, uint32_t n,
                    uint32_t seed, float* X, float* Y) {
  if (!m || !c || !norm || (n && (!X || !Y))) return MPM_ERR_INVALID_ARGUMENT;
  if (m->model < 4) return MPM_ERR_UNSUPPORTED;
  if (!n) return MPM_OK;
  const DevMaterial dm = to_dev_material(*m);
  float *dX = nullptr, *dY = nullptr;
  hipError_t st = hipMalloc((void**)&dX, sizeof(float) * 7 * n);
  if (st == hipSuccess) st = hipMalloc((void**)&dY, sizeof(float) * 6 * n);
  if (st == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_samples, dim3((n + TB - 1) / TB), dim3(TB), 0, 0, seed, n, c->strain_sigma, c->T_lo,
                       c->T_hi, c->dt, c->near_yield_fraction, dm, to_norm(*norm), dX, dY);
    st = hipGetLastError();
  }
  std::vector<float> hx((size_t)7 * n), hy((size_t)6 * n);
  if (st == hipSuccess) st = hipMemcpy(hx.data(), dX, sizeof(float) * hx.size(), hipMemcpyDeviceToHost);
  if (st == hipSuccess) st = hipMemcpy(hy.data(), dY, sizeof(float) * hy.size(), hipMemcpyDeviceToHost);
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  if (st != hipSuccess) return MPM_ERR_HIP;
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 7; ++k) X[(size_t)i * 7 + k] = hx[(size_t)k * n + i];
    for (int k = 0; k < 6; ++k) Y[(size_t)i * 6 + k] = hy[(size_t)k * n + i];
  }
  return MPM_OK;
}

int mpm_mlp_train(const mpm_material* m, const mpm_mlp_train_config* c, const mpm_mlp_norm* norm, float* params,
                  float* loss_history) {
  if (!m || !train_args_ok(c) || !norm || !params) return MPM_ERR_INVALID_ARGUMENT;
  if (m->model < 4) return MPM_ERR_UNSUPPORTED;
  const uint32_t n = c->samples, B = c->batch;
  const DevMaterial dm = to_dev_material(*m);
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MPM_ERR_HIP;
  Trainer t{};
  float *dX = nullptr, *dY = nullptr;
  hipError_t st = alloc_trainer(t, (int)B);
  if (st == hipSuccess) st = hipMalloc((void**)&dX, sizeof(float) * 7 * n);
  if (st == hipSuccess) st = hipMalloc((void**)&dY, sizeof(float) * 6 * n);
  std::vector<float> P;
  init_params(P, c->seed);
  if (st == hipSuccess) st = hipMemcpyAsync(t.P, P.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s);
  if (st == hipSuccess) st = hipMemsetAsync(t.m1, 0, sizeof(float) * MLP_NPARAM, s);
  if (st == hipSuccess) st = hipMemsetAsync(t.m2, 0, sizeof(float) * MLP_NPARAM, s);
  if (st == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_samples, dim3((n + TB - 1) / TB), dim3(TB), 0, s, c->seed, n, c->strain_sigma, c->T_lo,
                       c->T_hi, c->dt, c->near_yield_fraction, dm, to_norm(*norm), dX, dY);
    st = hipGetLastError();
  }
  const uint32_t per_epoch = n / B;
  const uint64_t total = (uint64_t)per_epoch * c->epochs;
  const float b1 = 0.9f, b2 = 0.999f;
  uint64_t step = 0;
  for (uint32_t ep = 0; ep < c->epochs && st == hipSuccess; ++ep) {
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, s, t.loss);
    for (uint32_t it = 0; it < per_epoch && st == hipSuccess; ++it, ++step) {
      // Each epoch visits the batches in a different (seeded) rotation.
      const uint32_t bi = (uint32_t)((it + (uint64_t)ep * 7919u) % per_epoch);
      const float* Xb = dX + (size_t)bi * B;
      const float* Yb = dY + (size_t)bi * B;
      forward(s, t, Xb, (int)n, (int)B);
      backward(s, t, Xb, (int)n, Yb, (int)n, (int)B);
      const double frac = (double)step / (double)std::max<uint64_t>(total - 1, 1);
      const float lr = (float)(c->lr_final + 0.5 * (c->lr - c->lr_final) * (1 + std::cos(M_PI * frac)));
      const double k1 = (double)step + 1;
      const float c1 = (float)(1.0 / (1.0 - std::pow(b1, k1))), c2 = (float)(1.0 / (1.0 - std::pow(b2, k1)));
      hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + TB - 1) / TB), dim3(TB), 0, s, t.P, t.G, t.m1, t.m2, MLP_NPARAM,
                         lr, b1, b2, 1e-8f, c1, c2);
      st = hipGetLastError();
    }
    if (loss_history && st == hipSuccess) {
      double l = 0;
      st = hipMemcpyAsync(&l, t.loss, sizeof(double), hipMemcpyDeviceToHost, s);
      if (st == hipSuccess) st = hipStreamSynchronize(s);
      loss_history[ep] = (float)(l / per_epoch);
    }
  }
  if (st == hipSuccess) st = hipMemcpyAsync(params, t.P, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s);
  if (st == hipSuccess) st = hipStreamSynchronize(s);
  free_trainer(t);
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  (void)hipStreamDestroy(s);
  return st == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

int mpm_mlp_loss_grad(const float* params, uint32_t n, const float* X, const float* Y, float* grad, float* loss) {
  if (!params || !n || !X || !Y || !grad) return MPM_ERR_INVALID_ARGUMENT;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MPM_ERR_HIP;
  Trainer t{};
  float *dX = nullptr, *dY = nullptr;
  hipError_t st = alloc_trainer(t, (int)n);
  if (st == hipSuccess) st = hipMalloc((void**)&dX, sizeof(float) * 7 * n);
  if (st == hipSuccess) st = hipMalloc((void**)&dY, sizeof(float) * 6 * n);
  std::vector<float> hx((size_t)7 * n), hy((size_t)6 * n);
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 7; ++k) hx[(size_t)k * n + i] = X[(size_t)i * 7 + k];
    for (int k = 0; k < 6; ++k) hy[(size_t)k * n + i] = Y[(size_t)i * 6 + k];
  }
  if (st == hipSuccess) st = hipMemcpyAsync(dX, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice, s);
  if (st == hipSuccess) st = hipMemcpyAsync(dY, hy.data(), sizeof(float) * hy.size(), hipMemcpyHostToDevice, s);
  if (st == hipSuccess) st = hipMemcpyAsync(t.P, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s);
  if (st == hipSuccess) {
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, s, t.loss);
    forward(s, t, dX, (int)n, (int)n);
    backward(s, t, dX, (int)n, dY, (int)n, (int)n);
    st = hipGetLastError();
  }
  double l = 0;
  if (st == hipSuccess) st = hipMemcpyAsync(grad, t.G, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s);
  if (st == hipSuccess) st = hipMemcpyAsync(&l, t.loss, sizeof(double), hipMemcpyDeviceToHost, s);
  if (st == hipSuccess) st = hipStreamSynchronize(s);
  if (loss) *loss = (float)l;
  free_trainer(t);
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  (void)hipStreamDestroy(s);
  return st == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

int mpm_mlp_infer(const float* params, uint32_t n, const float* X, float* Y) {
  if (!params || (n && (!X || !Y))) return MPM_ERR_INVALID_ARGUMENT;
  if (!n) return MPM_OK;
  MlpDev* host = new MlpDev;
  MlpNorm nrm{};
  pack_mlp(params, nrm, *host);
  MlpDev* dM = nullptr;
  float *dX = nullptr, *dY = nullptr;
  hipError_t st = hipMalloc((void**)&dM, sizeof(MlpDev));
  if (st == hipSuccess) st = hipMalloc((void**)&dX, sizeof(float) * 7 * n);
  if (st == hipSuccess) st = hipMalloc((void**)&dY, sizeof(float) * 6 * n);
  std::vector<float> hx((size_t)7 * n), hy((size_t)6 * n);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 7; ++k) hx[(size_t)k * n + i] = X[(size_t)i * 7 + k];
  if (st == hipSuccess) st = hipMemcpy(dM, host, sizeof(MlpDev), hipMemcpyHostToDevice);
  if (st == hipSuccess) st = hipMemcpy(dX, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice);
  if (st == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + 63) / 64), dim3(64), 0, 0, dM, n, dX, dY);
    st = hipGetLastError();
  }
  if (st == hipSuccess) st = hipMemcpy(hy.data(), dY, sizeof(float) * hy.size(), hipMemcpyDeviceToHost);
  if (st == hipSuccess)
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 6; ++k) Y[(size_t)i * 6 + k] = hy[(size_t)k * n + i];
  if (dM) (void)hipFree(dM);
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  delete host;
  return st == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

int mpm_surrogate_return_batch(const mpm_material* m, const float* params, const mpm_mlp_norm* norm, uint32_t n,
                               float dt, const float* F, const float* ep, const float* T, float* F_out, float* ep_out,
                               float* tau_out, int32_t* branch) {
  if (!m || !params || !norm || (n && (!F || !ep || !T || !F_out || !ep_out || !tau_out || !branch)))
    return MPM_ERR_INVALID_ARGUMENT;
  if (!(m->model == MPM_MODEL_J2 || (m->model == MPM_MODEL_THERMAL_J2 && m->th_H == 0.f)))
    return MPM_ERR_UNSUPPORTED;
  if (mpm_mlp_check_material(m, norm) != MPM_OK) return MPM_ERR_INVALID_ARGUMENT;
  if (!n) return MPM_OK;
  MlpDev* host = new MlpDev;
  pack_mlp(params, to_norm(*norm), *host);
  const DevMaterial dm = to_dev_material(*m);
  MlpDev* dM = nullptr;
  float *dF = nullptr, *dep = nullptr, *dT = nullptr, *dFo = nullptr, *depo = nullptr, *dtau = nullptr;
  int32_t* dbr = nullptr;
  hipError_t st = hipMalloc((void**)&dM, sizeof(MlpDev));
  auto a = [&](float** p, size_t k) {
    if (st == hipSuccess) st = hipMalloc((void**)p, sizeof(float) * k);
  };
  a(&dF, 9ull * n);
  a(&dep, n);
  a(&dT, n);
  a(&dFo, 9ull * n);
  a(&depo, n);
  a(&dtau, 6ull * n);
  if (st == hipSuccess) st = hipMalloc((void**)&dbr, sizeof(int32_t) * n);
  if (st == hipSuccess) st = hipMemcpy(dM, host, sizeof(MlpDev), hipMemcpyHostToDevice);
  if (st == hipSuccess) st = hipMemcpy(dF, F, sizeof(float) * 9 * n, hipMemcpyHostToDevice);
  if (st == hipSuccess) st = hipMemcpy(dep, ep, sizeof(float) * n, hipMemcpyHostToDevice);
  if (st == hipSuccess) st = hipMemcpy(dT, T, sizeof(float) * n, hipMemcpyHostToDevice);
  if (st == hipSuccess) {
    hipLaunchKernelGGL(k_sur_batch, dim3((n + 63) / 64), dim3(64), 0, 0, dM, dm, n, dt, dF, dep, dT, dFo, depo, dtau,
                       dbr);
    st = hipGetLastError();
  }
  if (st == hipSuccess) st = hipMemcpy(F_out, dFo, sizeof(float) * 9 * n, hipMemcpyDeviceToHost);
  if (st == hipSuccess) st = hipMemcpy(ep_out, depo, sizeof(float) * n, hipMemcpyDeviceToHost);
  if (st == hipSuccess) st = hipMemcpy(tau_out, dtau, sizeof(float) * 6 * n, hipMemcpyDeviceToHost);
  if (st == hipSuccess) st = hipMemcpy(branch, dbr, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
  for (void* p : {(void*)dM, (void*)dF, (void*)dep, (void*)dT, (void*)dFo, (void*)depo, (void*)dtau, (void*)dbr})
    if (p) (void)hipFree(p);
  delete host;
  return st == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
