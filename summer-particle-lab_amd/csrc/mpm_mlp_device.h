// Learned constitutive surrogate: network constants, the device record the
// fused step kernel reads, and the BF16 MFMA inference of one wave.
//
// The network maps the trial elastic logarithmic strain (6 components) and the
// temperature to the deviatoric part of the return-mapped Kirchhoff stress
// (6 components): 7 -> 64 -> 64 -> 64 -> 6 with hard-swish. The volumetric part is
// not learned: a J2 return leaves tr(log strain) unchanged, so the pressure
// K tr(eps) is computed exactly. Inference runs 16x16x32 BF16 MFMAs with FP32
// accumulation; activations between layers pass through LDS as BF16.
// No counterpart in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.h"

namespace mpm {

constexpr int MLP_IN = 7, MLP_HID = 64, MLP_OUT = 6;
// Packed FP32 parameters: W1[64x7], b1[64], W2[64x64], b2[64], W3[64x64],
// b3[64], W4[6x64], b4[6] (weights output-major).
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_HID * MLP_IN, MLP_W2 = MLP_B1 + MLP_HID,
              MLP_B2 = MLP_W2 + MLP_HID * MLP_HID, MLP_W3 = MLP_B2 + MLP_HID, MLP_B3 = MLP_W3 + MLP_HID * MLP_HID,
              MLP_W4 = MLP_B3 + MLP_HID, MLP_B4 = MLP_W4 + MLP_OUT * MLP_HID, MLP_NPARAM = MLP_B4 + MLP_OUT;
// MFMA A-operand tiles (16 output rows x 32 inputs) per layer: layer 1 has
// 4 row tiles x 1 k-step, layers 2 and 3 4 x 2, layer 4 1 x 2.
constexpr int MLP_TILES = 4 + 8 + 8 + 2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Input/output scaling and the training envelope.
struct MlpNorm {
  float in_scale;   // scale of the strain in units of the yield strain at T
  float T_center, T_scale;
  float out_scale;  // output unit over y(T) E (y the E-normalised yield stress)
  float eps_max;    // envelope: |eps_ij| <= eps_max
  float T_lo, T_hi; // envelope: T_lo <= T <= T_hi
  float min_yield_ratio;  // envelope: q >= min_yield_ratio * yield
  int32_t material; // material index the surrogate replaces (-1: none)
};

// Device record: BF16 weight tiles in MFMA A-operand lane order and FP32
// biases (padded to 16-row tiles).
struct MlpDev {
  bf16x8 tiles[MLP_TILES][64];
  float bias[3 * MLP_HID + 16];
  MlpNorm norm;
};

// Hard-swish z min(max(z/6 + 1/2, 0), 1): an FMA, a clamp and a multiply.
__device__ __forceinline__ float hswish_f(float z) {
  return z * __builtin_amdgcn_fmed3f(z * (1.f / 6.f) + 0.5f, 0.f, 1.f);
}

// Activation scratch of one wave: 64 particle rows x 72 BF16 (144 B rows).
constexpr int MLP_ACT_STRIDE = 72;
constexpr int MLP_WAVE_LDS = 64 * MLP_ACT_STRIDE * 2;

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Network forward for the 64 particles of this wave (one per lane). Lanes
// with use == false still take part in the MFMAs (their rows are zero) and
// get zeros back. All 64 lanes must be active.
//
// MFMA 16x16x32 BF16 operand layout (lane l, g = l / 16, c = l % 16): A holds
// row c, k = 8 g .. 8 g + 7; B holds column c, k = 8 g .. 8 g + 7; the
// accumulator holds rows 4 g .. 4 g + 3 of column c. Here A is a 16-row
// weight tile, B a 32-feature slice of 16 particles, so a lane's accumulator
// is 4 consecutive features of particle 16 nt + c. Each layer first reads its
// B operands for all 4 particle tiles into registers, so its outputs can
// overwrite the activations in place, and then loads every weight tile once.
__device__ __forceinline__ void mlp_forward_wave(const MlpDev* __restrict__ M, bool use, const float x[7],
                                                 float y[6], unsigned char* wave_lds) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  __bf16* act = reinterpret_cast<__bf16*>(wave_lds);
  {
    bf16x8 in;
#pragma unroll
    for (int v = 0; v < 8; ++v) in[v] = (__bf16)((use && v < 7) ? x[v] : 0.f);
    *reinterpret_cast<bf16x8*>(act + lane * MLP_ACT_STRIDE) = in;
  }
  wave_sync_lds();
  const bf16x8 zero8 = {};
  int tile = 0;
#pragma unroll
  for (int layer = 0; layer < 4; ++layer) {
    const int ksteps = layer == 0 ? 1 : 2;
    const int mtiles = layer == 3 ? 1 : 4;
    bf16x8 B[4][2];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const __bf16* row = act + (nt * 16 + c) * MLP_ACT_STRIDE;
      if (layer == 0) {
        B[nt][0] = g == 0 ? *reinterpret_cast<const bf16x8*>(row) : zero8;
      } else {
        B[nt][0] = *reinterpret_cast<const bf16x8*>(row + 8 * g);
        B[nt][1] = *reinterpret_cast<const bf16x8*>(row + 32 + 8 * g);
      }
    }
    wave_sync_lds();
#pragma unroll
    for (int mt = 0; mt < mtiles; ++mt) {
      bf16x8 A[2];
#pragma unroll
      for (int ks = 0; ks < ksteps; ++ks) A[ks] = M->tiles[tile + mt * ksteps + ks][lane];
      f32x4 bias;
#pragma unroll
      for (int r = 0; r < 4; ++r) bias[r] = M->bias[layer * MLP_HID + mt * 16 + 4 * g + r];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = bias;
#pragma unroll
        for (int ks = 0; ks < ksteps; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks], B[nt][ks], acc, 0, 0, 0);
        __bf16* row = act + (nt * 16 + c) * MLP_ACT_STRIDE;
        if (layer < 3) {
          typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
          bf16x4 h;
#pragma unroll
          for (int r = 0; r < 4; ++r) h[r] = (__bf16)hswish_f(acc[r]);
          *reinterpret_cast<bf16x4*>(row + mt * 16 + 4 * g) = h;
        } else if (g < 2) {
          // Outputs 4 g .. 4 g + 3 as FP32 in the first 32 bytes of the row.
          *reinterpret_cast<f32x4*>(reinterpret_cast<float*>(row) + 4 * g) = acc;
        }
      }
    }
    tile += mtiles * ksteps;
    wave_sync_lds();
  }
  const float* mine = reinterpret_cast<const float*>(act + lane * MLP_ACT_STRIDE);
#pragma unroll
  for (int k = 0; k < 6; ++k) y[k] = use ? mine[k] : 0.f;
  wave_sync_lds();
}

// ---- SVD-free strain algebra of the surrogate branch ----------------------
// Symmetric 3x3 matrices as 6-vectors (xx, yy, zz, xy, xz, yz).

// c = a b for commuting symmetric a, b (the product is then symmetric).
__device__ __forceinline__ void sym_mul(const float a[6], const float b[6], float c[6]) {
  c[0] = a[0] * b[0] + a[3] * b[3] + a[4] * b[4];
  c[1] = a[3] * b[3] + a[1] * b[1] + a[5] * b[5];
  c[2] = a[4] * b[4] + a[5] * b[5] + a[2] * b[2];
  c[3] = a[0] * b[3] + a[3] * b[1] + a[4] * b[5];
  c[4] = a[0] * b[4] + a[3] * b[5] + a[4] * b[2];
  c[5] = a[3] * b[4] + a[1] * b[5] + a[5] * b[2];
}

// Left Hencky strain eps = 1/2 log(F F^T) without a decomposition:
// eps = ln(J)/3 I + artanh(Z), Z = (Bb - I)(Bb + I)^-1 = I - 2 (Bb + I)^-1,
// Bb = J^(-2/3) F F^T. The eigenvalues of Z are tanh of the deviatoric
// principal strains, and the series Z + Z^3/3 + Z^5/5 + Z^7/7 leaves a
// relative error below z^9/9 (6e-8 at |z| = 0.2).
__device__ __forceinline__ void log_strain_series(const M3& F, float eps[6], float& lnJ) {
  const float J = det3(F);
  lnJ = __logf(J);  // NaN for J <= 0: the envelope test then fails
  const float s = __expf(-0.6666667f * lnJ);
  float S[6];
  S[0] = s * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) + 1.f;
  S[1] = s * (F[3] * F[3] + F[4] * F[4] + F[5] * F[5]) + 1.f;
  S[2] = s * (F[6] * F[6] + F[7] * F[7] + F[8] * F[8]) + 1.f;
  S[3] = s * (F[0] * F[3] + F[1] * F[4] + F[2] * F[5]);
  S[4] = s * (F[0] * F[6] + F[1] * F[7] + F[2] * F[8]);
  S[5] = s * (F[3] * F[6] + F[4] * F[7] + F[5] * F[8]);
  const float c0 = S[1] * S[2] - S[5] * S[5], c1 = S[0] * S[2] - S[4] * S[4], c2 = S[0] * S[1] - S[3] * S[3];
  const float c3 = S[4] * S[5] - S[3] * S[2], c4 = S[3] * S[5] - S[4] * S[1], c5 = S[3] * S[4] - S[0] * S[5];
  const float m2 = -2.f * __builtin_amdgcn_rcpf(S[0] * c0 + S[3] * c3 + S[4] * c4);
  const float Z[6] = {1.f + m2 * c0, 1.f + m2 * c1, 1.f + m2 * c2, m2 * c3, m2 * c4, m2 * c5};
  float W[6], R[6], T[6];
  sym_mul(Z, Z, W);
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = W[k] * (1.f / 7.f) + (k < 3 ? 0.2f : 0.f);
  sym_mul(W, R, T);
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = T[k] + (k < 3 ? (1.f / 3.f) : 0.f);
  sym_mul(W, R, T);
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = T[k] + (k < 3 ? 1.f : 0.f);
  sym_mul(Z, R, eps);
  const float v = lnJ * (1.f / 3.f);
  eps[0] += v;
  eps[1] += v;
  eps[2] += v;
}

// exp(D) of a small symmetric D: Taylor series to 5th order in Horner form
// (relative error below |d|^6/720, 9e-8 at |d| = 0.2).
__device__ __forceinline__ void sym_exp_series(const float D[6], float E[6]) {
  float R[6], T[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = D[k] * 0.2f + (k < 3 ? 1.f : 0.f);
  sym_mul(D, R, T);
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = T[k] * 0.25f + (k < 3 ? 1.f : 0.f);
  sym_mul(D, R, T);
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = T[k] * (1.f / 3.f) + (k < 3 ? 1.f : 0.f);
  sym_mul(D, R, T);
#pragma unroll
  for (int k = 0; k < 6; ++k) R[k] = T[k] * 0.5f + (k < 3 ? 1.f : 0.f);
  sym_mul(D, R, T);
#pragma unroll
  for (int k = 0; k < 6; ++k) E[k] = T[k] + (k < 3 ? 1.f : 0.f);
}

// Frobenius norm of a symmetric 6-vector.
__device__ __forceinline__ float sym_norm(const float a[6]) {
  return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + 2.f * (a[3] * a[3] + a[4] * a[4] + a[5] * a[5]));
}

// Initial yield stress (E-normalised) of the rate-independent, non-hardening
// J2 models the surrogate covers: model 4 (constant) and model 10 (linear
// thermal softening, H = 0).
__device__ __forceinline__ float surrogate_yield(float T, const DevMaterial& m) {
  return m.model == 4 ? m.pl_A : thermal_strength(0.f, T, m);
}

// Return-mapped state from the deviatoric Kirchhoff stress dtau (absolute
// units) of the network: eps_new = ln(J)/3 I + dtau / (2 mu); F <- exp(eps_new
// - eps_tr) F (the two are coaxial); tau = dtau + K ln(J) I; the plastic
// strain grows by sqrt(2/3) (|dev eps_tr| - |dev eps_new|).
__device__ __forceinline__ float surrogate_apply(M3& F, const float eps[6], float lnJ, const float dtau[6],
                                                 const DevMaterial& m, float tau[6]) {
  const float mean = (eps[0] + eps[1] + eps[2]) * (1.f / 3.f);
  const float dev_tr[6] = {eps[0] - mean, eps[1] - mean, eps[2] - mean, eps[3], eps[4], eps[5]};
  const float i2mu = 0.5f / m.mu;
  float dev_new[6], D[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    dev_new[k] = dtau[k] * i2mu;
    D[k] = dev_new[k] - dev_tr[k];
  }
  float X[6];
  sym_exp_series(D, X);
  M3 Fn;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Fn[c] = X[0] * F[c] + X[3] * F[3 + c] + X[4] * F[6 + c];
    Fn[3 + c] = X[3] * F[c] + X[1] * F[3 + c] + X[5] * F[6 + c];
    Fn[6 + c] = X[4] * F[c] + X[5] * F[3 + c] + X[2] * F[6 + c];
  }
  F = Fn;
  const float p = (m.lambda + 0.6666667f * m.mu) * lnJ;
#pragma unroll
  for (int k = 0; k < 6; ++k) tau[k] = dtau[k] + (k < 3 ? p : 0.f);
  return fmaxf(0.f, 0.8164966f * (sym_norm(dev_tr) - sym_norm(dev_new)));
}

// Elastic trial state: tau = 2 mu eps + lambda tr(eps) I, F unchanged.
__device__ __forceinline__ void hencky_tau_of_strain(const float eps[6], const DevMaterial& m, float tau[6]) {
  const float l = m.lambda * (eps[0] + eps[1] + eps[2]);
#pragma unroll
  for (int k = 0; k < 6; ++k) tau[k] = 2.f * m.mu * eps[k] + (k < 3 ? l : 0.f);
}

// Branches of the surrogate return of one lane.
enum SurBranch : int { SUR_NONE = -1, SUR_NETWORK = 0, SUR_ELASTIC = 1, SUR_FALLBACK = 2 };

// Surrogate return of the trial state F (updated in place with ep and the
// stress tau) for a lane holding a particle of the surrogate's material
// (mine). Inside the training envelope (where the series logarithm is also
// accurate), trial states inside the yield surface take the closed-form
// elastic branch and states past min_yield_ratio times the yield stress the
// network; the rest return SUR_FALLBACK for the caller's plain return. Wave-uniform:
// all 64 lanes of the wave must call it together.
__device__ __forceinline__ int surrogate_return(const MlpDev* __restrict__ M, bool mine, M3& F, float& ep, float T,
                                                const DevMaterial& mat, float tau[6], unsigned char* wave_lds) {
  const MlpNorm nm = M->norm;
  float eps[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, lnJ = 0.f;
  bool elastic = false, inenv = false, deep = false;
  if (mine) {
    log_strain_series(F, eps, lnJ);
    const float mean = (eps[0] + eps[1] + eps[2]) * (1.f / 3.f);
    const float dv[6] = {eps[0] - mean, eps[1] - mean, eps[2] - mean, eps[3], eps[4], eps[5]};
    const float q = 2.4494897427831781f * mat.G * sym_norm(dv) * __expf(-lnJ);
    const float y = surrogate_yield(T, mat);
    elastic = q <= y;
    deep = q >= nm.min_yield_ratio * y;
    inenv = T >= nm.T_lo && T <= nm.T_hi;
#pragma unroll
    for (int k = 0; k < 6; ++k) inenv = inenv && fabsf(eps[k]) <= nm.eps_max;
  }
  const bool use = mine && deep && inenv;
  float dtau[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (__any(use)) {
    float xin[7], yout[6];
    // the strain in units of the yield strain at T, the output in units of
    // the yield stress at T: the return of a perfectly plastic J2 material is
    // the same map at every temperature in these units
    const float ys = surrogate_yield(T, mat);
    const float kin = nm.in_scale * 2.4494897427831781f * mat.G / ys;
#pragma unroll
    for (int k = 0; k < 6; ++k) xin[k] = eps[k] * kin;
    xin[6] = (T - nm.T_center) * nm.T_scale;
    mlp_forward_wave(M, use, xin, yout, wave_lds);
    const float unit = nm.out_scale * ys * mat.E;
#pragma unroll
    for (int k = 0; k < 6; ++k) dtau[k] = yout[k] * unit;
  }
  if (use) {
    ep += surrogate_apply(F, eps, lnJ, dtau, mat, tau);
    return SUR_NETWORK;
  }
  if (mine && inenv && elastic) {
    hencky_tau_of_strain(eps, mat, tau);
    return SUR_ELASTIC;
  }
  return mine ? SUR_FALLBACK : SUR_NONE;
}

}  // namespace mpm
