// Device-side constitutive math for the MLS-MPM engine (FP32).
//
// Each routine restates the algorithm of a reference function; the arithmetic
// is written for FP32 registers, not transcribed:
//   stencil weights      src/three/mpm.js:25-32
//   polar factor         src/three/math.js:16-24 (Newton R <- (R + R^-T)/2)
//   fixed-corotated tau  src/three/math.js:25-32
//   principal stretches  src/three/constitutive.js:22-42 (Jacobi on F^T F); the
//                        rotations use the Golub-Van Loan symmetric Schur
//                        formula instead of atan2/cos/sin
//   thermal J2 / creep   src/three/thermal-plasticity.js:12-40
//   J2 / Johnson-Cook    src/three/metal-plasticity.js:12-38
//   Hencky return stress src/three/constitutive.js:81-86
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

// Material record as the kernels see it: Lame parameters in absolute units
// plus the reference's E-normalised strengths. G = mu/E, lr = lambda/mu.
struct DevMaterial {
  int32_t model;
  float E, mu, lambda, G, lr;
  float th_Tref, th_Thot, th_hotRatio, th_A, th_H, th_rate, th_Q, th_n, th_stress, th_kelvin;
  float th_inv_span;  // 1 / (Thot - Tref)
  float pl_A, pl_B, pl_n, pl_C, pl_rate0, pl_Tref, pl_Tm, pl_m;
  float sn_ys;     // snow stretch yield strain
  float sd_alpha;  // sand: sqrt(2/3) 2 sin(phi) / (3 - sin(phi))
  // cohesive Drucker-Prager (8) / crushable foam (9), strengths / E
  float pr_c, pr_slope, pr_dil, pr_T, pr_alpha, pr_plateau, pr_gamma, pr_densify, pr_beta, pr_epsD;
  int32_t fr_on;   // tensile damage on
  float fr_onset;  // damage onset (log stretch)
  float fr_inv;    // 1 / (failure - onset)
};

// The model a particle steps with: a failed particle (damage 1) is
// cohesionless sand (src/three/fracture.js:17).
__device__ __forceinline__ int effective_model(const DevMaterial& m, float dmg) {
  return (m.fr_on && dmg >= 1.f) ? 3 : m.model;
}

struct M3 {
  float a[9];
  __device__ __forceinline__ float& operator[](int i) { return a[i]; }
  __device__ __forceinline__ float operator[](int i) const { return a[i]; }
};

__device__ __forceinline__ float det3(const M3& a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// Quadratic B-spline weights of the three nodes b, b+1, b+2 at offset t=q-b.
__device__ __forceinline__ void bspline(float t, float w[3]) {
  const float a = 1.5f - t, b = t - 1.0f, c = t - 0.5f;
  w[0] = 0.5f * a * a;
  w[1] = 0.75f - b * b;
  w[2] = 0.5f * c * c;
}

// Newton polar iteration R <- (R + R^-T)/2 (at most 8 rounds; FP32 stop
// threshold on the summed absolute change), identity when near singular.
__device__ __forceinline__ void polar_rotation(const M3& F, M3& R) {
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = F[k];
  for (int it = 0; it < 8; ++it) {
    const float c0 = R[4] * R[8] - R[5] * R[7];
    const float c1 = R[5] * R[6] - R[3] * R[8];
    const float c2 = R[3] * R[7] - R[4] * R[6];
    const float d = R[0] * c0 + R[1] * c1 + R[2] * c2;
    if (fabsf(d) < 1e-10f) {
#pragma unroll
      for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.f : 0.f;
      return;
    }
    const float id = __builtin_amdgcn_rcpf(d);
    float inv[9];  // inverse transpose = cofactor / det
    inv[0] = c0 * id;
    inv[1] = c1 * id;
    inv[2] = c2 * id;
    inv[3] = (R[2] * R[7] - R[1] * R[8]) * id;
    inv[4] = (R[0] * R[8] - R[2] * R[6]) * id;
    inv[5] = (R[1] * R[6] - R[0] * R[7]) * id;
    inv[6] = (R[1] * R[5] - R[2] * R[4]) * id;
    inv[7] = (R[2] * R[3] - R[0] * R[5]) * id;
    inv[8] = (R[0] * R[4] - R[1] * R[3]) * id;
    float diff = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const float v = 0.5f * (R[k] + inv[k]);
      diff += fabsf(v - R[k]);
      R[k] = v;
    }
    if (diff < 3e-7f) break;
  }
}

// Fixed-corotated Kirchhoff stress 2 mu (F - R) F^T + lambda (J - 1) J I,
// returned as the symmetric 6-vector (xx, yy, zz, xy, xz, yz).
__device__ __forceinline__ void fixed_corotated_tau(const M3& F, float mu, float lambda, float t[6]) {
  M3 R;
  polar_rotation(F, R);
  const float J = det3(F);
  float D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = F[k] - R[k];
  const float vol = lambda * (J - 1.f) * J;
  const float m2 = 2.f * mu;
  // (F-R) F^T is symmetric for the exact polar factor; symmetrise.
  float S[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) S[a * 3 + b] = D[a * 3] * F[b * 3] + D[a * 3 + 1] * F[b * 3 + 1] + D[a * 3 + 2] * F[b * 3 + 2];
  t[0] = m2 * S[0] + vol;
  t[1] = m2 * S[4] + vol;
  t[2] = m2 * S[8] + vol;
  t[3] = mu * (S[1] + S[3]);
  t[4] = mu * (S[2] + S[6]);
  t[5] = mu * (S[5] + S[7]);
}

// Cyclic Jacobi eigen-decomposition of the symmetric A = F^T F. On return
// V holds the eigenvectors in its columns and lam the eigenvalues.
__device__ __forceinline__ void jacobi_eig3(float a00, float a11, float a22, float a01, float a02, float a12,
                                            float V[9], float lam[3]) {
#pragma unroll
  for (int k = 0; k < 9; ++k) V[k] = (k % 4 == 0) ? 1.f : 0.f;
  for (int sweep = 0; sweep < 6; ++sweep) {
    const float off = fabsf(a01) + fabsf(a02) + fabsf(a12);
    const float scale = fabsf(a00) + fabsf(a11) + fabsf(a22);
    if (!(off > 1e-8f * scale)) break;
    // Each rotation zeroes one off-diagonal entry (p,q) with c, s from
    // t = sign(theta)/(|theta| + sqrt(theta^2+1)), theta = (aqq - app)/(2 apq).
#define MPM_JACOBI_ROT(app, aqq, apq, akp, akq, P, Q)                               \
  if (apq != 0.f) {                                                                 \
    const float theta = (aqq - app) * __builtin_amdgcn_rcpf(2.f * apq);             \
    const float tt = copysignf(__builtin_amdgcn_rcpf(fabsf(theta) + __builtin_amdgcn_sqrtf(theta * theta + 1.f)), theta); \
    const float c = __builtin_amdgcn_rsqf(tt * tt + 1.f), s = tt * c;               \
    app = app - tt * apq;                                                           \
    aqq = aqq + tt * apq;                                                           \
    apq = 0.f;                                                                      \
    const float x = akp, y = akq;                                                   \
    akp = c * x - s * y;                                                            \
    akq = s * x + c * y;                                                            \
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {                                 \
      const float vx = V[k * 3 + P], vy = V[k * 3 + Q];                             \
      V[k * 3 + P] = c * vx - s * vy;                                               \
      V[k * 3 + Q] = s * vx + c * vy;                                               \
    }                                                                               \
  }
    MPM_JACOBI_ROT(a00, a11, a01, a02, a12, 0, 1)
    MPM_JACOBI_ROT(a00, a22, a02, a01, a12, 0, 2)
    MPM_JACOBI_ROT(a11, a22, a12, a01, a02, 1, 2)
#undef MPM_JACOBI_ROT
  }
  lam[0] = a00;
  lam[1] = a11;
  lam[2] = a22;
}

// Principal stretches of F: F = U diag(sigma) V^T with V from the Jacobi
// eigenvectors of F^T F, sigma = sqrt(max(1e-16, lambda)), U = F V / sigma.
__device__ __forceinline__ void svd3(const M3& F, float U[9], float V[9], float sigma[3]) {
  float a[6];
  a[0] = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  a[1] = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  a[2] = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  a[3] = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  a[4] = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  a[5] = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  float lam[3];
  jacobi_eig3(a[0], a[1], a[2], a[3], a[4], a[5], V, lam);
#pragma unroll
  for (int k = 0; k < 3; ++k) sigma[k] = sqrtf(fmaxf(1e-16f, lam[k]));
  const float inv[3] = {__builtin_amdgcn_rcpf(sigma[0]), __builtin_amdgcn_rcpf(sigma[1]),
                        __builtin_amdgcn_rcpf(sigma[2])};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      U[i * 3 + j] = (F[i * 3] * V[j] + F[i * 3 + 1] * V[3 + j] + F[i * 3 + 2] * V[6 + j]) * inv[j];
}

// Symmetric U diag(d) U^T as a 6-vector (xx, yy, zz, xy, xz, yz).
__device__ __forceinline__ void sym_from_principal(const float U[9], const float d[3], float t[6]) {
  t[0] = U[0] * U[0] * d[0] + U[1] * U[1] * d[1] + U[2] * U[2] * d[2];
  t[1] = U[3] * U[3] * d[0] + U[4] * U[4] * d[1] + U[5] * U[5] * d[2];
  t[2] = U[6] * U[6] * d[0] + U[7] * U[7] * d[1] + U[8] * U[8] * d[2];
  t[3] = U[0] * U[3] * d[0] + U[1] * U[4] * d[1] + U[2] * U[5] * d[2];
  t[4] = U[0] * U[6] * d[0] + U[1] * U[7] * d[1] + U[2] * U[8] * d[2];
  t[5] = U[3] * U[6] * d[0] + U[4] * U[7] * d[1] + U[5] * U[8] * d[2];
}

// F = U diag(s) V^T.
__device__ __forceinline__ void compose(const float U[9], const float s[3], const float V[9], M3& F) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      F[i * 3 + j] = U[i * 3] * s[0] * V[j * 3] + U[i * 3 + 1] * s[1] * V[j * 3 + 1] + U[i * 3 + 2] * s[2] * V[j * 3 + 2];
}

// theta = (T - Tref) / (Thot - Tref) as the IEEE quotient rounds it, from
// the correctly rounded reciprocal of the span (set on the host): the
// product corrected once by its exact residual (Markstein) -- a multiply and
// two fused multiply-adds in place of a division. Checked bit for bit against
// the division for every float T in [-300, 3000] over six spans and 2e8
// random normal pairs. A span that is not positive (no reciprocal) divides.
__device__ __forceinline__ float thermal_strength(float ep, float T, const DevMaterial& m) {
  const float a = T - m.th_Tref, span = m.th_Thot - m.th_Tref;
  float q;
  if (m.th_inv_span != 0.f) {
    const float q0 = a * m.th_inv_span;
    q = fmaf(fmaf(-q0, span, a), m.th_inv_span, q0);
  } else {
    q = a / span;
  }
  const float theta = fminf(1.f, fmaxf(0.f, q));
  return (m.th_A + m.th_H * ep) * (1.f - (1.f - m.th_hotRatio) * theta);
}

__device__ __forceinline__ float creep_log_rate(float q, float T, const DevMaterial& m) {
  if (m.th_rate == 0.f || q <= 0.f) return -INFINITY;
  const float act = fmaxf(-60.f, fminf(60.f, m.th_Q / 8.314462618f *
                                                (1.f / (m.th_Tref + m.th_kelvin) - 1.f / fmaxf(1.f, T + m.th_kelvin))));
  return logf(m.th_rate) + act + m.th_n * logf(q / m.th_stress);
}

__device__ __forceinline__ float flow_stress(int model, float ep, float rate, float T, const DevMaterial& m) {
  if (model == 4) return m.pl_A;
  if (model == 5) return m.pl_A + m.pl_B * ep;
  const float rateFactor = 1.f + m.pl_C * logf(fmaxf(1.f, rate / m.pl_rate0));
  const float theta = fminf(1.f, fmaxf(0.f, (T - m.pl_Tref) / (m.pl_Tm - m.pl_Tref)));
  const float hard = m.pl_A + m.pl_B * (ep > 0.f ? powf(ep, m.pl_n) : 0.f);
  return hard * rateFactor * (model == 6 ? 1.f - powf(theta, m.pl_m) : 1.f);
}

// Radial return on logarithmic principal strains e (in place) for models 4-7,
// 10, 11. Returns the plastic strain increment. G = mu/E (E-normalised units).
__device__ __forceinline__ float return_map(float e[3], int model, float ep, float dt, float T,
                                            const DevMaterial& m) {
  const float tr = e[0] + e[1] + e[2];
  const float mean = tr * (1.f / 3.f);
  float dev[3] = {e[0] - mean, e[1] - mean, e[2] - mean};
  const float length = sqrtf(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const float J = fmaxf(1e-6f, expf(tr));
  const float q = 2.4494897427831781f * m.G * length / J;
  const float slope = 3.f * m.G / J;
  float inc = 0.f;
  if (model == 10) {
    const float factor = thermal_strength(0.f, T, m) / m.th_A;
    inc = fmaxf(0.f, (q - thermal_strength(ep, T, m)) / (slope + m.th_H * factor));
  } else if (model == 11) {
    if (dt > 0.f && m.th_rate > 0.f && q > 1e-20f) {
      float lo = 0.f, hi = q / slope;
      for (int k = 0; k < 32; ++k) {
        const float mid = 0.5f * (lo + hi);
        if (logf(mid / dt) < creep_log_rate(fmaxf(0.f, q - slope * mid), T, m))
          lo = mid;
        else
          hi = mid;
      }
      inc = 0.5f * (lo + hi);
    }
  } else {
    const float y0 = flow_stress(model, ep, 0.f, T, m);
    if (q > y0) {
      if (model <= 5) {
        inc = (q - y0) / (slope + (model == 5 ? m.pl_B : 0.f));
      } else {
        float lo = 0.f, hi = q / slope;
        const float idt = 1.f / fmaxf(dt, 1e-12f);
        for (int it = 0; it < 24; ++it) {
          const float mid = 0.5f * (lo + hi);
          if (q - slope * mid > flow_stress(model, ep + mid, mid * idt, T, m))
            lo = mid;
          else
            hi = mid;
        }
        inc = 0.5f * (lo + hi);
      }
    }
  }
  const float scale = fmaxf(0.f, 1.f - 1.2247448713915890f * inc / fmaxf(length, 1e-20f));
#pragma unroll
  for (int k = 0; k < 3; ++k) e[k] = mean + scale * dev[k];
  return inc;
}

// Snow's compaction hardening clamp(exp(10 (1 - Jp)), .25, 4)
// (src/three/constitutive.js:48).
__device__ __forceinline__ float snow_hardening(float jp) {
  return fminf(fmaxf(__expf(10.f * (1.f - jp)), .25f), 4.f);
}

// Neo-Hookean Kirchhoff stress mu (F F^T - I) + lambda ln J I
// (src/three/constitutive.js:88-93).
__device__ __forceinline__ void neo_hookean_tau(const M3& F, float mu, float lambda, float t[6]) {
  const float p = lambda * logf(fmaxf(det3(F), 1e-6f)) - mu;
  t[0] = mu * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) + p;
  t[1] = mu * (F[3] * F[3] + F[4] * F[4] + F[5] * F[5]) + p;
  t[2] = mu * (F[6] * F[6] + F[7] * F[7] + F[8] * F[8]) + p;
  t[3] = mu * (F[0] * F[3] + F[1] * F[4] + F[2] * F[5]);
  t[4] = mu * (F[0] * F[6] + F[1] * F[7] + F[2] * F[8]);
  t[5] = mu * (F[3] * F[6] + F[4] * F[7] + F[5] * F[8]);
}

// Crushable-foam hardening plateau + gamma r - densify ln(1 - r^beta),
// r = ep / epsD, locked (1e20) at full densification
// (src/three/pressure-plasticity.js:11-14).
__device__ __forceinline__ float foam_strength(float ep, const DevMaterial& m) {
  const float r = ep / m.pr_epsD;
  if (r >= 1.f) return 1e20f;
  return m.pr_plateau + m.pr_gamma * r - m.pr_densify * logf(fmaxf(1e-30f, 1.f - powf(r, m.pr_beta)));
}

// Return of the pressure-dependent models on the principal log stretches e
// (in place): 8 cohesive Drucker-Prager with tension cut-off and
// non-associated dilatancy, 9 Deshpande-Fleck foam (associated; bisection on
// the multiplier, 32 doublings then 28 halvings). Cauchy measures with
// J = exp(tr e) jp; G = (mu/E)/J, K = (lambda/E + 2/3 mu/E)/J, which are the
// reference's .4/J and (2/3)/J at mu = lambda = .4 E. jp takes the plastic
// volume change. Returns the equivalent plastic strain increment.
// src/three/pressure-plasticity.js:15-45
__device__ __forceinline__ float pressure_return(float e[3], int model, float& jp, float ep, const DevMaterial& m) {
  const float tr = e[0] + e[1] + e[2];
  float dev[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) dev[k] = e[k] - tr * (1.f / 3.f);
  const float len = sqrtf(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const float J = fmaxf(1e-6f, expf(tr) * jp);
  const float G = m.G / J, K = m.G * (m.lr + 2.f / 3.f) / J;
  const float mean = K * tr, q = 2.4494897427831781f * G * len;
  float mn = mean, qn = q, inc = 0.f;
  if (model == 8) {
    const float t = q * 0.57735026918962576f, c = m.pr_c, a = m.pr_slope, b = m.pr_dil, T = m.pr_T;
    if (mean > T && t + a * T <= c) {
      mn = T;
    } else if (t + a * mean > c) {
      const float dl = (t + a * mean - c) / (G + a * K * b);
      mn = mean - K * b * dl;
      qn = 1.7320508075688772f * fmaxf(0.f, t - G * dl);
      if (mn > T) {
        mn = T;
        qn = 1.7320508075688772f * fmaxf(0.f, c - a * T);
      }
    }
  } else {
    const float a2 = m.pr_alpha * m.pr_alpha, B = 1.f + a2 * (1.f / 9.f), iB = 1.f / B;
    if (sqrtf((q * q + a2 * mean * mean) * iB) > foam_strength(ep, m)) {
      // residual(eta) = hat(eta) - strength(ep + eta B hat(eta))
      auto residual = [&](float eta, float& qe, float& me, float& dl) {
        qe = q / (1.f + 3.f * G * eta);
        me = mean / (1.f + K * a2 * eta);
        const float hat = sqrtf((qe * qe + a2 * me * me) * iB);
        dl = eta * B * hat;
        return hat - foam_strength(ep + dl, m);
      };
      float qe, me, dl, lo = 0.f, hi = 1.f / G;
      for (int k = 0; k < 32 && residual(hi, qe, me, dl) > 0.f; ++k) hi *= 2.f;
      for (int k = 0; k < 28; ++k) {
        const float mid = 0.5f * (lo + hi);
        if (residual(mid, qe, me, dl) > 0.f)
          lo = mid;
        else
          hi = mid;
      }
      residual(0.5f * (lo + hi), qe, me, dl);
      qn = qe;
      mn = me;
      inc = dl;
    }
  }
  const float trn = mn / K, scale = q > 1e-20f ? qn / q : 1.f;
  float d2 = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float ne = trn * (1.f / 3.f) + scale * dev[k];
    d2 += (e[k] - ne) * (e[k] - ne);
    e[k] = ne;
  }
  if (model == 8) inc = 0.81649658092772603f * sqrtf(d2);
  jp *= expf(tr - trn);
  return inc;
}

// Hencky strain eps = 1/2 log(F F^T) without a decomposition, for F near a
// rotation: with G = F - I (exact for diagonal entries in [0.5, 2]) and
// X = F F^T - I = G + G^T + G G^T (no digits cancel), log(I + X) is the
// series X - X^2/2 + X^3/3 - X^4/4 + X^5/5, by Horner steps whose factors
// are polynomials in X (so every product is symmetric). For |X|_F <= 1/32
// the truncation error is below (1/32)^6 / 6 < 2e-10 (shorter series for
// smaller strains, below). eps has F's left
// singular vectors and the log stretches as eigenvalues, so its trace and
// deviatoric Frobenius norm are those of the principal log strains. Returns
// false (eps unset) when |X|_F is above 1/32 or not finite. Symmetric
// layout {xx, yy, zz, xy, xz, yz}.
__device__ __forceinline__ bool hencky_small(const M3& F, float eps[6]) {
  const float g0 = F[0] - 1.f, g4 = F[4] - 1.f, g8 = F[8] - 1.f;
  const float g1 = F[1], g2 = F[2], g3 = F[3], g5 = F[5], g6 = F[6], g7 = F[7];
  float X[6];
  X[0] = 2.f * g0 + (g0 * g0 + g1 * g1 + g2 * g2);
  X[1] = 2.f * g4 + (g3 * g3 + g4 * g4 + g5 * g5);
  X[2] = 2.f * g8 + (g6 * g6 + g7 * g7 + g8 * g8);
  X[3] = (g1 + g3) + (g0 * g3 + g1 * g4 + g2 * g5);
  X[4] = (g2 + g6) + (g0 * g6 + g1 * g7 + g2 * g8);
  X[5] = (g5 + g7) + (g3 * g6 + g4 * g7 + g5 * g8);
  const float r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2] + 2.f * (X[3] * X[3] + X[4] * X[4] + X[5] * X[5]);
  const bool small = r2 <= 1.f / 1024.f;
  // Fewer terms where every active lane's strain is small enough that the
  // dropped ones are below FP32 rounding relative to X: |X|_F <= 2^-12 keeps
  // X - X^2/2 (the next term is |X|^2/3 <= 2^-24/3 of X), |X|_F <= 2^-7 up
  // to X^3/3 (|X|^3/4 <= 2^-23 of X). Decided per wave, so the lanes take one
  // path.
  const bool over12 = __ballot(small && r2 > 5.9604645e-8f) != 0ull;  // |X| > 2^-12
  const bool over7 = __ballot(small && r2 > 6.1035156e-5f) != 0ull;   // |X| > 2^-7
  if (!small) return false;
  auto mul = [&](const float A[6], float c, float R[6]) {
    R[0] = X[0] * A[0] + X[3] * A[3] + X[4] * A[4] + c;
    R[1] = X[3] * A[3] + X[1] * A[1] + X[5] * A[5] + c;
    R[2] = X[4] * A[4] + X[5] * A[5] + X[2] * A[2] + c;
    R[3] = X[0] * A[3] + X[3] * A[1] + X[4] * A[5];
    R[4] = X[0] * A[4] + X[3] * A[5] + X[4] * A[2];
    R[5] = X[3] * A[4] + X[1] * A[5] + X[5] * A[2];
  };
  float P[6], Q[6];
  if (over7) {
    // P = X/5 - I/4, then P <- X P + c I for c = 1/3, -1/2, 1; log = X P.
    const float P0[6] = {0.2f * X[0] - 0.25f, 0.2f * X[1] - 0.25f, 0.2f * X[2] - 0.25f,
                         0.2f * X[3], 0.2f * X[4], 0.2f * X[5]};
    mul(P0, 1.f / 3.f, Q);
    mul(Q, -0.5f, P);
    mul(P, 1.f, Q);
  } else if (over12) {
    // Q = X/3 - I/2, then X Q + I; log = X (X Q + I).
    const float Q0[6] = {X[0] * (1.f / 3.f) - 0.5f, X[1] * (1.f / 3.f) - 0.5f, X[2] * (1.f / 3.f) - 0.5f,
                         X[3] * (1.f / 3.f), X[4] * (1.f / 3.f), X[5] * (1.f / 3.f)};
    mul(Q0, 1.f, Q);
  } else {
    // log = X (I - X/2).
#pragma unroll
    for (int k = 0; k < 6; ++k) Q[k] = -0.5f * X[k] + (k < 3 ? 1.f : 0.f);
  }
  mul(Q, 0.f, P);
#pragma unroll
  for (int k = 0; k < 6; ++k) eps[k] = 0.5f * P[k];
  return true;
}

// Hencky stress 2 mu eps + lambda tr(eps) I of a symmetric strain.
__device__ __forceinline__ void hencky_stress(const float eps[6], const DevMaterial& m, float tau[6]) {
  const float tr = eps[0] + eps[1] + eps[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) tau[k] = 2.f * m.mu * eps[k] + m.lambda * tr;
#pragma unroll
  for (int k = 3; k < 6; ++k) tau[k] = 2.f * m.mu * eps[k];
}

// Elastic step of the radial-return models 4-7 and 10 (those whose return
// leaves the state alone when q <= the flow stress at zero increment), from
// the series strain: when the trial state is inside the yield surface the
// return keeps every principal strain, so F and ep stay and the stress is
// the Hencky stress of F's own strain. Same q, J and yield test as
// return_map. Returns false (nothing written) when the strain is too large
// for the series or the state yields; the caller then takes the full return.
__device__ __forceinline__ bool elastic_return(const M3& F, int model, float ep, float T, const DevMaterial& m,
                                               float tau[6]) {
  float eps[6];
  if (!hencky_small(F, eps)) return false;
  const float tr = eps[0] + eps[1] + eps[2];
  const float mean = tr * (1.f / 3.f);
  const float d0 = eps[0] - mean, d1 = eps[1] - mean, d2 = eps[2] - mean;
  // Only the yield test reads length and J, so the hardware square root and
  // exponential (a few ulp) serve, and q = sqrt(6) G length / J <= y is
  // tested as sqrt(6) G length <= y J (J > 0). A NaN fails either way.
  const float length =
      __builtin_amdgcn_sqrtf(d0 * d0 + d1 * d1 + d2 * d2 + 2.f * (eps[3] * eps[3] + eps[4] * eps[4] + eps[5] * eps[5]));
  const float J = fmaxf(1e-6f, __expf(tr));
  const float y = model == 10 ? thermal_strength(ep, T, m) : flow_stress(model, ep, 0.f, T, m);
  if (!(2.4494897427831781f * m.G * length <= y * J)) return false;
  hencky_stress(eps, m, tau);
  return true;
}

// Plastic projection of F (models 2-11): SVD, return on the principal
// stretches, recompose F, and the stress of the returned state on the
// returned principal directions as the next step's stress (6-vector):
// snow 2 mu (s - 1) s + lambda (J - 1) J times its hardening, the others
// Hencky 2 mu e + lambda tr(e) (src/three/constitutive.js:52-86). jp carries
// snow's Jp and the pressure models' plastic volume ratio, dil sand's
// dilation. `model` is the particle's effective model. PRESSURE = false
// compiles the pressure models (8, 9) out, for kernels launched only when no
// material uses them. Returns the plastic strain increment.
template <bool PRESSURE = true>
__device__ __forceinline__ float plastic_project(M3& F, int model, float ep, float& jp, float& dil, float dt,
                                                 float T, const DevMaterial& m, float tau[6]) {
  if ((model >= 4 && model <= 7) || model == 10) {
    if (elastic_return(F, model, ep, T, m, tau)) return 0.f;
  }
  float U[9], V[9], s[3];
  svd3(F, U, V, s);
  float sp[3];
  float inc;
  if (model == 2) {
    // Snow: clamp the stretches to [1 - ys, 1 + 0.3 ys]; Jp takes the
    // clamped-away volume (src/three/constitutive.js:56-58).
#pragma unroll
    for (int k = 0; k < 3; ++k) sp[k] = fminf(fmaxf(s[k], 1.f - m.sn_ys), 1.f + m.sn_ys * .3f);
    jp = fminf(fmaxf(jp * (s[0] * s[1] * s[2]) / (sp[0] * sp[1] * sp[2]), .6f), 20.f);
  } else if (model == 3) {
    // Sand: non-cohesive logarithmic Drucker-Prager with a dilation history
    // (src/three/constitutive.js:65-73); (3 lambda + 2 mu) / (2 mu) is the
    // reference's 2.5 at lambda = mu.
    float e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) e[k] = logf(fmaxf(s[k], 1e-6f));
    const float tr = e[0] + e[1] + e[2] + dil;
    if (tr >= 0.f) {
      sp[0] = sp[1] = sp[2] = 1.f;
      dil = tr;
    } else {
      dil = 0.f;
      float dv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) dv[k] = e[k] - tr * (1.f / 3.f);
      const float len = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      const float delta = fmaxf(0.f, len + (1.5f * m.lr + 1.f) * tr * m.sd_alpha);
      const float f = delta / (len != 0.f ? len : 1.f);
#pragma unroll
      for (int k = 0; k < 3; ++k) sp[k] = expf(e[k] - f * dv[k]);
    }
  } else {
    float e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) e[k] = logf(fmaxf(s[k], 1e-6f));
    if (PRESSURE && (model == 8 || model == 9))
      inc = pressure_return(e, model, jp, ep, m);
    else
      inc = return_map(e, model, ep, dt, T, m);
#pragma unroll
    for (int k = 0; k < 3; ++k) sp[k] = expf(e[k]);
  }
  if (model == 2 || model == 3) {
    const float l0 = logf(s[0] / sp[0]), l1 = logf(s[1] / sp[1]), l2 = logf(s[2] / sp[2]);
    inc = sqrtf(l0 * l0 + l1 * l1 + l2 * l2);
  }
  compose(U, sp, V, F);
  float d[3];
  if (model == 2) {
    const float J = sp[0] * sp[1] * sp[2], h = snow_hardening(jp);
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = h * (2.f * m.mu * (sp[k] - 1.f) * sp[k] + m.lambda * (J - 1.f) * J);
  } else {
    const float e0 = logf(sp[0]), e1 = logf(sp[1]), e2 = logf(sp[2]);
    const float tr = e0 + e1 + e2;
    d[0] = 2.f * m.mu * e0 + m.lambda * tr;
    d[1] = 2.f * m.mu * e1 + m.lambda * tr;
    d[2] = 2.f * m.mu * e2 + m.lambda * tr;
  }
  sym_from_principal(U, d, tau);
  return inc;
}

// Compatibility form for the models without history state (Jp = 1, no
// dilation).
__device__ __forceinline__ float plastic_project(M3& F, float ep, float dt, float T, const DevMaterial& m,
                                                 float tau[6]) {
  float jp = 1.f, dil = 0.f;
  return plastic_project(F, m.model, ep, jp, dil, dt, T, m, tau);
}

// Hencky stress 2 mu log(sigma) + lambda tr from F's own decomposition (the
// first-step stress of a plastic model before any return has been cached).
__device__ __forceinline__ void hencky_tau(const M3& F, const DevMaterial& m, float tau[6]) {
  float eps[6];
  if (hencky_small(F, eps)) {
    hencky_stress(eps, m, tau);
    return;
  }
  float U[9], V[9], s[3];
  svd3(F, U, V, s);
  float e[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) e[k] = logf(fmaxf(s[k], 1e-6f));
  const float tr = e[0] + e[1] + e[2];
  float d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = 2.f * m.mu * e[k] + m.lambda * tr;
  sym_from_principal(U, d, tau);
}

// Step stress from F alone (no cached return): constitutiveStress with the
// material's moduli, snow times its hardening (src/three/mpm.js:33-37,
// src/three/constitutive.js:87-99).
__device__ __forceinline__ void elastic_tau(const M3& F, int model, const DevMaterial& m, float jp, float tau[6]) {
  if (model == 1) {
    neo_hookean_tau(F, m.mu, m.lambda, tau);
  } else if (model == 0 || model == 2) {
    const float h = model == 2 ? snow_hardening(jp) : 1.f;
    fixed_corotated_tau(F, m.mu * h, m.lambda * h, tau);
  } else {
    hencky_tau(F, m, tau);
  }
}

__device__ __forceinline__ void elastic_tau(const M3& F, const DevMaterial& m, float jp, float tau[6]) {
  elastic_tau(F, m.model, m, jp, tau);
}

// Tensile-strain damage after the F update: the largest principal log
// stretch between onset and failure raises the damage (never lowers it); at
// 1 the particle fails and F resets to the identity. Returns true on failure
// in this call. src/three/fracture.js:11-18
__device__ __forceinline__ bool update_damage(M3& F, float& dmg, const DevMaterial& m) {
  if (!m.fr_on || dmg >= 1.f) return false;
  float U[9], V[9], s[3];
  svd3(F, U, V, s);
  const float strain = fmaxf(0.f, logf(fmaxf(fmaxf(fmaxf(s[0], s[1]), s[2]), 1e-8f)));
  dmg = fmaxf(dmg, fminf(1.f, (strain - m.fr_onset) * m.fr_inv));
  if (dmg < 1.f) return false;
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = (k % 4 == 0) ? 1.f : 0.f;
  return true;
}

}  // namespace mpm
