// TEST INFRASTRUCTURE ONLY -- the oracle, never the product.
//
// FP64 CPU restatement of the reference MLS-MPM step of Summer-Particle-Lab
// (`MPM3` in src/three/mpm.js) and of every function it calls on the solid
// path, plus the seeding rules used to build its scenes. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
// and only as the checker (or as the timed CPU baseline). The product path in
// summer-particle-lab_amd/ never links or calls it.
//
// Each function cites the reference file:line it restates. Where the product
// generalises a reference constant (mu = lambda = 0.4 E, separating walls,
// gravity added by the caller), the oracle takes the same generalisation as a
// parameter and reduces to the literal reference arithmetic when that
// parameter holds the reference value (see `Part` and `Scene` below).
//
// Parity pins (what checks each function; tests/test_oracle_*.py):
//   svd3, compose           reconstruction/orthogonality on the reference's own
//                           matrices (tests/constitutive.test.mjs:10-16) and
//                           singular values vs LAPACK.
//   rotation, stress3       rotation invariance and shear signs
//                           (tests/three.test.mjs:8-13); polar factor vs the
//                           SVD polar factor; closed form for F=diag(s,1,1).
//   constitutiveStress      objectivity (tests/constitutive.test.mjs:17-25).
//   thermalReturn           yield consistency, isochoric flow, cooling, hot
//                           ratio 1, extreme T (tests/thermal-plasticity.test.mjs
//                           :12-20); closed-form n=1 creep (:28-31).
//   metalReturn, flowStress yield consistency, limiting cases
//                           (tests/metal-plasticity.test.mjs:12-42).
//   step                    exact uniform translation (tests/three.test.mjs:14-21,
//                           tests/discretization.test.mjs:38-46), mass and
//                           momentum identities, brute-force weights.
//   random, jitterPoint,    numpy restatement bit-for-bit, bounds and spacing
//   geometryPoints          (tests/sampling.test.mjs:9-25), counts and mass
//                           (tests/discretization.test.mjs:15-23).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- math.js
inline double det3(const double* a) {  // src/three/math.js:9
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
inline double clampd(double x, double a, double b) { return std::max(a, std::min(b, x)); }

// src/three/math.js:10-15; returns false where the reference returns null.
bool inverse_transpose(const double* a, double* out) {
  const double d = det3(a);
  if (std::fabs(d) < 1e-10) return false;
  out[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  out[1] = (a[5] * a[6] - a[3] * a[8]) / d;
  out[2] = (a[3] * a[7] - a[4] * a[6]) / d;
  out[3] = (a[2] * a[7] - a[1] * a[8]) / d;
  out[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  out[5] = (a[1] * a[6] - a[0] * a[7]) / d;
  out[6] = (a[1] * a[5] - a[2] * a[4]) / d;
  out[7] = (a[2] * a[3] - a[0] * a[5]) / d;
  out[8] = (a[0] * a[4] - a[1] * a[3]) / d;
  return true;
}

// Newton polar iteration R <- (R + R^-T)/2, at most 8 rounds, stop when the
// summed absolute change is below 1e-8; identity when singular.
// src/three/math.js:16-24
void rotation(const double* F, double* R) {
  for (int k = 0; k < 9; ++k) R[k] = F[k];
  double inv[9];
  for (int n = 0; n < 8; ++n) {
    if (!inverse_transpose(R, inv)) {
      for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.0 : 0.0;
      return;
    }
    double diff = 0;
    for (int k = 0; k < 9; ++k) {
      const double v = .5 * (R[k] + inv[k]);
      diff += std::fabs(v - R[k]);
      R[k] = v;
    }
    if (diff < 1e-8) break;
  }
}

// Fixed-corotated Kirchhoff stress 2mu(F-R)F^T + lambda(J-1)J I.
// src/three/math.js:25-32
void stress3(const double* F, double mu, double lambda, double* t) {
  double R[9];
  rotation(F, R);
  const double J = det3(F);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double v = 0;
      for (int k = 0; k < 3; ++k) v += 2 * mu * (F[a * 3 + k] - R[a * 3 + k]) * F[b * 3 + k];
      if (a == b) v += lambda * (J - 1) * J;
      t[a * 3 + b] = v;
    }
}

// ---------------------------------------------------------- constitutive.js
struct Svd {
  double u[9], v[9], sigma[3];
};

// Cyclic Jacobi diagonalisation of F^T F (at most 10 sweeps, pairs (0,1),
// (0,2), (1,2), skip threshold 1e-13(1+|app|+|aqq|)); sigma = sqrt(max(1e-16,
// lambda)); U = F V / sigma. src/three/constitutive.js:22-42
Svd svd3(const double* F) {
  Svd d;
  double a[9] = {0};
  double* v = d.v;
  for (int k = 0; k < 9; ++k) v[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) a[i * 3 + j] += F[k * 3 + i] * F[k * 3 + j];
  static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 10; ++sweep) {
    bool changed = false;
    for (const auto& pq : pairs) {
      const int p = pq[0], q = pq[1];
      const double apq = a[p * 3 + q];
      if (std::fabs(apq) < 1e-13 * (1 + std::fabs(a[p * 3 + p]) + std::fabs(a[q * 3 + q]))) continue;
      changed = true;
      const double theta = .5 * std::atan2(2 * apq, a[q * 3 + q] - a[p * 3 + p]);
      const double c = std::cos(theta), s = std::sin(theta);
      const double app = a[p * 3 + p], aqq = a[q * 3 + q];
      a[p * 3 + p] = c * c * app - 2 * s * c * apq + s * s * aqq;
      a[q * 3 + q] = s * s * app + 2 * s * c * apq + c * c * aqq;
      a[p * 3 + q] = a[q * 3 + p] = 0;
      for (int k = 0; k < 3; ++k) {
        if (k != p && k != q) {
          const double x = a[k * 3 + p], y = a[k * 3 + q];
          a[k * 3 + p] = a[p * 3 + k] = c * x - s * y;
          a[k * 3 + q] = a[q * 3 + k] = s * x + c * y;
        }
        const double x = v[k * 3 + p], y = v[k * 3 + q];
        v[k * 3 + p] = c * x - s * y;
        v[k * 3 + q] = s * x + c * y;
      }
    }
    if (!changed) break;
  }
  for (int k = 0; k < 3; ++k) d.sigma[k] = std::sqrt(std::max(1e-16, a[k * 3 + k]));
  for (int k = 0; k < 9; ++k) d.u[k] = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) d.u[i * 3 + j] += F[i * 3 + k] * v[k * 3 + j] / d.sigma[j];
  return d;
}

// U diag(sigma) V^T. src/three/constitutive.js:43-47
void compose(const Svd& d, const double* sigma, double* F) {
  for (int k = 0; k < 9; ++k) F[k] = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) F[i * 3 + j] += d.u[i * 3 + k] * sigma[k] * d.v[j * 3 + k];
}

inline double hardening(double jp) { return clampd(std::exp(10 * (1 - jp)), .25, 4); }  // constitutive.js:48

}  // namespace

extern "C" {

// ------------------------------------------------------------ parameter records
// Thermal J2 / creep record, strengths normalised by E; temperatures in the
// unit of the particle temperature. `kelvin_offset` is the constant the
// reference adds inside the Arrhenius term (273.15 for degC input).
// src/three/thermal-plasticity.js:3-4
typedef struct {
  double Tref, Thot, hotRatio, A, H, rate, Q, n, stress, kelvin_offset;
} OrcThermal;

// Johnson-Cook record, strengths normalised by E. src/three/metal-plasticity.js:3-4
typedef struct {
  double A, B, n, C, rate0, Tref, Tm, m;
} OrcPlastic;

// One part (material). `E` > 0 fixes Young's modulus for every particle of the
// part; E == 0 takes the reference's per-particle E = m/(64 V)*16000*stiffness
// (src/three/mpm.js:34). `G`, `L` are mu/E and lambda/E; the reference
// hard-codes G = L = 0.4 (src/three/mpm.js:37, thermal-plasticity.js:24).
// Pressure-dependent record (models 8, 9): Cauchy strengths / E, tension
// positive. src/three/pressure-plasticity.js:2-3
typedef struct {
  double cohesion, slope, dilatancy, tension, alpha, plateau, gamma, densify, beta, epsD;
} OrcPressure;

// Tensile-strain damage record. src/three/fracture.js:6-9
typedef struct {
  int32_t enabled;
  double onset, failure;
} OrcFracture;

typedef struct {
  double E, G, L;
  OrcThermal th;
  OrcPlastic pl;
  OrcPressure pr;
  OrcFracture fr;
} OrcPart;

// Caller-owned scene (all arrays FP64, per-particle 3x3 tensors row-major,
// 9 per particle as in src/three/mpm.js:9-13). Grid arrays hold N^3 nodes.
typedef struct {
  int count;
  double *x, *y, *z, *vx, *vy, *vz;
  const double *mass, *volume, *temperature;
  double* damage;                       // may be NULL (no fracture)
  const double *yield_strain, *friction;  // may be NULL (defaults .025, 35)
  int32_t* model;                       // a failed particle becomes sand (3)
  const int32_t* part;
  const uint8_t* pinned;
  double *F, *C, *Jp, *dilation, *stress, *plastic, *tau, *cachedTau;
  uint8_t* cacheValid;
  int nparts;
  const OrcPart* parts;
  double stiffness;
  double lower[3], upper[3], origin[3], H;
  int N;
  double *gmass, *gv;   // N^3 and 3 N^3
  uint8_t* gfixed;      // N^3
  // generalisations (reference values: 64, {0,0,0}, {0,0,0,0,0,0}, 1e-12)
  double vol_scale;     // particle volume = vol_scale * volume[p]
  double grid_gravity[3];
  int32_t wall_mode[6]; // faces -x,+x,-y,+y,-z,+z: 0 separating, 1 sticky, 2 slip
  double mass_eps;
  // the platen (s.press, src/three/mpm.js:97-101): nodes at or above
  // press_y - H/2 not held by a pinned particle move down at most at
  // press_velocity; press_impulse accumulates (before - after) mass
  int32_t press_enabled;
  double press_y, press_velocity, press_impulse;
} OrcScene;

// -------------------------------------------------------------- sampling.js
// uint32 hash of (seed, index, attempt, axis) mapped to [-1, 1).
// src/three/sampling.js:8-12
double orc_random(uint32_t seed, int32_t index, int32_t attempt, int32_t axis) {
  uint32_t h = seed + (uint32_t)(index + 1) * 0x9e3779b9u + (uint32_t)(attempt * 3 + axis + 1) * 0x85ebca6bu;
  h = (h ^ (h >> 16)) * 0x7feb352du;
  h = (h ^ (h >> 15)) * 0x846ca68bu;
  return (double)(h ^ (h >> 16)) / 4294967296.0 * 2 - 1;
}

// Bounded seeded jitter with face reflection and up to 8 retries.
// shape: 0 box, 1 ellipsoid, 2 cylinder. src/three/sampling.js:13-29
void orc_jitter_point(const double* p, int index, const double* lo, const double* hi, int shape,
                      int fixed, int enabled, double strength, uint32_t seed, double spacing,
                      double* out) {
  for (int a = 0; a < 3; ++a) out[a] = p[a];
  if (!enabled || strength == 0) return;
  const double amplitude = spacing * strength;
  for (int attempt = 0; attempt < 8; ++attempt) {
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = (fixed && a == 0) ? 0 : orc_random(seed, index, attempt, a);
    const double scale = amplitude / std::max(1.0, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    double c[3], q[3];
    for (int a = 0; a < 3; ++a) {
      double x = p[a] + d[a] * scale;
      if (x < lo[a]) x = 2 * lo[a] - x;
      if (x > hi[a]) x = 2 * hi[a] - x;
      c[a] = x;
      q[a] = (x - (lo[a] + hi[a]) / 2) / ((hi[a] - lo[a]) / 2);
    }
    const double r = shape == 1 ? q[0] * q[0] + q[1] * q[1] + q[2] * q[2] : q[0] * q[0] + q[2] * q[2];
    if (shape == 0 || r <= 1 + 1e-12) {
      for (int a = 0; a < 3; ++a) out[a] = c[a];
      return;
    }
  }
}

// ----------------------------------------------------------- discretization.js
// Background grid of a domain: origin snapped to the world grid minus two
// guard cells, nodes = ceil((upper-origin)/h)+3. `strict` applies the
// reference's validation (h in [4,16], <= 60 cells/axis, ...); returns 0 on
// success, -1 when the reference would throw. src/three/discretization.js:4-15
int orc_domain(const double* lower, const double* upper, double h, int n, int strict,
               double* origin, int32_t* nodes, int32_t* cells) {
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(lower[a]) || !std::isfinite(upper[a])) return -1;
  if (!std::isfinite(h) || h <= 0 || n < 1) return -1;
  if (strict && (h < 4 || h > 16 || n > 4 || h / n < 1 || h / n > 4)) return -1;
  for (int a = 0; a < 3; ++a) {
    const double size = upper[a] - lower[a];
    cells[a] = (int32_t)(std::ceil(upper[a] / h - 1e-10) - std::floor(lower[a] / h + 1e-10));
    if (size < h) return -1;
    if (strict && (lower[a] < -2400 || upper[a] > 2400 || cells[a] > 60)) return -1;
  }
  for (int a = 0; a < 3; ++a) {
    origin[a] = std::floor(lower[a] / h) * h - 2 * h;
    nodes[a] = (int32_t)(std::ceil((upper[a] - origin[a]) / h) + 3);
  }
  return 0;
}

// Lattice of an axis-aligned, unrotated geometry: points (k+1/2)*spacing with
// k from ceil(lo/s-.5-1e-10) to ceil(hi/s-.5-1e-10)-1, z outermost, kept when
// inside the shape. Returns the point count (or -1 if over capacity).
// src/three/discretization.js:33-48 (rotation = 0)
int64_t orc_geometry_points(const double* center, const double* size, int shape, double spacing,
                            int64_t capacity, double* out) {
  int64_t first[3], last[3];
  for (int a = 0; a < 3; ++a) {
    const double lo = center[a] - size[a] * .5, hi = center[a] + size[a] * .5;
    first[a] = (int64_t)std::ceil(lo / spacing - .5 - 1e-10);
    last[a] = (int64_t)std::ceil(hi / spacing - .5 - 1e-10) - 1;
  }
  int64_t count = 0;
  for (int64_t z = first[2]; z <= last[2]; ++z)
    for (int64_t y = first[1]; y <= last[1]; ++y)
      for (int64_t x = first[0]; x <= last[0]; ++x) {
        const double p[3] = {(x + .5) * spacing, (y + .5) * spacing, (z + .5) * spacing};
        double q[3];
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
          q[a] = 2 * (p[a] - center[a]) / size[a];
          if (std::fabs(q[a]) > 1 + 1e-10) inside = false;
        }
        if (inside && shape == 1) inside = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] <= 1 + 1e-10;
        if (inside && shape == 2) inside = q[0] * q[0] + q[2] * q[2] <= 1 + 1e-10;
        if (!inside) continue;
        if (count >= capacity) return -1;
        if (out)
          for (int a = 0; a < 3; ++a) out[count * 3 + a] = p[a];
        ++count;
      }
  return count;
}

// --------------------------------------------------------- constitutive laws
void orc_svd3(const double* F, double* u, double* v, double* sigma) {
  Svd d = svd3(F);
  std::memcpy(u, d.u, sizeof d.u);
  std::memcpy(v, d.v, sizeof d.v);
  std::memcpy(sigma, d.sigma, sizeof d.sigma);
}
void orc_compose(const double* u, const double* v, const double* sigma, double* F) {
  Svd d;
  std::memcpy(d.u, u, sizeof d.u);
  std::memcpy(d.v, v, sizeof d.v);
  compose(d, sigma, F);
}
void orc_rotation(const double* F, double* R) { rotation(F, R); }
void orc_stress3(const double* F, double mu, double lambda, double* t) { stress3(F, mu, lambda, t); }

// Thermal J2 strength (A + H ep)(1 - (1-hotRatio) theta).
// src/three/thermal-plasticity.js:12-15
double orc_thermal_strength(double ep, double T, const OrcThermal* p) {
  const double theta = std::max(0.0, std::min(1.0, (T - p->Tref) / (p->Thot - p->Tref)));
  return (p->A + p->H * ep) * (1 - (1 - p->hotRatio) * theta);
}

// Bounded Arrhenius/Norton log rate. src/three/thermal-plasticity.js:17-21
double orc_creep_log_rate(double q, double T, const OrcThermal* p) {
  if (p->rate == 0 || q <= 0) return -INFINITY;
  const double activation = std::max(
      -60.0, std::min(60.0, p->Q / 8.314462618 *
                                (1 / (p->Tref + p->kelvin_offset) - 1 / std::max(1.0, T + p->kelvin_offset))));
  return std::log(p->rate) + activation + p->n * std::log(q / p->stress);
}

// Closed-form thermal J2 radial return (model 10) or backward-Euler creep
// (model 11) on principal stretches. G = mu/E (0.4 in the reference).
// src/three/thermal-plasticity.js:22-40
void orc_thermal_return(const double* sigma, int model, double ep, double dt, double T,
                        const OrcThermal* p, double G, double* out_sigma, double* out_increment) {
  double e[3], dev[3];
  for (int k = 0; k < 3; ++k) e[k] = std::log(std::max(sigma[k], 1e-6));
  const double tr = e[0] + e[1] + e[2];
  for (int k = 0; k < 3; ++k) dev[k] = e[k] - tr / 3;
  const double length = std::sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double J = std::max(1e-6, std::exp(tr));
  const double q = std::sqrt(6.0) * G * length / J;
  const double slope = (G == .4 ? 1.2 : 3 * G) / J;
  double increment = 0;
  if (model == 10) {
    const double factor = orc_thermal_strength(0, T, p) / p->A;
    increment = std::max(0.0, (q - orc_thermal_strength(ep, T, p)) / (slope + p->H * factor));
  } else if (dt > 0 && p->rate > 0 && q > 1e-20) {
    double lo = 0, hi = q / slope;
    for (int k = 0; k < 32; ++k) {
      const double mid = (lo + hi) / 2;
      if (std::log(mid / dt) < orc_creep_log_rate(std::max(0.0, q - slope * mid), T, p))
        lo = mid;
      else
        hi = mid;
    }
    increment = (lo + hi) / 2;
  }
  const double scale = std::max(0.0, 1 - std::sqrt(1.5) * increment / std::max(length, 1e-20));
  for (int k = 0; k < 3; ++k) out_sigma[k] = std::exp(tr / 3 + scale * dev[k]);
  *out_increment = increment;
}

// J2 / Johnson-Cook flow stress. src/three/metal-plasticity.js:12-19
double orc_flow_stress(int model, double ep, double rate, double T, const OrcPlastic* p) {
  if (model == 4) return p->A;
  if (model == 5) return p->A + p->B * ep;
  const double rateFactor = 1 + p->C * std::log(std::max(1.0, rate / p->rate0));
  const double theta = std::max(0.0, std::min(1.0, (T - p->Tref) / (p->Tm - p->Tref)));
  return (p->A + p->B * std::pow(ep, p->n)) * rateFactor * (model == 6 ? 1 - std::pow(theta, p->m) : 1);
}

// J2 radial return with linear hardening (4, 5) or a 24-step bisection for
// rate/temperature dependence (6, 7). src/three/metal-plasticity.js:20-38
void orc_metal_return(const double* sigma, int model, double ep, double dt, double T,
                      const OrcPlastic* p, double G, double* out_sigma, double* out_increment) {
  double e[3], dev[3];
  for (int k = 0; k < 3; ++k) e[k] = std::log(std::max(sigma[k], 1e-6));
  const double trace = e[0] + e[1] + e[2];
  for (int k = 0; k < 3; ++k) dev[k] = e[k] - trace / 3;
  const double length = std::sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double J = std::max(1e-6, std::exp(trace));
  const double q = std::sqrt(6.0) * G * length / J;
  double increment = 0;
  if (q > orc_flow_stress(model, ep, 0, T, p)) {
    const double slope = (G == .4 ? 1.2 : 3 * G) / J;
    if (model <= 5) {
      increment = (q - orc_flow_stress(model, ep, 0, T, p)) / (slope + (model == 5 ? p->B : 0));
    } else {
      double lo = 0, hi = q / slope;
      for (int it = 0; it < 24; ++it) {
        const double mid = (lo + hi) / 2;
        if (q - slope * mid > orc_flow_stress(model, ep + mid, mid / std::max(dt, 1e-12), T, p))
          lo = mid;
        else
          hi = mid;
      }
      increment = (lo + hi) / 2;
    }
  }
  const double scale = std::max(0.0, 1 - std::sqrt(1.5) * increment / std::max(length, 1e-20));
  for (int k = 0; k < 3; ++k) out_sigma[k] = std::exp(trace / 3 + scale * dev[k]);
  *out_increment = increment;
}

// Crushable-foam hardening curve; 1e20 at and beyond full densification.
// src/three/pressure-plasticity.js:11-14
double orc_foam_strength(double ep, const OrcPressure* p) {
  const double r = ep / p->epsD;
  if (r >= 1) return 1e20;
  return p->plateau + p->gamma * r - p->densify * std::log(std::max(1e-30, 1 - std::pow(r, p->beta)));
}

// Return map of the pressure-dependent models on principal stretches:
// 8 cohesive Drucker-Prager with a tension cut-off and non-associated
// dilatancy, 9 Deshpande-Fleck crushable foam (associated, bisection on the
// multiplier). Hencky strains, Cauchy measures with J = exp(tr e) * jp.
// G, L are mu/E and lambda/E of the part (the reference's .4 and .4 give its
// G = .4/J, K = (2/3)/J). Outputs the returned stretches, the new jp (total J
// kept) and the equivalent plastic strain increment.
// src/three/pressure-plasticity.js:15-45
void orc_pressure_return(const double* sigma, int model, double jp, double ep, const OrcPressure* p, double G0,
                         double L0, double* out_sigma, double* out_jp, double* out_increment) {
  double e[3], dev[3];
  for (int k = 0; k < 3; ++k) e[k] = std::log(std::max(sigma[k], 1e-6));
  const double trace = e[0] + e[1] + e[2];
  for (int k = 0; k < 3; ++k) dev[k] = e[k] - trace / 3;
  const double length = std::sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double J = std::max(1e-6, std::exp(trace) * jp);
  const double G = G0 / J, K = (L0 + 2 * G0 / 3) / J;
  const double mean = K * trace, q = std::sqrt(6.0) * G * length;
  double meanNew = mean, qNew = q, increment = 0;
  if (model == 8) {
    const double t = q / std::sqrt(3.0), c = p->cohesion, a = p->slope, b = p->dilatancy, T = p->tension;
    if (mean > T && t + a * T <= c) {
      meanNew = T;
    } else if (t + a * mean > c) {
      const double dl = (t + a * mean - c) / (G + a * K * b);
      meanNew = mean - K * b * dl;
      qNew = std::sqrt(3.0) * std::max(0.0, t - G * dl);
      if (meanNew > T) {
        meanNew = T;
        qNew = std::sqrt(3.0) * std::max(0.0, c - a * T);
      }
    }
  } else {
    const double a2 = p->alpha * p->alpha, B = 1 + a2 / 9;
    const double trial = std::sqrt((q * q + a2 * mean * mean) / B);
    if (trial > orc_foam_strength(ep, p)) {
      auto state = [&](double eta, double* qn, double* mn, double* dl) {
        *qn = q / (1 + 3 * G * eta);
        *mn = mean / (1 + K * a2 * eta);
        const double hat = std::sqrt((*qn * *qn + a2 * *mn * *mn) / B);
        *dl = eta * B * hat;
        return hat - orc_foam_strength(ep + *dl, p);
      };
      double qn, mn, dl, lo = 0, hi = 1 / G;
      for (int k = 0; k < 32 && state(hi, &qn, &mn, &dl) > 0; ++k) hi *= 2;
      for (int k = 0; k < 28; ++k) {
        const double mid = (lo + hi) / 2;
        if (state(mid, &qn, &mn, &dl) > 0)
          lo = mid;
        else
          hi = mid;
      }
      state((lo + hi) / 2, &qn, &mn, &dl);
      qNew = qn;
      meanNew = mn;
      increment = dl;
    }
  }
  const double traceNew = meanNew / K, scale = q > 1e-20 ? qNew / q : 1;
  double ne[3];
  for (int k = 0; k < 3; ++k) ne[k] = traceNew / 3 + scale * dev[k];
  if (model == 8) {
    double d2 = 0;
    for (int k = 0; k < 3; ++k) d2 += (e[k] - ne[k]) * (e[k] - ne[k]);
    increment = std::sqrt(2.0 / 3) * std::sqrt(d2);
  }
  for (int k = 0; k < 3; ++k) out_sigma[k] = std::exp(ne[k]);
  *out_jp = jp * std::exp(trace - traceNew);
  *out_increment = increment;
}

// Kirchhoff stress of a law with unit shear modulus: mu = 1, lambda = lr
// (lr = lambda/mu; 1 in the reference). src/three/constitutive.js:92-106
void orc_constitutive_stress(const double* F, int model, double mu, double lambda, double jp, double* tau) {
  if (model == 0 || model == 2) {
    const double h = model == 2 ? hardening(jp) : 1;
    stress3(F, mu * h, lambda * h, tau);
    return;
  }
  for (int k = 0; k < 9; ++k) tau[k] = 0;
  if (model == 1) {
    const double pressure = lambda * std::log(std::max(det3(F), 1e-6));
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        for (int k = 0; k < 3; ++k) tau[a * 3 + b] += mu * F[a * 3 + k] * F[b * 3 + k];
        if (a == b) tau[a * 3 + b] += pressure - mu;
      }
  } else {
    Svd d = svd3(F);
    double e[3];
    for (int k = 0; k < 3; ++k) e[k] = std::log(std::max(d.sigma[k], 1e-6));
    const double trace = e[0] + e[1] + e[2];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 3; ++k)
          tau[a * 3 + b] += d.u[a * 3 + k] * (2 * mu * e[k] + lambda * trace) * d.u[b * 3 + k];
  }
}

// Hencky (or snow) stress on the returned principal directions, unit shear
// modulus, lambda/mu = lr. src/three/constitutive.js:81-86
void orc_principal_stress(const double* u, const double* sigma, int model, double lr, double* tau) {
  const double J = sigma[0] * sigma[1] * sigma[2];
  double e[3], diag[3];
  for (int k = 0; k < 3; ++k) e[k] = std::log(sigma[k]);
  const double trace = e[0] + e[1] + e[2];
  for (int k = 0; k < 3; ++k)
    diag[k] = model == 2 ? 2 * (sigma[k] - 1) * sigma[k] + (J - 1) * J : 2 * e[k] + lr * trace;
  for (int k = 0; k < 9; ++k) tau[k] = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int k = 0; k < 3; ++k) tau[a * 3 + b] += u[a * 3 + k] * diag[k] * u[b * 3 + k];
}

// Von Mises of a Kirchhoff stress divided by J. src/three/constitutive.js:88-90
double orc_von_mises(const double* t, double J) {
  return std::sqrt(std::max(0.0, .5 * ((t[0] - t[4]) * (t[0] - t[4]) + (t[4] - t[8]) * (t[4] - t[8]) +
                                      (t[8] - t[0]) * (t[8] - t[0])) +
                                     3 * (t[1] * t[1] + t[2] * t[2] + t[5] * t[5]))) /
         std::max(1e-6, J);
}

// Plastic projection for models 2-11. Returns -1 for an unknown model. Outputs F, jp, dilation, increment and the
// returned decomposition (u, v, projected sigma). src/three/constitutive.js:52-78
int orc_project_plastic(const double* F, int model, double jp, double dilation, double yieldStrain,
                        double friction, double ep, double dt, double T, const OrcPart* part,
                        double* outF, double* out_jp, double* out_dilation, double* out_increment,
                        double* out_u, double* out_v, double* out_sigma) {
  if (model < 2) {
    std::memcpy(outF, F, 9 * sizeof(double));
    *out_jp = jp;
    *out_dilation = dilation;
    *out_increment = 0;
    return 0;
  }
  if (model > 11) return -1;
  Svd d = svd3(F);
  const double* s = d.sigma;
  double projected[3];
  double metal = NAN;
  if (model == 2) {
    for (int k = 0; k < 3; ++k) projected[k] = clampd(s[k], 1 - yieldStrain, 1 + yieldStrain * .3);
    jp = clampd(jp * (s[0] * s[1] * s[2]) / (projected[0] * projected[1] * projected[2]), .6, 20);
  } else if (model >= 10) {
    orc_thermal_return(s, model, ep, dt, T, &part->th, part->G, projected, &metal);
  } else if (model >= 8) {
    orc_pressure_return(s, model, jp, ep, &part->pr, part->G, part->L, projected, &jp, &metal);
  } else if (model >= 4) {
    orc_metal_return(s, model, ep, dt, T, &part->pl, part->G, projected, &metal);
  } else {
    double e[3];
    for (int k = 0; k < 3; ++k) e[k] = std::log(std::max(s[k], 1e-6));
    const double trace = e[0] + e[1] + e[2] + dilation;
    if (trace >= 0) {
      projected[0] = projected[1] = projected[2] = 1;
      dilation = trace;
    } else {
      dilation = 0;
      double dev[3];
      for (int k = 0; k < 3; ++k) dev[k] = e[k] - trace / 3;
      const double length = std::sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
      const double sn = std::sin(friction * M_PI / 180);
      const double alpha = std::sqrt(2.0 / 3) * 2 * sn / (3 - sn);
      const double delta = std::max(0.0, length + 2.5 * trace * alpha);
      for (int k = 0; k < 3; ++k) projected[k] = std::exp(e[k] - delta / (length != 0 ? length : 1) * dev[k]);
    }
  }
  double inc;
  if (!std::isnan(metal)) {
    inc = metal;
  } else {
    double l[3];
    for (int k = 0; k < 3; ++k) l[k] = std::log(s[k] / projected[k]);
    inc = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
  }
  compose(d, projected, outF);
  *out_jp = jp;
  *out_dilation = dilation;
  *out_increment = inc;
  std::memcpy(out_u, d.u, sizeof d.u);
  std::memcpy(out_v, d.v, sizeof d.v);
  std::memcpy(out_sigma, projected, 3 * sizeof(double));
  return 0;
}

// ----------------------------------------------------------------- MPM3
static inline double particle_E(const OrcScene* s, int p) {
  const OrcPart& part = s->parts[s->part[p]];
  if (part.E > 0) return part.E;
  return s->mass[p] / (64 * s->volume[p]) * 16000 * s->stiffness;  // src/three/mpm.js:34
}

// Tensile-strain damage after the F update: the largest principal log
// stretch between onset and failure maps to damage in [0, 1], which never
// decreases; at 1 the particle becomes cohesionless sand with F = I, Jp = 1,
// no dilation and no cached stress. src/three/fracture.js:11-18
void orc_update_damage(const OrcScene* s, int p) {
  if (!s->damage) return;
  const OrcFracture& f = s->parts[s->part[p]].fr;
  if (!f.enabled || s->damage[p] >= 1) return;
  const Svd d = svd3(s->F + 9 * (size_t)p);
  double strain = 0;
  for (int k = 0; k < 3; ++k) strain = std::max(strain, std::log(std::max(d.sigma[k], 1e-8)));
  s->damage[p] = std::max(s->damage[p], std::min(1.0, (strain - f.onset) / (f.failure - f.onset)));
  if (s->damage[p] >= 1) {
    s->model[p] = 3;
    double* F = s->F + 9 * (size_t)p;
    for (int k = 0; k < 9; ++k) F[k] = (k % 4 == 0) ? 1 : 0;
    s->Jp[p] = 1;
    s->dilation[p] = 0;
    s->cacheValid[p] = 0;
  }
}

// Stress for this step from F (or the cached return stress for models >= 2)
// times G E h (1 - damage); von Mises diagnostic. src/three/mpm.js:33-39
void orc_update_stress(const OrcScene* s, int p) {
  const OrcPart& part = s->parts[s->part[p]];
  const double* F = s->F + p * 9;
  const double E = particle_E(s, p);
  const int model = s->model[p];
  const double h = model == 2 ? hardening(s->Jp[p]) : 1;
  const double lr = part.L / part.G;
  double local[9];
  const double* t;
  if (s->cacheValid[p] && model >= 2) {
    t = s->cachedTau + p * 9;
  } else {
    orc_constitutive_stress(F, model, 1, lr, s->Jp[p], local);
    t = local;
  }
  const double dmg = s->damage ? s->damage[p] : 0;
  const double soft = dmg < 1 ? 1 - dmg : 1;
  for (int k = 0; k < 9; ++k) s->tau[p * 9 + k] = t[k] * part.G * E * h * soft;
  s->stress[p] = orc_von_mises(t, det3(F) * ((model >= 8 && model < 10) ? s->Jp[p] : 1)) * part.G * h;
}

void orc_refresh_diagnostics(const OrcScene* s) {  // src/three/mpm.js:40-42
  for (int p = 0; p < s->count; ++p) orc_update_stress(s, p);
}

// CFL-type stable step. `stiffness` enters through 19200*stiffness = (lambda+
// 2 mu)/rho in reference units; for a part with fixed E the same wave speed is
// sqrt((lambda+2mu)/rho) with rho = m/V. `cap` replaces H/8/240 and `cfl`
// replaces 1.4/8 when positive. src/three/mpm.js:43-55
double orc_stable_step(const OrcScene* s, double cap, double cfl) {
  double mx = 1, grad = 1;
  for (int p = 0; p < s->count; ++p) {
    double row = 0, col = 0, g = 0;
    for (int a = 0; a < 3; ++a) {
      double r = 0, c = 0;
      for (int b = 0; b < 3; ++b) {
        r += std::fabs(s->F[p * 9 + a * 3 + b]);
        c += std::fabs(s->F[p * 9 + b * 3 + a]);
        g += std::fabs(s->C[p * 9 + a * 3 + b]);
      }
      row = std::max(row, r);
      col = std::max(col, c);
    }
    const int model = s->model[p];
    const double hard = model == 2 ? hardening(s->Jp[p]) : 1;
    const double compression =
        (model == 1 || (model >= 8 && model < 10))
            ? 1 / std::max(.05, det3(s->F + p * 9) * ((model >= 8 && model < 10) ? s->Jp[p] : 1))
            : 1;
    const OrcPart& part = s->parts[s->part[p]];
    double c2;
    if (part.E > 0) {
      const double rho = s->mass[p] / (s->vol_scale * s->volume[p]);
      c2 = (part.L + 2 * part.G) * part.E * hard / rho;
    } else {
      c2 = 19200 * s->stiffness * hard;
    }
    const double speed = std::sqrt(s->vx[p] * s->vx[p] + s->vy[p] * s->vy[p] + s->vz[p] * s->vz[p]);
    mx = std::max(mx, speed + std::sqrt(c2) * std::max(std::max(1.0, row * col), compression));
    grad = std::max(grad, g);
  }
  const double H = s->H;
  return std::min(std::min(cap > 0 ? cap : H / 8 / 240, (cfl > 0 ? cfl * H : 1.4 * H / 8) / mx), .08 / grad);
}

// Particle gravity as the reference's caller adds it before the MPM step.
// src/three/simulation.js:337-344 (constant gravity mode)
void orc_apply_particle_gravity(const OrcScene* s, const double* g, double dt) {
  for (int i = 0; i < s->count; ++i) {
    if (s->pinned[i]) continue;
    s->vx[i] += g[0] * dt;
    s->vy[i] += g[1] * dt;
    s->vz[i] += g[2] * dt;
  }
}

// One MLS-MPM/APIC step: stress, P2G of mass and affine momentum, grid
// velocity with wall rules, G2P with C and F update, plastic return.
// Every particle in the scene is a solid (no fluid, no fracture, no separate
// part grids, no platen). Returns -1 for an unsupported model.
// src/three/mpm.js:56-134
static bool parts_valid(const OrcScene* s) {
  for (int p = 0; p < s->count; ++p)
    if (s->part[p] < 0 || s->part[p] >= s->nparts) return false;
  return true;
}

// Quadratic B-spline stencil of particle p: base node, offsets and weights
// per axis. src/three/mpm.js:25-32
static void stencil3(const OrcScene* s, int p, int b[3], double f[3], double w[9]) {
  const double H = s->H;
  const int N = s->N;
  for (int a = 0; a < 3; ++a) {
    const double value = a == 0 ? s->x[p] : a == 1 ? s->y[p] : s->z[p];
    const double q = (value - s->origin[a]) / H;
    b[a] = (int)clampd(std::floor(q - .5), 0, N - 3);
    const double t = q - b[a];
    f[a] = t;
    w[a * 3] = .5 * (1.5 - t) * (1.5 - t);
    w[a * 3 + 1] = .75 - (t - 1) * (t - 1);
    w[a * 3 + 2] = .5 * (t - .5) * (t - .5);
  }
}

// First half of MPM3.step: this step's stress, P2G and the grid update
// (src/three/mpm.js:61-102). The grid then holds node masses and the
// velocities a caller may change before orc_step_from_grid (the
// externalGridContact hook, :104).
int orc_step_to_grid(const OrcScene* s, double dt) {
  if (!parts_valid(s)) return -1;
  const double H = s->H;
  const int N = s->N;
  const double* O = s->origin;
  const int n = s->count;
  for (int p = 0; p < n; ++p) orc_update_stress(s, p);  // :61
  const size_t nodes = (size_t)N * N * N;
  std::fill(s->gmass, s->gmass + nodes, 0.0);  // :64
  std::fill(s->gv, s->gv + 3 * nodes, 0.0);
  std::fill(s->gfixed, s->gfixed + nodes, (uint8_t)0);
  for (int p = 0; p < n; ++p) {  // P2G :66-85
    int b[3];
    double f[3], w[9];
    stencil3(s, p, b, f, w);
    const double m = s->mass[p], V = s->vol_scale * s->volume[p];
    double A[9];
    const double* C = s->C + p * 9;
    const double* tau = s->tau + p * 9;
    for (int k = 0; k < 9; ++k) A[k] = m * C[k] - dt * V * 4 / (H * H) * tau[k];
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
          const size_t node = ((size_t)(b[2] + k) * N + b[1] + j) * N + b[0] + i;
          const double weight = w[i] * w[3 + j] * w[6 + k];
          const double x = (i - f[0]) * H, y = (j - f[1]) * H, z = (k - f[2]) * H;
          s->gmass[node] += weight * m;
          s->gv[node * 3] += weight * (m * s->vx[p] + A[0] * x + A[1] * y + A[2] * z);
          s->gv[node * 3 + 1] += weight * (m * s->vy[p] + A[3] * x + A[4] * y + A[5] * z);
          s->gv[node * 3 + 2] += weight * (m * s->vz[p] + A[6] * x + A[7] * y + A[8] * z);
          if (s->pinned[p] && weight > 1e-8) s->gfixed[node] = 1;
        }
  }
  for (size_t node = 0; node < nodes; ++node) {  // grid :89-102
    const double mass = s->gmass[node];
    if (!(mass > s->mass_eps)) continue;
    const double q[3] = {O[0] + (double)(node % N) * H, O[1] + (double)((node / N) % N) * H,
                         O[2] + (double)(node / ((size_t)N * N)) * H};
    bool sticky = false;
    for (int a = 0; a < 3; ++a) {
      if ((s->wall_mode[2 * a] == 1 && q[a] < s->lower[a] + H) ||
          (s->wall_mode[2 * a + 1] == 1 && q[a] > s->upper[a] - H))
        sticky = true;
    }
    for (int a = 0; a < 3; ++a) {
      double v = s->gv[node * 3 + a] / mass + s->grid_gravity[a] * dt;
      const double lo = s->lower[a], hi = s->upper[a];
      const bool lo_hit = q[a] < lo + H && (s->wall_mode[2 * a] == 2 || v < 0);
      const bool hi_hit = q[a] > hi - H && (s->wall_mode[2 * a + 1] == 2 || v > 0);
      if (s->gfixed[node] || sticky || lo_hit || hi_hit) v = 0;
      s->gv[node * 3 + a] = v;
    }
    if (s->press_enabled && q[1] >= s->press_y - H * .5 && !s->gfixed[node]) {  // :97-101
      const double before = s->gv[node * 3 + 1], after = std::min(before, s->press_velocity);
      s->gv[node * 3 + 1] = after;
      const_cast<OrcScene*>(s)->press_impulse += (before - after) * mass;
    }
  }
  return 0;
}

// Second half of MPM3.step: G2P from the grid velocity plus the particle's
// part delta (the contactGridDeltas hook, src/three/mpm.js:114; deltas may be
// NULL, or hold one 3 N^3 array or NULL per part), F update, damage,
// advection and the plastic return (:107-131). Positions must be those of
// orc_step_to_grid.
int orc_step_from_grid(const OrcScene* s, double dt, const double* const* deltas) {
  const double H = s->H;
  const int N = s->N;
  const int n = s->count;
  for (int p = 0; p < n; ++p) {  // G2P :107-123
    if (s->pinned[p]) continue;
    int b[3];
    double f[3], w[9];
    stencil3(s, p, b, f, w);
    const double* delta = deltas ? deltas[s->part[p]] : nullptr;
    double vx = 0, vy = 0, vz = 0, C[9] = {0};
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
          const size_t node = ((size_t)(b[2] + k) * N + b[1] + j) * N + b[0] + i;
          const double weight = w[i] * w[3 + j] * w[6 + k];
          const double dx = (i - f[0]) * 4 / H, dy = (j - f[1]) * 4 / H, dz = (k - f[2]) * 4 / H;
          const double x = weight * (s->gv[node * 3] + (delta ? delta[node * 3] : 0)),
                       y = weight * (s->gv[node * 3 + 1] + (delta ? delta[node * 3 + 1] : 0)),
                       z = weight * (s->gv[node * 3 + 2] + (delta ? delta[node * 3 + 2] : 0));
          vx += x;
          vy += y;
          vz += z;
          C[0] += x * dx; C[1] += x * dy; C[2] += x * dz;
          C[3] += y * dx; C[4] += y * dy; C[5] += y * dz;
          C[6] += z * dx; C[7] += z * dy; C[8] += z * dz;
        }
    double F[9];
    std::memcpy(F, s->F + p * 9, sizeof F);
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double value = F[a * 3 + c];
        for (int k = 0; k < 3; ++k) value += dt * C[a * 3 + k] * F[k * 3 + c];
        s->F[p * 9 + a * 3 + c] = value;
      }
    orc_update_damage(s, p);
    std::memcpy(s->C + p * 9, C, sizeof C);
    s->vx[p] = vx;
    s->vy[p] = vy;
    s->vz[p] = vz;
    s->x[p] += dt * vx;
    s->y[p] += dt * vy;
    s->z[p] += dt * vz;
  }
  for (int p = 0; p < n; ++p) {  // plastic return :126-131
    if (s->pinned[p] || s->model[p] < 2) continue;
    const OrcPart& part = s->parts[s->part[p]];
    const double ys = s->yield_strain ? s->yield_strain[p] : .025;
    const double fr = s->friction ? s->friction[p] : 35;
    double F[9], jp, dil, inc, u[9], v[9], sg[3];
    if (orc_project_plastic(s->F + p * 9, s->model[p], s->Jp[p], s->dilation[p], ys, fr, s->plastic[p], dt,
                            s->temperature[p], &part, F, &jp, &dil, &inc, u, v, sg) != 0)
      return -1;
    std::memcpy(s->F + p * 9, F, sizeof F);
    s->Jp[p] = jp;
    s->dilation[p] = dil;
    s->plastic[p] += inc;
    orc_principal_stress(u, sg, s->model[p], part.L / part.G, s->cachedTau + p * 9);
    s->cacheValid[p] = 1;
  }
  return 0;
}

// One MPM3.step. src/three/mpm.js:56-134
int orc_step(const OrcScene* s, double dt) {
  const int rc = orc_step_to_grid(s, dt);
  return rc ? rc : orc_step_from_grid(s, dt, nullptr);
}

// P2G only (stress + scatter), leaving the grid holding mass and momentum;
// used to compare a single transfer. src/three/mpm.js:61-85
void orc_p2g(const OrcScene* s, double dt) {
  const double H = s->H;
  const int N = s->N;
  const double* O = s->origin;
  for (int p = 0; p < s->count; ++p) orc_update_stress(s, p);
  const size_t nodes = (size_t)N * N * N;
  std::fill(s->gmass, s->gmass + nodes, 0.0);
  std::fill(s->gv, s->gv + 3 * nodes, 0.0);
  std::fill(s->gfixed, s->gfixed + nodes, (uint8_t)0);
  for (int p = 0; p < s->count; ++p) {
    int b[3];
    double f[3], w[9];
    for (int a = 0; a < 3; ++a) {
      const double value = a == 0 ? s->x[p] : a == 1 ? s->y[p] : s->z[p];
      const double q = (value - O[a]) / H;
      b[a] = (int)clampd(std::floor(q - .5), 0, N - 3);
      const double t = q - b[a];
      f[a] = t;
      w[a * 3] = .5 * (1.5 - t) * (1.5 - t);
      w[a * 3 + 1] = .75 - (t - 1) * (t - 1);
      w[a * 3 + 2] = .5 * (t - .5) * (t - .5);
    }
    const double m = s->mass[p], V = s->vol_scale * s->volume[p];
    double A[9];
    for (int k = 0; k < 9; ++k) A[k] = m * s->C[p * 9 + k] - dt * V * 4 / (H * H) * s->tau[p * 9 + k];
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
          const size_t node = ((size_t)(b[2] + k) * N + b[1] + j) * N + b[0] + i;
          const double weight = w[i] * w[3 + j] * w[6 + k];
          const double x = (i - f[0]) * H, y = (j - f[1]) * H, z = (k - f[2]) * H;
          s->gmass[node] += weight * m;
          s->gv[node * 3] += weight * (m * s->vx[p] + A[0] * x + A[1] * y + A[2] * z);
          s->gv[node * 3 + 1] += weight * (m * s->vy[p] + A[3] * x + A[4] * y + A[5] * z);
          s->gv[node * 3 + 2] += weight * (m * s->vz[p] + A[6] * x + A[7] * y + A[8] * z);
          if (s->pinned[p] && weight > 1e-8) s->gfixed[node] = 1;
        }
  }
}

// Smallest det F over the particles (starting at 1). src/three/mpm.js:135
double orc_min_j(const OrcScene* s) {
  double v = 1;
  for (int i = 0; i < s->count; ++i) v = std::min(v, det3(s->F + i * 9));
  return v;
}

// Mechanical totals: momentum with the APIC spin term, kinetic and affine
// energy, D = H^2/4 in scene units. out = {mass, px, py, pz, Lx, Ly, Lz,
// kinetic, affine}. src/simulation/mechanical-diagnostics.js:16-37
void orc_mechanics(const OrcScene* s, double* out) {
  const double D = s->H * s->H / 4;
  for (int k = 0; k < 9; ++k) out[k] = 0;
  for (int i = 0; i < s->count; ++i) {
    const double m = s->mass[i];
    const double p[3] = {s->x[i], s->y[i], s->z[i]}, v[3] = {s->vx[i], s->vy[i], s->vz[i]};
    const double P[3] = {v[0] * m, v[1] * m, v[2] * m};
    const double* C = s->C + i * 9;
    double Hm[3] = {p[1] * P[2] - p[2] * P[1], p[2] * P[0] - p[0] * P[2], p[0] * P[1] - p[1] * P[0]};
    double cc = 0;
    for (int k = 0; k < 9; ++k) cc += C[k] * C[k];
    const double spin[3] = {C[7] - C[5], C[2] - C[6], C[3] - C[1]};
    for (int k = 0; k < 3; ++k) Hm[k] += m * D * spin[k];
    out[0] += m;
    for (int k = 0; k < 3; ++k) {
      out[1 + k] += P[k];
      out[4 + k] += Hm[k];
    }
    out[7] += .5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    out[8] += .5 * m * D * cc;
  }
}


// ------------------------------------------------------------ frame wrapper
// The all-solid branch of Simulation3.step / substep around MPM3.step
// (src/three/simulation.js:305-315, 330-388) with regularizeSolids
// (src/three/solid-sampling.js:5-42). Gravity modes as gravityAcceleration
// (src/three/gravity-options.js:7-16): 0 none, 1 uniform, 2 attract,
// 3 repel. No fluid, drag, press or thermal exchange.
typedef struct {
  double time_step_scale;
  int32_t max_substeps;
  double velocity_limit;
  double bounds_lo[3], bounds_hi[3];
  int32_t anti_clump;
  int32_t gravity_mode;
  double gravity_strength;
  double gravity_direction[3];
  double gravity_center[3];
  double gravity_softening;
} OrcFrame;

static void gravity_at(const OrcFrame* f, double x, double y, double z, double a[3]) {
  a[0] = a[1] = a[2] = 0;
  if (f->gravity_mode == 1) {
    for (int k = 0; k < 3; ++k) a[k] = f->gravity_direction[k] * f->gravity_strength * 24;
  } else if (f->gravity_mode == 2 || f->gravity_mode == 3) {
    const double d[3] = {f->gravity_center[0] - x, f->gravity_center[1] - y, f->gravity_center[2] - z};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + f->gravity_softening * f->gravity_softening;
    const double sc = (f->gravity_mode == 3 ? -1 : 1) * 24000 * f->gravity_strength / std::pow(r2, 1.5);
    for (int k = 0; k < 3; ++k) a[k] = d[k] * sc;
  }
}

// regularizeSolids: pairs of solid particles of one part closer than the mean
// of their spacings 2.8 cbrt(clamp(volume det F, 1/64, 8)) are pushed apart
// by (target - r) gain / (r (1/m_i + 1/m_j)) weighted by 1/m; the
// displacements are scaled down together so the largest is at most 8 dt,
// then applied to unpinned particles and clamped to the bounds. All pairs
// are visited (the reference's neighbour cells cover every pair closer than
// the largest spacing, 5.6).
void orc_regularize(const OrcScene* s, double dt, const double* lo, const double* hi) {
  const int n = s->count;
  std::vector<double> d(3 * (size_t)n, 0.0), sp(n), im(n);
  for (int i = 0; i < n; ++i) {
    const double* F = s->F + 9 * (size_t)i;
    const double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                     F[2] * (F[3] * F[7] - F[4] * F[6]);
    sp[i] = 2.8 * std::cbrt(clampd(s->volume[i] * J, .015625, 8.0));
    im[i] = s->pinned[i] ? 0.0 : 1.0 / s->mass[i];
  }
  const double gain = 1 - std::exp(-40 * dt);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (s->part[i] != s->part[j]) continue;
      const double total = im[i] + im[j];
      if (!total) continue;
      double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j], dz = s->z[i] - s->z[j];
      double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double target = (sp[i] + sp[j]) * .5;
      if (r >= target) continue;
      if (r < 1e-8) {
        const uint32_t hash = ((uint32_t)(i + 1) * 73856093u) ^ ((uint32_t)(j + 1) * 19349663u);
        dx = (hash & 1023) / 511.5 - 1;
        dy = ((hash >> 10) & 1023) / 511.5 - 1;
        dz = ((hash >> 20) & 1023) / 511.5 - 1;
        r = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double q = target * gain / (r * total);
        dx *= q;
        dy *= q;
        dz *= q;
      } else {
        const double q = (target - r) * gain / (r * total);
        dx *= q;
        dy *= q;
        dz *= q;
      }
      d[3 * (size_t)i] += dx * im[i];
      d[3 * (size_t)i + 1] += dy * im[i];
      d[3 * (size_t)i + 2] += dz * im[i];
      d[3 * (size_t)j] -= dx * im[j];
      d[3 * (size_t)j + 1] -= dy * im[j];
      d[3 * (size_t)j + 2] -= dz * im[j];
    }
  double scale = 1;
  const double budget = 8 * dt;
  for (int i = 0; i < n; ++i)
    scale = std::max(scale, std::sqrt(d[3 * (size_t)i] * d[3 * (size_t)i] + d[3 * (size_t)i + 1] * d[3 * (size_t)i + 1] +
                                      d[3 * (size_t)i + 2] * d[3 * (size_t)i + 2]) /
                                budget);
  for (int i = 0; i < n; ++i) {
    if (s->pinned[i]) continue;
    s->x[i] = clampd(s->x[i] + d[3 * (size_t)i] / scale, lo[0], hi[0]);
    s->y[i] = clampd(s->y[i] + d[3 * (size_t)i + 1] / scale, lo[1], hi[1]);
    s->z[i] = clampd(s->z[i] + d[3 * (size_t)i + 2] / scale, lo[2], hi[2]);
  }
}

// One frame of dt: substeps h = min(remaining, stableStep * scale) while more
// than 1e-9 remains and fewer than max_substeps were taken. Returns the
// substep count; *advanced receives the time advanced.
int orc_advance_frame(const OrcScene* s, const OrcFrame* f, double dt, double* advanced) {
  if (!parts_valid(s)) return -1;
  const int n = s->count;
  double remaining = dt;
  int steps = 0;
  std::vector<double> ox(n), oy(n), oz(n);
  while (remaining > 1e-9 && steps < f->max_substeps) {
    const double h = std::min(remaining, orc_stable_step(s, 0, 0) * f->time_step_scale);
    for (int i = 0; i < n; ++i) {
      ox[i] = s->x[i];
      oy[i] = s->y[i];
      oz[i] = s->z[i];
      if (s->pinned[i]) continue;
      double a[3];
      gravity_at(f, s->x[i], s->y[i], s->z[i], a);
      s->vx[i] += a[0] * h;
      s->vy[i] += a[1] * h;
      s->vz[i] += a[2] * h;
    }
    if (orc_step(s, h) != 0) return -1;
    const double L = f->velocity_limit;
    for (int i = 0; i < n; ++i) {
      if (s->pinned[i]) {
        s->x[i] = ox[i];
        s->y[i] = oy[i];
        s->z[i] = oz[i];
      } else {
        s->x[i] = clampd(s->x[i], f->bounds_lo[0], f->bounds_hi[0]);
        s->y[i] = clampd(s->y[i], f->bounds_lo[1], f->bounds_hi[1]);
        s->z[i] = clampd(s->z[i], f->bounds_lo[2], f->bounds_hi[2]);
      }
      s->vx[i] = clampd((s->x[i] - ox[i]) / h, -L, L);
      s->vy[i] = clampd((s->y[i] - oy[i]) / h, -L, L);
      s->vz[i] = clampd((s->z[i] - oz[i]) / h, -L, L);
    }
    if (f->anti_clump) orc_regularize(s, h, f->bounds_lo, f->bounds_hi);
    remaining -= h;
    ++steps;
  }
  if (advanced) *advanced = dt - remaining;
  return steps;
}

}  // extern "C"

// ------------------------------------------------------- learned surrogate
// FP64 reference of the constitutive surrogate network (no counterpart in the
// reference): 7 -> 64 -> 64 -> 64 -> 6, hard-swish z min(max(z/6 + 1/2, 0), 1)
// between layers. Parameters are
// packed W1[64x7], b1[64], W2[64x64], b2[64], W3[64x64], b3[64], W4[6x64],
// b4[6] (weights row-major, output-major), 9222 values.
namespace {
constexpr int kIn = 7, kHid = 64, kOut = 6;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kHid, kOffW4 = kOffB3 + kHid,
              kOffB4 = kOffW4 + kOut * kHid, kParams = kOffB4 + kOut;
inline double act(double z) { return z <= -3 ? 0 : (z >= 3 ? z : z * (z + 3) / 6); }
inline double dact(double z) { return z < -3 ? 0 : (z > 3 ? 1 : z / 3 + 0.5); }
void layer(const double* W, const double* b, int out, int in, const double* x, double* z) {
  for (int o = 0; o < out; ++o) {
    double acc = b[o];
    for (int i = 0; i < in; ++i) acc += W[o * in + i] * x[i];
    z[o] = acc;
  }
}
}  // namespace

extern "C" {
int orc_mlp_param_count(void) { return kParams; }

void orc_mlp_forward(const double* P, int n, const double* x, double* y) {
  double z1[kHid], h1[kHid], z2[kHid], h2[kHid], z3[kHid], h3[kHid];
  for (int s = 0; s < n; ++s) {
    layer(P + kOffW1, P + kOffB1, kHid, kIn, x + s * kIn, z1);
    for (int k = 0; k < kHid; ++k) h1[k] = act(z1[k]);
    layer(P + kOffW2, P + kOffB2, kHid, kHid, h1, z2);
    for (int k = 0; k < kHid; ++k) h2[k] = act(z2[k]);
    layer(P + kOffW3, P + kOffB3, kHid, kHid, h2, z3);
    for (int k = 0; k < kHid; ++k) h3[k] = act(z3[k]);
    layer(P + kOffW4, P + kOffB4, kOut, kHid, h3, y + s * kOut);
  }
}

// Mean squared error over n samples (sum over the 6 outputs, mean over
// samples) and its gradient with respect to the packed parameters.
double orc_mlp_loss_grad(const double* P, int n, const double* x, const double* t, double* grad) {
  for (int k = 0; k < kParams; ++k) grad[k] = 0;
  double loss = 0;
  double z1[kHid], h1[kHid], z2[kHid], h2[kHid], z3[kHid], h3[kHid], y[kOut];
  double dy[kOut], d3[kHid], d2[kHid], d1[kHid];
  for (int s = 0; s < n; ++s) {
    const double* xs = x + s * kIn;
    layer(P + kOffW1, P + kOffB1, kHid, kIn, xs, z1);
    for (int k = 0; k < kHid; ++k) h1[k] = act(z1[k]);
    layer(P + kOffW2, P + kOffB2, kHid, kHid, h1, z2);
    for (int k = 0; k < kHid; ++k) h2[k] = act(z2[k]);
    layer(P + kOffW3, P + kOffB3, kHid, kHid, h2, z3);
    for (int k = 0; k < kHid; ++k) h3[k] = act(z3[k]);
    layer(P + kOffW4, P + kOffB4, kOut, kHid, h3, y);
    for (int o = 0; o < kOut; ++o) {
      const double e = y[o] - t[s * kOut + o];
      loss += e * e;
      dy[o] = 2 * e / n;
    }
    for (int o = 0; o < kOut; ++o) {
      grad[kOffB4 + o] += dy[o];
      for (int i = 0; i < kHid; ++i) grad[kOffW4 + o * kHid + i] += dy[o] * h3[i];
    }
    for (int i = 0; i < kHid; ++i) {
      double acc = 0;
      for (int o = 0; o < kOut; ++o) acc += P[kOffW4 + o * kHid + i] * dy[o];
      d3[i] = acc * dact(z3[i]);
    }
    for (int o = 0; o < kHid; ++o) {
      grad[kOffB3 + o] += d3[o];
      for (int i = 0; i < kHid; ++i) grad[kOffW3 + o * kHid + i] += d3[o] * h2[i];
    }
    for (int i = 0; i < kHid; ++i) {
      double acc = 0;
      for (int o = 0; o < kHid; ++o) acc += P[kOffW3 + o * kHid + i] * d3[o];
      d2[i] = acc * dact(z2[i]);
    }
    for (int o = 0; o < kHid; ++o) {
      grad[kOffB2 + o] += d2[o];
      for (int i = 0; i < kHid; ++i) grad[kOffW2 + o * kHid + i] += d2[o] * h1[i];
    }
    for (int i = 0; i < kHid; ++i) {
      double acc = 0;
      for (int o = 0; o < kHid; ++o) acc += P[kOffW2 + o * kHid + i] * d2[o];
      d1[i] = acc * dact(z1[i]);
    }
    for (int o = 0; o < kHid; ++o) {
      grad[kOffB1 + o] += d1[o];
      for (int i = 0; i < kIn; ++i) grad[kOffW1 + o * kIn + i] += d1[o] * xs[i];
    }
  }
  return loss / n;
}

// ------------------------------------------------------------ frame snapshot
// Preview pick of the service's pack kernel: one sample per block of
// `stride` consecutive particles, chosen by an integer hash of the block
// index so the preview shows no lattice stripes. gpu/solver.py:990-996
uint32_t orc_preview_pick(uint32_t row, uint32_t stride, uint32_t n) {
  uint32_t h = row * 747796405u + 2891336453u;
  h = ((h >> ((h >> 28) + 4u)) ^ h) * 277803737u;
  h = (h >> 22) ^ h;
  const uint32_t span = std::min(stride, n - row * stride);
  return h % span;
}

typedef struct {
  uint32_t count, render_count, sample_stride;
  int64_t solids, liquids;
  int64_t models[12];
  double mass, min_temperature, max_temperature, min_J, max_stress, max_plastic;
  int64_t nonfinite;
} OrcSnapshot;

// Frame snapshot of an all-solid scene: stride = max(1, ceil(n / limit)),
// rows = ceil(n / stride); row r holds x, y, z, material code, T, pinned,
// stress, plastic of its picked particle (code 11 + model for the MPM solid
// material 11, else the code itself; `codes` per part, NULL: 11). Totals:
// mass, per-model counts, min/max T, min det F (from 1), max stress and
// plastic strain (from 0), particles with |x|^2 or |v|^2 not below 1e20 or
// non-finite. The stress is the scene's diagnostic (orc_refresh_diagnostics).
// gpu/solver.py:986-1032, src/three/simulation.js:390-400
void orc_snapshot(const OrcScene* s, uint32_t limit, const int32_t* codes, float* data, OrcSnapshot* out) {
  const uint32_t n = (uint32_t)s->count;
  OrcSnapshot r{};
  r.count = n;
  r.sample_stride = std::max<uint32_t>(1, (uint32_t)((n + (uint64_t)limit - 1) / limit));
  r.render_count = (uint32_t)((n + (uint64_t)r.sample_stride - 1) / r.sample_stride);
  r.min_temperature = 1e30;
  r.max_temperature = -1e30;
  r.min_J = 1;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t row = i / r.sample_stride;
    if (data && i % r.sample_stride == orc_preview_pick(row, r.sample_stride, n)) {
      float* d = data + 8 * (size_t)row;
      const int32_t c = codes ? codes[s->part[i]] : 11;
      d[0] = (float)s->x[i];
      d[1] = (float)s->y[i];
      d[2] = (float)s->z[i];
      d[3] = (float)(c == 11 ? 11 + s->model[i] : c);
      d[4] = (float)s->temperature[i];
      d[5] = (float)s->pinned[i];
      d[6] = (float)s->stress[i];
      d[7] = (float)s->plastic[i];
    }
    const double x2 = s->x[i] * s->x[i] + s->y[i] * s->y[i] + s->z[i] * s->z[i];
    const double v2 = s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i];
    if (!(x2 < 1e20 && v2 < 1e20)) ++r.nonfinite;
    r.mass += s->mass[i];
    ++r.solids;
    if (s->model[i] >= 0 && s->model[i] < 12) ++r.models[s->model[i]];
    r.min_J = std::min(r.min_J, det3(s->F + 9 * (size_t)i));
    r.min_temperature = std::min(r.min_temperature, s->temperature[i]);
    r.max_temperature = std::max(r.max_temperature, s->temperature[i]);
    r.max_stress = std::max(r.max_stress, s->stress[i]);
    r.max_plastic = std::max(r.max_plastic, s->plastic[i]);
  }
  *out = r;
}
}  // extern "C"
