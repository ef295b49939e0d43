// FP64 CPU restatement of the reference's 2D elastic MLS-MPM (SolidMPM,
// src/solid-mpm.js) and of the all-solid branch of its 2D driver
// (Simulation.step / substep, src/simulation.js:240-300), for testing the
// device engine. TEST INFRASTRUCTURE ONLY: the product never links it.
//
// Storage follows the reference: particle x, y, vx, vy, mass and volume are
// float32 (src/simulation.js:18-19), every write rounds to float32; F and C
// are float64 (src/solid-mpm.js:22).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
constexpr double DX = 8, ORIGIN = -592, AREA = 12.96;  // src/solid-mpm.js:4
constexpr int N = 149;

inline double clampd(double x, double a, double b) { return std::max(a, std::min(b, x)); }
}  // namespace

extern "C" {

// Kirchhoff stress P F^T of the 2D fixed-corotated law with the closed-form
// polar rotation (r, t) = (a + d, c - b) / |(a + d, c - b)| (identity when
// that norm is below 1e-12); out = (t00, t01, t10, t11). src/solid-mpm.js:10-18
void orc2_solid_stress(double a, double b, double c, double d, double mu, double lambda, double* out) {
  const double J = a * d - b * c, q = std::hypot(a + d, c - b);
  const double r = q > 1e-12 ? (a + d) / q : 1, t = q > 1e-12 ? (c - b) / q : 0;
  const double vol = lambda * (J - 1) * J;
  out[0] = 2 * mu * ((a - r) * a + (b + t) * b) + vol;
  out[1] = 2 * mu * ((a - r) * c + (b + t) * d);
  out[2] = 2 * mu * ((c - t) * a + (d - r) * b);
  out[3] = 2 * mu * ((c - t) * c + (d - r) * d) + vol;
}

// Caller-owned 2D scene of MPM solid particles. F, C hold 4 values per
// particle (00, 01, 10, 11). Walls: the reference's `boundary` setting and
// its scene-dependent box (|x| < bound, floor < y < top).
typedef struct {
  int count;
  float *x, *y, *vx, *vy;
  const float *mass, *volume;
  const uint8_t* pinned;
  double *F, *C;
  double stiffness;  // settings.solidStiffness
  int walls;
  double bound, floor, top;
  int active_nodes;  // nodes with mass of the last step
} OrcScene2;

// Quadratic B-spline stencil of one position. src/solid-mpm.js:31-37
static void stencil2(double x, double y, int& bx, int& by, double& fx, double& fy, double wx[3], double wy[3]) {
  const double px = (x - ORIGIN) / DX, py = (y - ORIGIN) / DX;
  bx = (int)clampd(std::floor(px - .5), 0, N - 3);
  by = (int)clampd(std::floor(py - .5), 0, N - 3);
  fx = px - bx;
  fy = py - by;
  wx[0] = .5 * (1.5 - fx) * (1.5 - fx);
  wx[1] = .75 - (fx - 1) * (fx - 1);
  wx[2] = .5 * (fx - .5) * (fx - .5);
  wy[0] = .5 * (1.5 - fy) * (1.5 - fy);
  wy[1] = .75 - (fy - 1) * (fy - 1);
  wy[2] = .5 * (fy - .5) * (fy - .5);
}

// CFL-type step: min(1/240, .18 DX / wave, .1 / max(1, |C|_1)) with
// wave = max(1, |v| + sqrt(19200 stiffness) max(1, |F|_F)^2).
// src/solid-mpm.js:38-47
double orc2_stable_step(const OrcScene2* s) {
  double wave = 1, gradient = 0;
  const double sound = std::sqrt(19200 * s->stiffness);
  for (int p = 0; p < s->count; ++p) {
    const double* F = s->F + 4 * (size_t)p;
    const double* C = s->C + 4 * (size_t)p;
    const double stretch = std::max(1.0, std::sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3]));
    wave = std::max(wave, std::hypot((double)s->vx[p], (double)s->vy[p]) + sound * stretch * stretch);
    gradient = std::max(gradient, std::fabs(C[0]) + std::fabs(C[1]) + std::fabs(C[2]) + std::fabs(C[3]));
  }
  return std::min(std::min(1.0 / 240, .18 * DX / wave), .1 / std::max(1.0, gradient));
}

// One MLS-MPM step: P2G with the fixed-corotated stress (mu = lambda = 0.4 E,
// E = m / (12.96 V) 16000 stiffness), grid velocities with pinned and wall
// rules, G2P with APIC C, F update and advection, positions clamped inside
// the walls. src/solid-mpm.js:48-92
void orc2_step(OrcScene2* s, double dt) {
  const size_t nn = (size_t)N * N;
  std::vector<double> mass(nn, 0.0), ux(nn, 0.0), uy(nn, 0.0), fixed(nn, 0.0);
  double wx[3], wy[3], fx, fy;
  int bx, by;
  for (int p = 0; p < s->count; ++p) {
    stencil2(s->x[p], s->y[p], bx, by, fx, fy, wx, wy);
    const double m = s->mass[p], V0 = AREA * s->volume[p], rho = m / V0, E = rho * 16000 * s->stiffness;
    const double* F = s->F + 4 * (size_t)p;
    const double* C = s->C + 4 * (size_t)p;
    double tau[4];
    orc2_solid_stress(F[0], F[1], F[2], F[3], .4 * E, .4 * E, tau);
    const double factor = -dt * V0 * 4 / (DX * DX);
    const double a00 = m * C[0] + factor * tau[0], a01 = m * C[1] + factor * tau[1];
    const double a10 = m * C[2] + factor * tau[2], a11 = m * C[3] + factor * tau[3];
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        const size_t n = (size_t)(by + j) * N + bx + i;
        const double w = wx[i] * wy[j], dx = (i - fx) * DX, dy = (j - fy) * DX;
        mass[n] += w * m;
        ux[n] += w * (m * s->vx[p] + a00 * dx + a01 * dy);
        uy[n] += w * (m * s->vy[p] + a10 * dx + a11 * dy);
        if (s->pinned[p]) fixed[n] += w * m;
      }
  }
  s->active_nodes = 0;
  for (size_t n = 0; n < nn; ++n) {
    if (!(mass[n] > 1e-12)) continue;
    ++s->active_nodes;
    ux[n] /= mass[n];
    uy[n] /= mass[n];
    if (fixed[n] > 1e-9) {
      ux[n] = uy[n] = 0;
      continue;
    }
    if (s->walls) {
      const double x = ORIGIN + (double)(n % N) * DX, y = ORIGIN + (double)(n / N) * DX;
      if ((x < -s->bound + DX && ux[n] < 0) || (x > s->bound - DX && ux[n] > 0)) ux[n] = 0;
      if ((y < s->floor + DX && uy[n] < 0) || (y > s->top - DX && uy[n] > 0)) uy[n] = 0;
    }
  }
  for (int p = 0; p < s->count; ++p) {
    if (s->pinned[p]) continue;
    stencil2(s->x[p], s->y[p], bx, by, fx, fy, wx, wy);
    double vx = 0, vy = 0, a = 0, b = 0, c = 0, d = 0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        const size_t n = (size_t)(by + j) * N + bx + i;
        const double w = wx[i] * wy[j], dx = (i - fx) * 4 / DX, dy = (j - fy) * 4 / DX;
        vx += w * ux[n];
        vy += w * uy[n];
        a += w * ux[n] * dx;
        b += w * ux[n] * dy;
        c += w * uy[n] * dx;
        d += w * uy[n] * dy;
      }
    double* C = s->C + 4 * (size_t)p;
    double* F = s->F + 4 * (size_t)p;
    C[0] = a;
    C[1] = b;
    C[2] = c;
    C[3] = d;
    const double f = F[0], g = F[1], h = F[2], k = F[3];
    F[0] = (1 + dt * a) * f + dt * b * h;
    F[1] = (1 + dt * a) * g + dt * b * k;
    F[2] = dt * c * f + (1 + dt * d) * h;
    F[3] = dt * c * g + (1 + dt * d) * k;
    s->vx[p] = (float)vx;
    s->vy[p] = (float)vy;
    s->x[p] = (float)(s->x[p] + dt * vx);  // the unrounded velocity (src/solid-mpm.js:89)
    s->y[p] = (float)(s->y[p] + dt * vy);
    if (s->walls) {
      s->x[p] = (float)clampd(s->x[p], -s->bound + .1, s->bound - .1);
      s->y[p] = (float)clampd(s->y[p], s->floor + .1, s->top - .1);
    }
  }
}

// Deformation statistics: count, min/max det F, largest Green strain norm
// .5 |(a^2 + c^2 - 1, b^2 + d^2 - 1, sqrt2 (a b + c d))|.
// out = {count, minJ, maxJ, strain}. src/solid-mpm.js:106-115
void orc2_stats(const OrcScene2* s, double* out) {
  double mn = INFINITY, mx = 0, strain = 0;
  for (int p = 0; p < s->count; ++p) {
    const double* F = s->F + 4 * (size_t)p;
    const double a = F[0], b = F[1], c = F[2], d = F[3], J = a * d - b * c;
    mn = std::min(mn, J);
    mx = std::max(mx, J);
    const double e0 = a * a + c * c - 1, e1 = b * b + d * d - 1, e2 = std::sqrt(2.0) * (a * b + c * d);
    strain = std::max(strain, .5 * std::sqrt(e0 * e0 + e1 * e1 + e2 * e2));
  }
  out[0] = s->count;
  out[1] = s->count ? mn : 1;
  out[2] = s->count ? mx : 1;
  out[3] = strain;
}

// One frame of the 2D driver for a scene of MPM solids only: dt clamped to
// [0, .05]; substeps h = min(remaining, stable step) while time remains and
// fewer than 96 have run; each substep adds gravity to unpinned velocities
// (mode 0 none, 1 uniform (0, -1), 2 towards the centre -(x, y) / sqrt(x^2 +
// y^2 + 400)) times 27 strength, zeroes pinned velocities, clamps v to
// +-180 (no damping for MPM solids) and runs the MPM step. Returns the
// substep count; *advanced receives the time solved.
// src/simulation.js:240-300
int orc2_advance(OrcScene2* s, double dt, int gravity_mode, double strength, double* advanced) {
  dt = clampd(dt, 0, .05);
  double remaining = dt;
  int steps = 0;
  while (remaining > 1e-9 && steps < 96) {
    const double h = std::min(remaining, orc2_stable_step(s));
    for (int i = 0; i < s->count; ++i) {
      if (s->pinned[i]) {
        s->vx[i] = s->vy[i] = 0;
        continue;
      }
      double gx = 0, gy = 0;
      if (gravity_mode == 2) {
        const double r = std::sqrt((double)s->x[i] * s->x[i] + (double)s->y[i] * s->y[i] + 400);
        gx = -s->x[i] / r;
        gy = -s->y[i] / r;
      } else if (gravity_mode == 1) {
        gy = -1;
      }
      const double force = 27 * strength;
      s->vx[i] = (float)(s->vx[i] + gx * force * h);
      s->vy[i] = (float)(s->vy[i] + gy * force * h);
      s->vx[i] = (float)clampd(s->vx[i], -180, 180);
      s->vy[i] = (float)clampd(s->vy[i], -180, 180);
    }
    orc2_step(s, h);
    remaining -= h;
    ++steps;
  }
  *advanced = dt - remaining;
  return steps;
}

}  // extern "C"
