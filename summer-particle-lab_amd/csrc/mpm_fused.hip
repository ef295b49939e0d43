// Fused particle kernel of the MLS-MPM step for gfx950.
//
// One launch per step does G2P of step s and P2G of step s+1 for a chunk of
// 256 cell-sorted particles per workgroup (one particle per lane):
//   1. gather v and C from the grid velocity (src/three/mpm.js:107-117); the
//      27-node neighbourhoods of the chunk are first staged in LDS as one box
//      of nodes, so the lanes read LDS instead of issuing 27 global loads;
//   2. F <- (I + dt C) F, x += dt v (src/three/mpm.js:118-122), the plastic
//      return for models >= 2 (src/three/mpm.js:126-131) and the stress of the
//      next step (src/three/mpm.js:33-39);
//   3. scatter of mass and APIC momentum into the next grid
//      (src/three/mpm.js:66-85). The lanes of a wave are grouped by base cell,
//      their node contributions transposed through LDS in that order and
//      summed over each run of particles sharing a base cell; each run sum is
//      added to an LDS box of nodes, and the box is added to the global grid
//      with one atomic per node and channel.
// Node mass is 64-bit fixed point: a particle's 27 parts sum exactly to its
// own fixed-point mass (the centre node takes the remainder), so the grid
// total equals the particle total bit for bit in any summation order.
//
// MODE 0: P2G from the stored state (first step of a call, after a re-bin).
// MODE 1: G2P of this step and P2G of the next (steady state); only x, F
//         and ep are written back, v and C stay in registers.
// MODE 2: G2P only, writing the full state (last step of a call).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "mpm_internal.h"
#include "mpm_mlp_device.h"
#include "mpm_wave.h"

namespace mpm {
namespace {

constexpr int FB = 256;            // particles per chunk of the single-scene launch
constexpr int TR_STRIDE = 65;      // 64 lanes + 1 entry: rows start in different banks
constexpr int CHUNK_MAX = 16;      // lanes per scatter chunk (run pieces summed by one lane), at most
constexpr int TR_TAIL = 16;        // zero entries after the last row
constexpr int SORT_PASSES = 8;     // cells of a wave grouped by minimum passes before P2G
static_assert(CHUNK_MAX <= TR_TAIL, "the zero tail covers a chunk read past the last row");

// Shape of the scatter: chunks of CH lanes and R nodes transposed per pass.
// Dense cells (long runs) take 16-lane chunks; cells of about 8 particles
// take 8-lane chunks, whose task reads carry half the dead entries (measured
// at 8 particles per cell: 1214 against 1314 us of the fused kernel; at 64,
// 16 lanes 132.7 against 136.9 us). A pass's tasks are R rows x chunks: at
// 8 particles per cell a wave holds about 9 chunks, so 6 rows (five passes,
// 54 tasks) keep each pass within one round of the wave's 64 lanes; at 64
// per cell it holds about 4, and 7 rows (four passes) do. The per-wave
// buffer holds R transpose rows, the zero tail and the chunk table
// (int4 {first lane, last lane, box base, grid base} or, packed, int2
// {first lane | (lanes - 1) << 6 | box base << 12, grid base}); the node box
// takes 2.5 (sparse) or 2 (dense) nodes per particle, so that four
// workgroups fit a CU's LDS. The dense shape packs its chunk table into 8
// bytes per chunk to fit; the sparse one keeps 16 bytes, whose fields need
// no unpacking. Measured at four waves per SIMD, each pair on one GPU:
// large scene (8-lane chunks) 1021 us with 6 rows against 1032 with 5;
// thermal boxes (16-lane chunks) 115.7 us with 7 rows against 119.9 with 6
// and 125.2 with 5.
template <int CH>
struct Scatter {
  static constexpr int CHUNK = CH;
  static constexpr int R_PASS = CH == 16 ? 7 : 6;
  static constexpr int NPASS = (27 + R_PASS - 1) / R_PASS;
  static constexpr int ROWS = R_PASS;  // transpose rows in the wave buffer
  static constexpr int ENTRIES = ROWS * TR_STRIDE + TR_TAIL;
  static constexpr bool PACKED = CH == 16;  // chunk table entries of 8 bytes
  static constexpr int WAVE_BYTES = ENTRIES * 16 + 64 * (PACKED ? 8 : 16);  // rows, zero tail, chunk table
  static constexpr int ACC_PER_PARTICLE_X2 = CH == 16 ? 4 : 5;  // node box nodes per particle, times two
  static_assert(CH % 8 == 0 && CH <= CHUNK_MAX, "scatter pass shape");
  static_assert(WAVE_BYTES % 16 == 0, "wave buffers must stay 16-byte aligned");
  static_assert((ROWS - 1) * TR_STRIDE + 63 + CHUNK_MAX <= ENTRIES,
                "the run-sum reads past a row's last lane stay inside the rows and the zero tail");
  static_assert(ROWS + TR_TAIL + CHUNK_MAX <= 64,
                "one store per lane zeroes the pads, the tail and the row past the passes");
};

// Chunk table entries, unpacked (int4) or packed (int2).
__device__ __forceinline__ int4 make_chunk(int4*, int first, int last, int box, int grid) {
  return make_int4(first, last, box, grid);
}
__device__ __forceinline__ int2 make_chunk(int2*, int first, int last, int box, int grid) {
  return make_int2(first | ((last - first) << 6) | (box << 12), grid);
}
__device__ __forceinline__ int chunk_first(int4 c) { return c.x; }
__device__ __forceinline__ int chunk_first(int2 c) { return c.x & 63; }
__device__ __forceinline__ int chunk_span(int4 c) { return c.y - c.x; }  // lanes - 1
__device__ __forceinline__ int chunk_span(int2 c) { return (c.x >> 6) & 63; }
__device__ __forceinline__ int chunk_box(int4 c) { return c.z; }
__device__ __forceinline__ int chunk_box(int2 c) { return c.x >> 12; }
__device__ __forceinline__ int chunk_grid(int4 c) { return c.w; }
__device__ __forceinline__ int chunk_grid(int2 c) { return c.y; }

// Node g of the scatter order: the 26 nodes other than the centre (n = 13,
// offset (1, 1, 1)) in index order, then the centre, which takes the
// remainder of the fixed-point mass. Node index n = ii + 3 j + 9 k.
__host__ __device__ constexpr int scatter_node(int g) { return g == 26 ? 13 : g + (g >= 13 ? 1 : 0); }

// LDS plan of a workgroup of FBK particles (one per lane) for scatter chunks
// of CH lanes: a node box for the scatter (32-bit fixed-point mass and
// momentum x, y, z planes, 16 B per node) and one buffer per wave (in the
// surrogate variant at least the network's activations, which reuse it); the
// gather box (16 B per node) reuses the whole buffer.
template <int FBK, bool SUR, int CH>
struct Shape {
  static constexpr int WAVES = FBK / 64;
  static constexpr int PATCH_ACC = Scatter<CH>::ACC_PER_PARTICLE_X2 * FBK / 2;
  static constexpr int ACC_BYTES = 16 * PATCH_ACC;
  static constexpr int WAVE_BYTES = SUR && MLP_WAVE_LDS > Scatter<CH>::WAVE_BYTES ? (MLP_WAVE_LDS + 15) / 16 * 16
                                                                                   : Scatter<CH>::WAVE_BYTES;
  static constexpr int LDS_BYTES = ACC_BYTES + WAVES * WAVE_BYTES;
  static constexpr int PATCH_VEL = LDS_BYTES / 16;
  static_assert(ACC_BYTES % 16 == 0, "wave buffers must stay 16-byte aligned");
};

__device__ __forceinline__ float ld(const float* P, uint32_t cap, int k, uint32_t i) {
  return P[(size_t)k * cap + i];
}
__device__ __forceinline__ void st(float* P, uint32_t cap, int k, uint32_t i, float v) {
  P[(size_t)k * cap + i] = v;
}

__device__ __forceinline__ int base_axis(float x, float O, float invH, int N, float& t) {
  const float q = __fmul_rn(__fsub_rn(x, O), invH);
  int b = (int)floorf(__fsub_rn(q, 0.5f));
  b = b < 0 ? 0 : (b > N - 3 ? N - 3 : b);
  t = __fsub_rn(q, (float)b);
  return b;
}

// Minimum (MAX = false) or maximum over the 64 lanes of a wave, all lanes
// active; the result is wave-uniform.
template <bool MAX>
__device__ __forceinline__ int wave_extreme(int x) {
  return wave_reduce(x, MAX ? (int)0x80000000 : 0x7fffffff, [](int a, int b) { return MAX ? max(a, b) : min(a, b); });
}

// Block-wide box of base cells: lo[a] = min b[a], hi[a] = max b[a] over the
// lanes with `use` set. Returns false when no lane is in use. With `top`
// set, *top becomes the block's maximum of every lane's *top in the same
// pass (s_box holds 7 words per wave). Called with every lane of the block
// active, at most once per s_box: no barrier guards s_box against a
// previous call's readers.
template <int WAVES>
__device__ __forceinline__ bool block_box(const int b[3], bool use, int lo[3], int hi[3], int* s_box,
                                          int* top = nullptr) {
  int mn[3], mx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = wave_extreme<false>(use ? b[a] : 0x7fffffff);
    mx[a] = wave_extreme<true>(use ? b[a] : -1);
  }
  const int wtop = top ? wave_extreme<true>(*top) : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s_box[wave * 6 + a] = mn[a];
      s_box[wave * 6 + 3 + a] = mx[a];
    }
    if (top) s_box[WAVES * 6 + wave] = wtop;
  }
  __syncthreads();
  if (top) {
    int t = s_box[WAVES * 6];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) t = max(t, s_box[WAVES * 6 + w]);
    *top = t;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = s_box[a];
    hi[a] = s_box[3 + a];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) {
      lo[a] = min(lo[a], s_box[w * 6 + a]);
      hi[a] = max(hi[a], s_box[w * 6 + 3 + a]);
    }
  }
  return hi[0] >= 0;
}

// A box of base cells that is non-empty and inside [0, N - 3] on every axis,
// so its nodes (up to hi + 2) are grid nodes. An empty box (no lane in use)
// fails.
__device__ __forceinline__ bool box_inside(const int lo[3], const int hi[3], int N) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) ok = ok && lo[a] >= 0 && lo[a] <= hi[a] && hi[a] <= N - 3;
  return ok;
}

// Exponent e of a finite x with |x| < 2^e (0 for x = 0).
__device__ __forceinline__ int frexp_exponent(float x) { return __builtin_amdgcn_frexp_expf(x); }

// Box coordinates (lx, ly, lz) of entry p of an ex x ey x ez box, without
// integer division. For 0 <= p < 2^21 and any divisor d >= 1,
// (p + 1/2) / d lies at least 0.5 / d inside [q, q + 1). The float product
// (p + 1/2) * (1/d), with the reciprocal within 1 ulp (the hardware
// reciprocal) and one rounding, is
// off by a relative 2^-22 at most, that is by less than
// (p + 1/2) / d * 2^-22 < 0.5 / d, so its truncation is the exact quotient q.
// Boxes hold at most a few thousand entries.
struct BoxIndex {
  int ex, exy;
  float inv_ex, inv_exy;
  __device__ __forceinline__ BoxIndex(int ex_, int ey_) : ex(ex_), exy(ex_ * ey_) {
    inv_ex = __builtin_amdgcn_rcpf((float)ex);
    inv_exy = __builtin_amdgcn_rcpf((float)exy);
  }
  __device__ __forceinline__ void operator()(int p, int& lx, int& ly, int& lz) const {
    lz = (int)(((float)p + 0.5f) * inv_exy);
    const int rest = p - lz * exy;
    ly = (int)(((float)rest + 0.5f) * inv_ex);
    lx = rest - ly * ex;
  }
};

// Material record, read through the scalar path when the wave shares it.
__device__ __forceinline__ DevMaterial load_material(const DevMaterial* __restrict__ mats, uint32_t id) {
  const uint32_t first = __builtin_amdgcn_readfirstlane(id);
  if (__all(id == first)) return mats[first];
  return mats[id];
}

// The stable-step test of the reference's stableStep (src/three/mpm.js:43-55)
// as the GPU service applies it after a step (gpu/mixed_solver.py:383): true
// when this particle's wave speed |v| + c max(1, |F|row |F|col, compression),
// c = sqrt((lambda + 2 mu) hardening / rho), or its sum |C| is above what the
// step size allows (the limits of StepParams). A value that is not a number
// fails too.
__device__ __forceinline__ bool cfl_exceeded(const M3& F, const float C[9], const float v[3], float m, float vol,
                                             const DevMaterial& mat, int model, float jp, const StepParams& sp) {
  float row = 0.f, col = 0.f, gsum = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    row = fmaxf(row, fabsf(F[a * 3]) + fabsf(F[a * 3 + 1]) + fabsf(F[a * 3 + 2]));
    col = fmaxf(col, fabsf(F[a]) + fabsf(F[3 + a]) + fabsf(F[6 + a]));
    gsum += fabsf(C[a * 3]) + fabsf(C[a * 3 + 1]) + fabsf(C[a * 3 + 2]);
  }
  const float hard = model == 2 ? snow_hardening(jp) : 1.f;
  float comp = 1.f;
  if (model == 1 || model == 8 || model == 9)
    comp = __builtin_amdgcn_rcpf(fmaxf(.05f, det3(F) * (model == 1 ? 1.f : jp)));
  // Hardware reciprocal and square root (about 1 ulp): the limits carry the
  // reference's 0.99999 margin, far wider than their error.
  const float c = __builtin_amdgcn_sqrtf((mat.lambda + 2.f * mat.mu) * hard * vol * __builtin_amdgcn_rcpf(m));
  const float wave =
      __builtin_amdgcn_sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + c * fmaxf(fmaxf(1.f, row * col), comp);
  return !(wave <= sp.wave_limit) || !(gsum <= sp.grad_limit);
}

// The invalid state of the GPU service's check (gpu/mixed_solver.py:384-388):
// det F outside [1e-6, 1e6] or not a number, plastic strain not a number, a
// velocity component not finite or above 1e5 in magnitude, a position not a
// number or more than one cell outside the domain box. Every comparison is
// false for a NaN, so a NaN fails.
__device__ __forceinline__ bool state_invalid(const M3& F, float ep, const float v[3], const float x[3],
                                              const GridDims& g) {
  const float J = det3(F);
  bool ok = J >= 1e-6f && J <= 1e6f && ep == ep;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    ok = ok && fabsf(v[a]) <= 1e5f && x[a] >= g.lower[a] - g.H && x[a] <= g.upper[a] + g.H;
  return !ok;
}

// The particle work of one chunk of FBK cell-sorted particles in slots
// [first, min(first + FBK, end)). Grid pointers are those of the chunk's
// scene; momentum planes are `pstride` floats apart.
template <int MODE, bool SUR, int FBK, bool EXT, int CH, bool RG = false>
__device__ __forceinline__ void fused_chunk(float* __restrict__ P, uint32_t* __restrict__ meta, uint32_t cap,
                                            uint32_t first, uint32_t end, const DevMaterial* __restrict__ mats,
                                            const GridDims& g, const StepParams& sp, size_t pstride,
                                            const float4* __restrict__ vel, unsigned long long* __restrict__ acc_m,
                                            float* __restrict__ acc_p, uint8_t* __restrict__ fixed,
                                            uint8_t* __restrict__ pieces, const MlpDev* __restrict__ mlp,
                                            unsigned long long* __restrict__ sur_counts, unsigned char* s_raw,
                                            int* s_box, uint32_t* __restrict__ status,
                                            const float4* __restrict__ deltas = nullptr,
                                            const int32_t* __restrict__ delta_slot = nullptr,
                                            const Regather& rg = Regather{}) {
  static_assert(!RG || MODE == 1, "a re-gather is a fused step");
  using S = Shape<FBK, SUR, CH>;
  constexpr int TR_ROWS = Scatter<CH>::ROWS;
  constexpr int TR_ENTRIES = Scatter<CH>::ENTRIES;
  constexpr int CHUNK = Scatter<CH>::CHUNK;
  constexpr int R_PASS = Scatter<CH>::R_PASS;
  constexpr int NPASS = Scatter<CH>::NPASS;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t i = first + tid;
  const bool valid = i < end;
  // The slot this lane reads (after a re-bin sort, through the new order)
  // and the buffers it writes.
  const uint32_t ic = RG ? rg.order[valid ? i : first] : (valid ? i : first);
  float* const Po = RG ? rg.planes_out : P;
  uint32_t* const meta_o = RG ? rg.meta_out : meta;
  const int N = g.N;
  const float H = g.H, invH = g.invH, dt = sp.dt;

  // The loads that need no material go first, so they are in flight
  // together with the meta word.
  float x[3] = {ld(P, cap, P_X, ic), ld(P, cap, P_Y, ic), ld(P, cap, P_Z, ic)};
  M3 F;
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = ld(P, cap, P_F0 + k, ic);
  const float m = MODE != 2 ? ld(P, cap, P_MASS, ic) : 0.f;
  const float vol = MODE != 2 ? ld(P, cap, P_VOL, ic) : 0.f;
  const uint32_t mt = meta[ic];
  const uint32_t flags = mt >> 8;
  const bool pinned = (flags & FLAG_PINNED) != 0;
  const DevMaterial mat = load_material(mats, mt & 255u);
  float v[3] = {0.f, 0.f, 0.f}, Cm[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float tau[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float ep = 0.f, jp = 1.f, dil = 0.f;
  uint32_t new_flags = flags;
  // Damage and the model it implies (a failed particle steps as sand). EXT
  // variants carry damage and the pressure models; the others are launched
  // only when no material uses them.
  float dmg = 0.f;
  int model = mat.model;
  if constexpr (EXT) {
    dmg = mat.fr_on ? ld(P, cap, P_DMG, ic) : 0.f;
    model = effective_model(mat, dmg);
  }
  // Plastic strain and temperature of a return in G2P, loaded with the
  // first loads when some material of the scene returns (models 2 and up, a
  // particle that may fail into sand): the material record arrives while the
  // gather runs.
  const bool hist = MODE != 0 && sp.returns;
  const float ep0 = hist ? ld(P, cap, P_EP, ic) : 0.f;
  const float temp0 = hist ? ld(P, cap, P_TEMP, ic) : 0.f;

  if (MODE == 0 || pinned) {
    // State as stored: stress of this step from F (or the cached return).
    v[0] = ld(P, cap, P_VX, ic);
    v[1] = ld(P, cap, P_VY, ic);
    v[2] = ld(P, cap, P_VZ, ic);
#pragma unroll
    for (int k = 0; k < 9; ++k) Cm[k] = ld(P, cap, P_C0 + k, ic);
    if (model >= 2 && (flags & FLAG_CACHE)) {
#pragma unroll
      for (int k = 0; k < 6; ++k) tau[k] = ld(P, cap, P_T0 + k, ic);
    } else {
      elastic_tau(F, model, mat, model == 2 ? ld(P, cap, P_JP, ic) : 1.f, tau);
    }
  }

  if (MODE != 0) {
    // ---------------------------------------------------------------- G2P
    float t[3], w[3][3];
    int b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      b[a] = base_axis(x[a], g.O[a], invH, N, t[a]);
      bspline(t[a], w[a]);
    }
    int lo[3], hi[3];
    const bool gather = valid && !pinned;
    block_box<S::WAVES>(b, gather, lo, hi, s_box);
    const int ex = hi[0] - lo[0] + 3, ey = hi[1] - lo[1] + 3, ez = hi[2] - lo[2] + 3;
    const bool staged = box_inside(lo, hi, N) && ex * ey * ez <= S::PATCH_VEL;
    float4* s_vel = reinterpret_cast<float4*>(s_raw);
    if (staged) {
      // Up to four nodes per lane and round, their loads issued together.
      const int cnt = ex * ey * ez;
      const BoxIndex at(ex, ey);
      for (int p0 = tid; p0 < cnt; p0 += 4 * FBK) {
        float4 nv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int p = p0 + u * FBK;
          if (p < cnt) {
            int lx, ly, lz;
            at(p, lx, ly, lz);
            nv[u] = vel[((size_t)(lo[2] + lz) * N + (lo[1] + ly)) * N + (lo[0] + lx)];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (p0 + u * FBK < cnt) s_vel[p0 + u * FBK] = nv[u];
      }
    }
    __syncthreads();
    M3 Fn = F;
    // Per-part grid velocity deltas of a split step (src/three/mpm.js:114).
    const float4* dl = nullptr;
    if constexpr (MODE == 2) {
      if (deltas) {
        const int slot = delta_slot[mt & 255u];
        if (slot >= 0) dl = deltas + (size_t)slot * N * N * N;
      }
    }
    if (gather) {
      // The weighted sums of the node velocities and of their first moments,
      // V = sum w v, X, Y, Z = sum w v (offset along x, y, z), contracted
      // axis by axis: each row of three nodes along x, then the rows along
      // y for each k, then along z (the first term of each sum is a product,
      // not an add to zero).
      typedef float f4 __attribute__((ext_vector_type(4)));
      const float wx2 = 2.f * w[0][2], wy2 = 2.f * w[1][2], wz2 = 2.f * w[2][2];
      auto contract = [&](auto node, float Vo[3], float Xo[3], float Yo[3], float Zo[3]) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float a0[3], a1[3], ay[3];
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const f4 n0 = node(k, j, 0), n1 = node(k, j, 1), n2 = node(k, j, 2);
            const float r0[3] = {w[0][0] * n0.x + w[0][1] * n1.x + w[0][2] * n2.x,
                                 w[0][0] * n0.y + w[0][1] * n1.y + w[0][2] * n2.y,
                                 w[0][0] * n0.z + w[0][1] * n1.z + w[0][2] * n2.z};
            const float r1[3] = {w[0][1] * n1.x + wx2 * n2.x, w[0][1] * n1.y + wx2 * n2.y,
                                 w[0][1] * n1.z + wx2 * n2.z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              if (j == 0) {
                a0[c] = w[1][0] * r0[c];
                a1[c] = w[1][0] * r1[c];
              } else {
                a0[c] += w[1][j] * r0[c];
                a1[c] += w[1][j] * r1[c];
              }
              if (j == 1) ay[c] = w[1][1] * r0[c];
              if (j == 2) ay[c] += wy2 * r0[c];
            }
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            if (k == 0) {
              Vo[c] = w[2][0] * a0[c];
              Xo[c] = w[2][0] * a1[c];
              Yo[c] = w[2][0] * ay[c];
            } else {
              Vo[c] += w[2][k] * a0[c];
              Xo[c] += w[2][k] * a1[c];
              Yo[c] += w[2][k] * ay[c];
            }
            if (k == 1) Zo[c] = w[2][1] * a0[c];
            if (k == 2) Zo[c] += wz2 * a0[c];
          }
        }
      };
      float V[3], X[3], Y[3], Z[3];
      if (staged) {
        // Through an LDS-qualified pointer, so the reads are ds_read (a
        // pointer shared with the global path would make them flat loads).
        typedef const __attribute__((address_space(3))) f4 lds_f4;
        lds_f4* sv = (lds_f4*)s_vel;
        const int r00 = ((b[2] - lo[2]) * ey + (b[1] - lo[1])) * ex + (b[0] - lo[0]);
        contract([&](int k, int j, int i) { return sv[r00 + (k * ey + j) * ex + i]; }, V, X, Y, Z);
      } else {
        const f4* gv = reinterpret_cast<const f4*>(vel);
        contract([&](int k, int j, int i) { return gv[((size_t)(b[2] + k) * N + (b[1] + j)) * N + b[0] + i]; }, V,
                 X, Y, Z);
      }
      if constexpr (MODE == 2) {
        // The gather is linear in the node velocity: the part's grid deltas
        // add their own weighted sums.
        if (dl) {
          const f4* gd = reinterpret_cast<const f4*>(dl);
          float dV[3], dX[3], dY[3], dZ[3];
          contract([&](int k, int j, int i) { return gd[((size_t)(b[2] + k) * N + (b[1] + j)) * N + b[0] + i]; },
                   dV, dX, dY, dZ);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            V[c] += dV[c];
            X[c] += dX[c];
            Y[c] += dY[c];
            Z[c] += dZ[c];
          }
        }
      }
      const float s4 = 4.f * invH;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cm[c * 3 + 0] = s4 * (X[c] - V[c] * t[0]);
        Cm[c * 3 + 1] = s4 * (Y[c] - V[c] * t[1]);
        Cm[c * 3 + 2] = s4 * (Z[c] - V[c] * t[2]);
        v[c] = V[c];
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          Fn[a * 3 + c] =
              F[a * 3 + c] + dt * (Cm[a * 3] * F[c] + Cm[a * 3 + 1] * F[3 + c] + Cm[a * 3 + 2] * F[6 + c]);
#pragma unroll
      for (int a = 0; a < 3; ++a) x[a] += dt * V[a];
    }
    __syncthreads();  // the LDS box is free: the surrogate and the scatter reuse it

    bool failed = false;
    if constexpr (EXT) {
      if (gather && mat.fr_on) {
        failed = update_damage(Fn, dmg, mat);
        if (failed) model = 3;
      }
    }

    bool done = false;
    if constexpr (SUR) {
      // Learned return for the surrogate's material (J2 without hardening).
      const bool mine = gather && (int)(mt & 255u) == mlp->norm.material && (model == 4 || model == 10);
      float Tp = 0.f;
      if (mine) {
        Tp = temp0;
        ep = ep0;
      }
      const int br = surrogate_return(mlp, mine, Fn, ep, Tp, mat, tau, s_raw + S::ACC_BYTES + (tid >> 6) * S::WAVE_BYTES);
      if (br == SUR_NETWORK || br == SUR_ELASTIC) {
        new_flags |= FLAG_CACHE;
        F = Fn;
        done = true;
      }
      const uint64_t bu = __ballot(br == SUR_NETWORK), be = __ballot(br == SUR_ELASTIC),
                     bf = __ballot(br == SUR_FALLBACK);
      if (lane == 0 && (bu | be | bf)) {
        unsigned long long* slot = sur_counts + 4 * (blockIdx.x % SUR_SLOTS);
        if (bu) atomicAdd(slot + 0, (unsigned long long)__popcll(bu));
        if (be) atomicAdd(slot + 1, (unsigned long long)__popcll(be));
        if (bf) atomicAdd(slot + 2, (unsigned long long)__popcll(bf));
      }
    }
    if (gather && !done) {
      if (model >= 2) {
        ep = ep0;
        if (!failed) {
          if (model == 2 || model == 8 || model == 9) jp = ld(P, cap, P_JP, ic);
          if (model == 3) dil = ld(P, cap, P_DIL, ic);
        }
        ep += plastic_project<EXT>(Fn, model, ep, jp, dil, dt, temp0, mat, tau);
        new_flags |= FLAG_CACHE;
      } else if (MODE == 1) {
        elastic_tau(Fn, model, mat, 1.f, tau);
      }
      F = Fn;
    }

    if (valid && !pinned) {
      st(Po, cap, P_X, i, x[0]);
      st(Po, cap, P_Y, i, x[1]);
      st(Po, cap, P_Z, i, x[2]);
#pragma unroll
      for (int k = 0; k < 9; ++k) st(Po, cap, P_F0 + k, i, F[k]);
      if (model >= 2) st(Po, cap, P_EP, i, ep);
      if (model == 2 || model == 8 || model == 9 || failed) st(Po, cap, P_JP, i, jp);
      if (model == 3) st(Po, cap, P_DIL, i, dil);
      if (EXT && mat.fr_on) st(Po, cap, P_DMG, i, dmg);
      if (RG || new_flags != flags) meta_o[i] = (mt & 255u) | (new_flags << 8);
      if (MODE == 2) {
        st(P, cap, P_VX, i, v[0]);
        st(P, cap, P_VY, i, v[1]);
        st(P, cap, P_VZ, i, v[2]);
#pragma unroll
        for (int k = 0; k < 9; ++k) st(P, cap, P_C0 + k, i, Cm[k]);
        if (model >= 2) {
#pragma unroll
          for (int k = 0; k < 6; ++k) st(P, cap, P_T0 + k, i, tau[k]);
        }
      }
    }
    if constexpr (RG) {
      // The rest of what carries over a step, which this step does not
      // write: mass, volume, temperature, the plastic strain of the elastic
      // models, the history planes not updated here; a pinned particle
      // keeps its stored state and moves every plane.
      if (valid) {
        rg.ids_out[i] = rg.ids[ic];
        if (pinned) {
          for (int k = 0; k < NPLANES; ++k) st(Po, cap, k, i, ld(P, cap, k, ic));
          meta_o[i] = mt;
        } else {
          st(Po, cap, P_MASS, i, m);
          st(Po, cap, P_VOL, i, vol);
          st(Po, cap, P_TEMP, i, hist ? temp0 : ld(P, cap, P_TEMP, ic));
          if (!(model >= 2)) st(Po, cap, P_EP, i, hist ? ep0 : ld(P, cap, P_EP, ic));
          if (rg.history) {
            if (!(model == 2 || model == 8 || model == 9 || failed)) st(Po, cap, P_JP, i, ld(P, cap, P_JP, ic));
            if (model != 3) st(Po, cap, P_DIL, i, ld(P, cap, P_DIL, ic));
            if (!(EXT && mat.fr_on)) st(Po, cap, P_DMG, i, ld(P, cap, P_DMG, ic));
          }
        }
      }
    }
    if (sp.guard) {
      const bool bad = gather && state_invalid(F, ep, v, x, g);
      if (__any(bad) && lane == 0u) atomicOr(status, STATUS_INVALID);
    }
    if (MODE == 2) return;
  }

  // -------------------------------------------------------------------- P2G
  // The node box is cleared whole here (the gather's box, which shared it,
  // is done with it): the barrier of the block box below then also orders
  // the clearing before the scatter's atomics.
  {
    typedef int i4 __attribute__((ext_vector_type(4)));
    i4* z = reinterpret_cast<i4*>(s_raw);
    for (int p = (int)tid; p < S::ACC_BYTES / 16; p += FBK) z[p] = i4{0, 0, 0, 0};
  }
  if (sp.guard) {
    // The stable step of the state this step starts from.
    bool slow = false;
    if (valid) {
      // Jp as this step left it (written above for these models)
      const float jpc = (model == 2 || model == 8 || model == 9) ? ld(Po, cap, P_JP, RG ? i : ic) : 1.f;
      slow = cfl_exceeded(F, Cm, v, m, vol, mat, model, jpc, sp);
    }
    if (__any(slow) && lane == 0u) atomicOr(status, STATUS_CFL);
  }
  if (g.gravity_on_particles && !pinned) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] += g.gravity[a] * dt;
  }
  // The step stress carries (1 - damage) below failure (src/three/mpm.js:38).
  const float k4 = dt * vol * 4.f * invH * invH * ((EXT && mat.fr_on && dmg < 1.f) ? 1.f - dmg : 1.f);
  const float T9[9] = {tau[0], tau[3], tau[4], tau[3], tau[1], tau[5], tau[4], tau[5], tau[2]};
  float A[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) A[k] = m * Cm[k] - k4 * T9[k];
  float t[3], w[3][3];
  int b[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b[a] = base_axis(x[a], g.O[a], invH, N, t[a]);
    bspline(t[a], w[a]);
  }
  // Momentum of node (i,j,k): w (m v + A ((i,j,k) - t) h).
  float q0[3], c0[3], c1[3], c2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    c0[a] = A[a * 3 + 0] * H;
    c1[a] = A[a * 3 + 1] * H;
    c2[a] = A[a * 3 + 2] * H;
    q0[a] = m * v[a] - (c0[a] * t[0] + c1[a] * t[1] + c2[a] * t[2]);
  }
  // Weights of a lane past the end are zero, so all its parts are zero.
  float wyz[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float wz = valid ? w[2][k] : 0.f;
#pragma unroll
    for (int j = 0; j < 3; ++j) wyz[j][k] = w[1][j] * wz;
  }
  // Particle mass in fixed-point units, at most 2^24: a part is at most
  // 0.75^3 of it (below 2^23) and the centre remainder at least 0.125 Q - 13,
  // so every part fits a signed 24-bit factor and every sum of up to 64 parts
  // fits 32 bits.
  const float Qf = valid ? rintf(m * sp.mass_scale) : 0.f;
  const int Q = (int)Qf;

  // Base cell of this lane (the linear index of its base node) and the
  // lane's position in the wave's cell order: lanes are grouped by cell
  // (ascending, the lanes of one cell in lane order), so each cell's
  // particles form one run however far the slot order has drifted since the
  // last re-bin. Up to SORT_PASSES cells are placed by minimum passes; the
  // lanes left over, then the lanes past the end, follow in lane order. Any
  // order gives the same sums (a run only has to hold one cell); grouping
  // only cuts the number of chunks.
  const uint32_t key = valid ? ((uint32_t)b[2] * (uint32_t)N + (uint32_t)b[1]) * (uint32_t)N + (uint32_t)b[0]
                             : 0xffffffffu - lane;
  const uint64_t below = (1ull << lane) - 1ull;
  const uint64_t valid_lanes = __ballot(valid);
  // A wave already in cell order (the usual case soon after a re-bin; valid
  // lanes are a prefix) keeps its lane order and skips the passes.
  const uint32_t kleft = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)key, DPP_WAVE_SHR1, 0xf, 0xf, false);
  const bool in_order = __ballot(valid && lane > 0u && kleft > key) == 0ull;
  int dest = (int)lane, placed = 0;
  if (!in_order) {
    dest = -1;
    uint64_t rem = valid_lanes;
    for (int pass = 0; pass < SORT_PASSES && rem; ++pass) {
      const bool mine = (rem >> lane) & 1ull;
      const uint32_t kmin = (uint32_t)wave_extreme<false>(mine ? (int)key : 0x7fffffff);
      const uint64_t cell = __ballot(mine && key == kmin);
      if (mine && key == kmin) dest = placed + __popcll(cell & below);
      placed += __popcll(cell);
      rem &= ~cell;
    }
    const uint64_t rest = ~__ballot(dest >= 0);
    if (dest < 0) dest = placed + __popcll(rest & below);
  }

  // The node box sums momentum in 32-bit fixed point: integer LDS atomics
  // run at a small fraction of the cost of FP32 ones (measured: four more
  // FP32 atomics per run sum add 49% to the kernel at 8 particles per cell,
  // four more integer ones 4%). A lane's momentum part at a node is at most
  // w bnd with bnd = max_c |q0| + 2 (|c0| + |c1| + |c2|) (offsets <= 2) and
  // w <= 0.75^3 < 0.4219, so a node's sum over the FBK <= 256 particles of
  // the block is below 108 bmax, bmax the block's largest bnd. With
  // bmax < 2^e the scale qs = 2^(24 - e) keeps every node sum, and the run
  // sums' rounding (half a unit each, at most 128 per node), below 2^31;
  // the rounding step is at most 2^-23 bmax, the FP32 step of the largest
  // part. A non-finite bound sends the block to the global FP32 path, which
  // carries the NaN or infinity as before.
  float bnd = 0.f;
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float s = fabsf(q0[c]) + 2.f * (fabsf(c0[c]) + fabsf(c1[c]) + fabsf(c2[c]));
      bnd = s <= 3.0e38f ? fmaxf(bnd, s) : __int_as_float(0x7f800000);
    }
  }
  int top = __float_as_int(bnd);  // non-negative floats order as their bits
  int lo[3], hi[3];
  block_box<S::WAVES>(b, valid, lo, hi, s_box + 7 * S::WAVES, &top);  // the gather's box words stay apart
  const float bmax = __int_as_float(top);
  static_assert(FBK <= 256, "the fixed-point bound counts at most 256 particles per node");
  const int qexp = min(126, 24 - frexp_exponent(bmax));
  const float qs = __int_as_float((qexp + 127) << 23), inv_qs = __int_as_float((127 - qexp) << 23);
  int ex = hi[0] - lo[0] + 3, ey = hi[1] - lo[1] + 3, ez = hi[2] - lo[2] + 3;
  int cnt = ex * ey * ez;
  bool staged = box_inside(lo, hi, N) && cnt <= S::PATCH_ACC && bmax <= 3.0e38f;
  // A block whose box does not fit gives each wave a box of its own in a
  // quarter of the buffer: when the block's particles straddle a jump of the
  // cell order (the next row of 4x4x4 blocks), usually one wave holds the
  // jump and only that wave takes the global atomics.
  constexpr int WAVE_ACC = S::PATCH_ACC / S::WAVES;
  const bool per_wave = S::WAVES > 1 && !staged && bmax <= 3.0e38f;  // uniform over the block
  int box0 = 0;
  if (per_wave) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = wave_extreme<false>(valid ? b[a] : 0x7fffffff);
      hi[a] = wave_extreme<true>(valid ? b[a] : -1);
    }
    if (hi[0] < 0) {  // no particle in this wave: an empty box
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = 0;
        hi[a] = -1;
      }
    }
    ex = hi[0] - lo[0] + 3;
    ey = hi[1] - lo[1] + 3;
    ez = hi[2] - lo[2] + 3;
    cnt = ex * ey * ez;
    staged = box_inside(lo, hi, N) && cnt <= WAVE_ACC;  // uniform over the wave
    box0 = (int)(tid >> 6) * WAVE_ACC;
  }
  const int box_first = per_wave ? (int)lane : (int)tid, box_stride = per_wave ? 64 : FBK;
  int* s_q = reinterpret_cast<int*>(s_raw) + box0;                                  // momentum x, y, z planes
  uint32_t* s_mq = reinterpret_cast<uint32_t*>(s_raw + 12 * S::PATCH_ACC) + box0;  // fixed-point mass
  const int lbase = ((b[2] - lo[2]) * ey + (b[1] - lo[1])) * ex + (b[0] - lo[0]);
  const int gbase = (int)key;  // linear index of the base node

  // Per-wave LDS: TR_ROWS transpose rows (the nodes of a pass) of 64 particles,
  // each entry {fixed-point mass part (as bits), momentum x, y, z}, a zero
  // tail, and the chunk table {first lane, last lane, box base, grid base
  // (-1 when the chunk holds no particle)}. The run sums read whole chunks
  // and mask by multiplying with 0 or 1, so every entry they can reach past
  // a chunk's end -- the next lanes, a row's pad entry, the tail -- must hold
  // a finite value: the pads and the tail are zeroed here (the buffer held
  // the gather box or surrogate activations before), the lanes are written
  // in every slab.
  unsigned char* wbuf = s_raw + S::ACC_BYTES + (tid >> 6) * S::WAVE_BYTES;
  float4* s_tr = reinterpret_cast<float4*>(wbuf);
  using ChunkT = typename std::conditional<Scatter<CH>::PACKED, int2, int4>::type;
  ChunkT* s_chunk = reinterpret_cast<ChunkT*>(wbuf + TR_ENTRIES * 16);

  // The cell order as seen from each position: position p reads the key and
  // box base of the lane placed at p (through the first 512 bytes of the
  // transpose rows, free until the first slab is written). Valid lanes take
  // positions [0, count).
  int2 at = make_int2((int)key, lbase);
  if (!in_order) {
    int2* s_sort = reinterpret_cast<int2*>(wbuf);
    s_sort[dest] = at;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    at = s_sort[lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  const uint32_t skey = (uint32_t)at.x;
  const bool svalid = (int)lane < __popcll(valid_lanes);

  // Runs of equal cell in position order, cut into chunks of at most CHUNK
  // positions.
  const uint32_t kprev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)skey, DPP_WAVE_SHR1, 0xf, 0xf, false);
  const uint64_t run_starts = __ballot(lane == 0u || kprev != skey);
  const uint64_t upto = (lane == 63u) ? ~0ull : ((2ull << lane) - 1ull);
  const int run_start = 63 - __clzll(run_starts & upto);
  const uint64_t chunk_starts = __ballot(((int)lane - run_start) % CHUNK == 0);
  const int nchunks = __popcll(chunk_starts);

  // Zeroed once per launch: the pad entry of every row, the tail, and, when
  // a pass uses fewer rows than the buffer holds, the first CHUNK entries of
  // the first row no pass writes (the run sums of the last row in use read
  // into it).
  {
    constexpr int NZ_EXTRA = R_PASS < TR_ROWS ? CHUNK : 0;
    const int l = (int)lane;
    if (l < TR_ROWS + TR_TAIL + NZ_EXTRA) {
      const int idx = l < TR_ROWS ? l * TR_STRIDE + 64
                                  : (l < TR_ROWS + TR_TAIL ? TR_ROWS * TR_STRIDE + (l - TR_ROWS)
                                                           : R_PASS * TR_STRIDE + (l - TR_ROWS - TR_TAIL));
      s_tr[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  if ((chunk_starts >> lane) & 1ull) {
    const int c = __popcll(chunk_starts & (upto >> 1));  // chunks before this position
    const uint64_t later = chunk_starts & ~upto;
    const int next = later ? __ffsll((long long)later) - 1 : 64;
    s_chunk[c] = make_chunk(s_chunk, (int)lane, next - 1, at.y, svalid ? (int)skey : -1);  // grid base -1: no particle
  }

  // Sum of one (row, chunk) task of a pass, for node n, added to the node
  // box (or to the grid when the box does not fit). All CHUNK entries are
  // read unconditionally, so the loads issue back to back; entries past the
  // chunk's last lane (at most CHUNK - 1 past the row: the row's pad, the
  // next row, or the zero tail, all finite) are masked out by the bits of
  // `in`. Mass parts are below 2^23 and sit in an entry's low 24 bits (above
  // them an entry may hold the float bits of the 2^23 of its rounding), so
  // the 24-bit multiply-add reads and sums them exactly.
  auto run_task = [&](int row, int n, const ChunkT ch) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4* src = reinterpret_cast<const f4*>(s_tr + row * TR_STRIDE + chunk_first(ch));
    const uint32_t in = (2u << chunk_span(ch)) - 1u;
    int sq = 0;
    float px = 0.f, py = 0.f, pz = 0.f;
#pragma unroll
    for (int h = 0; h < CHUNK; h += 8) {
      f4 vals[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) vals[q] = src[h + q];
      int qfirst = 0;
      if (h == 0) {  // a chunk holds at least its first entry
        sq = __float_as_int(vals[0].x) & 0xffffff;
        px = vals[0].y;
        py = vals[0].z;
        pz = vals[0].w;
        qfirst = 1;
      }
#pragma unroll
      for (int q = qfirst; q < 8; ++q) {
        const uint32_t mi = (in >> (h + q)) & 1u;
        const float mf = (float)mi;
        sq += __mul24(__float_as_int(vals[q].x), (int)mi);
        px = fmaf(vals[q].y, mf, px);
        py = fmaf(vals[q].z, mf, py);
        pz = fmaf(vals[q].w, mf, pz);
      }
    }
    if (chunk_grid(ch) < 0) return;
    const int k = n / 9, j = (n / 3) % 3, ii = n % 3;
    const unsigned long long mq = (unsigned long long)(long long)sq;
    if (staged) {
      const int ln = chunk_box(ch) + (k * ey + j) * ex + ii;
      atomicAdd(&s_mq[ln], (uint32_t)sq);
      atomicAdd(&s_q[ln], __float2int_rn(px * qs));
      atomicAdd(&s_q[S::PATCH_ACC + ln], __float2int_rn(py * qs));
      atomicAdd(&s_q[2 * S::PATCH_ACC + ln], __float2int_rn(pz * qs));
    } else {
      const size_t node = (size_t)(uint32_t)chunk_grid(ch) + ((size_t)k * N + j) * N + ii;
      atomicAdd(&acc_m[node], mq);
      atomicAdd(&acc_p[node], px);
      atomicAdd(&acc_p[pstride + node], py);
      atomicAdd(&acc_p[2 * pstride + node], pz);
      pieces[node >> PIECE_SHIFT] = 1;
    }
  };
  // The chunk of this lane's first task is the same in every full pass:
  // read once.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const ChunkT ch0 = s_chunk[min((int)lane / R_PASS, 63)];

  // Nodes a pinned particle touches with weight above 1e-8 are pinned
  // (src/three/mpm.js:83).
  if (__any(pinned && valid)) {
    if (pinned && valid) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int ii = 0; ii < 3; ++ii)
            if (w[0][ii] * wyz[j][k] > 1e-8f) fixed[(size_t)gbase + ((size_t)k * N + j) * N + ii] = 1;
    }
  }
  // Transpose through LDS: in each pass, each lane writes its particle's
  // contributions to R_PASS nodes (in scatter_node order) as a column; then
  // every lane sums whole (row, chunk) tasks and adds each chunk sum to the
  // node box with one LDS atomic per channel. R_PASS rows of chunks of CHUNK
  // lanes keep a pass's tasks within one round of the wave's 64 lanes up to
  // 64 / R_PASS chunks (9 at 7 rows: a wave of 8 particles per cell). The
  // centre node goes last and takes the remainder of the fixed-point mass,
  // so the 27 parts sum exactly to Q.
  float c0x2[3], c1x2[3], c2x2[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    c0x2[c] = 2.f * c0[c];
    c1x2[c] = 2.f * c1[c];
    c2x2[c] = 2.f * c2[c];
  }
  uint32_t qsum = 0u;  // the other parts' entries, each 2^23 (as float bits) plus the part
#pragma unroll
  for (int p = 0; p < NPASS; ++p) {
    constexpr int R_LAST = 27 - (NPASS - 1) * R_PASS;
    const int rows = p == NPASS - 1 ? R_LAST : R_PASS;
#pragma unroll
    for (int r = 0; r < R_PASS; ++r) {
      const int g = p * R_PASS + r;
      if (g < 27) {
        const int n = scatter_node(g);
        const int k = n / 9, j = (n / 3) % 3, ii = n % 3;
        const float wt = w[0][ii] * wyz[j][k];
        int qm;
        if (g == 26) {
          qm = (int)((uint32_t)Q + 26u * 0x4b000000u - qsum);
        } else {
          // rint(wt Qf) for 0 <= wt Qf < 2^23: the sum with 2^23 rounds to
          // the nearest integer and its low mantissa bits hold it. The entry
          // keeps the float's bits: the run sums read the part from the low
          // 24 bits.
          qm = __float_as_int(fmaf(wt, Qf, 8388608.f));
          qsum += (uint32_t)qm;
        }
        // q0 + ii c0 + j c1 + k c2 without multiplying by zero offsets (the
        // row part is shared by the three nodes of a row).
        float pc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float rowv = q0[c];
          if (j) rowv += j == 1 ? c1[c] : c1x2[c];
          if (k) rowv += k == 1 ? c2[c] : c2x2[c];
          if (ii) rowv += ii == 1 ? c0[c] : c0x2[c];
          pc[c] = wt * rowv;
        }
        s_tr[r * TR_STRIDE + dest] = make_float4(__int_as_float(qm), pc[0], pc[1], pc[2]);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int ntask = rows * nchunks;
    for (int t = (int)lane; t < ntask; t += 64) {
      ChunkT ch = ch0;
      if (t >= 64 || rows != R_PASS) ch = s_chunk[t / rows];
      const int row = t % rows;
      run_task(row, scatter_node(p * R_PASS + row), ch);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (staged || per_wave) {  // uniform over the block
    __syncthreads();
    const BoxIndex at(ex, ey);
    for (int p = box_first; p < (staged ? cnt : 0); p += box_stride) {
      const unsigned long long mq = s_mq[p];
      const int qx = s_q[p], qy = s_q[S::PATCH_ACC + p], qz = s_q[2 * S::PATCH_ACC + p];
      if (mq != 0ull || qx != 0 || qy != 0 || qz != 0) {
        const float px = (float)qx * inv_qs, py = (float)qy * inv_qs, pz = (float)qz * inv_qs;
        int lx, ly, lz;
        at(p, lx, ly, lz);
        const size_t node = ((size_t)(lo[2] + lz) * N + (lo[1] + ly)) * N + (lo[0] + lx);
        atomicAdd(&acc_m[node], mq);
        atomicAdd(&acc_p[node], px);
        atomicAdd(&acc_p[pstride + node], py);
        atomicAdd(&acc_p[2 * pstride + node], pz);
        pieces[node >> PIECE_SHIFT] = 1;
      }
    }
  }
}

// Workgroups go to the chip's 8 XCDs in turn (workgroup b to XCD b mod 8),
// and each XCD has its own L2. The chunk of workgroup b is chosen so that
// every XCD takes one contiguous range of chunks: neighbouring chunks of the
// cell order, whose node boxes overlap, then share an L2.
constexpr uint32_t XCDS = 8;
__device__ __forceinline__ uint32_t xcd_chunk() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t x = b % XCDS, k = b / XCDS, q = nb / XCDS, r = nb % XCDS;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Single scene: one chunk of FB particles per workgroup. Four workgroups per
// CU fit the LDS (4 x 39.6 KB of 160 KB) and the register budget is held to
// the same four waves per SIMD (the compiler then fits 116 VGPRs without
// spills; left at three waves it took about 148, and four waves ran the large
// scene's fused step 13% faster). The surrogate variant's wave buffer holds
// the network's activations (47.2 KB per workgroup) and its MFMA operands
// need about 160 VGPRs: three waves.
// MODE 3 is the fused step of mode 1 re-gathering the particles after a
// re-bin sort (its own kernel, so that profiles keep it apart).
template <int MODE, bool SUR, bool EXT, int CH>
__global__ __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(SUR ? 3 : 4, SUR ? 3 : 4))) void k_fused(
    float* __restrict__ P, uint32_t* __restrict__ meta, uint32_t cap, uint32_t n,
    const DevMaterial* __restrict__ mats, GridDims g, StepParams sp, const float4* __restrict__ vel,
    unsigned long long* __restrict__ acc_m, float* __restrict__ acc_p, uint8_t* __restrict__ fixed,
    uint8_t* __restrict__ pieces, const MlpDev* __restrict__ mlp, unsigned long long* __restrict__ sur_counts,
    uint32_t* __restrict__ status, const float4* __restrict__ deltas, const int32_t* __restrict__ delta_slot,
    Regather rg) {
  __shared__ __align__(16) unsigned char s_raw[Shape<FB, SUR, CH>::LDS_BYTES];
  __shared__ int s_box[2 * 7 * Shape<FB, SUR, CH>::WAVES];
  fused_chunk<MODE == 3 ? 1 : MODE, SUR, FB, EXT, CH, MODE == 3>(
      P, meta, cap, xcd_chunk() * FB, n, mats, g, sp, (size_t)g.N * g.N * g.N, vel, acc_m, acc_p, fixed, pieces, mlp,
      sur_counts, s_raw, s_box, status, deltas, delta_slot, rg);
}

// Many scenes in one launch: one workgroup per chunk of the bucket's chunk
// list (chunks never straddle scenes). A persistent form (workgroups taking
// chunks from a counter) kept the loop's invariants live across chunks and
// needed about 45 more VGPRs, which spilled at three waves per SIMD. Four
// waves per SIMD, as the single-scene kernel.
template <int MODE, int FBK, bool EXT, int CH>
__global__ __launch_bounds__(FBK) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_fused_batch(
    float* __restrict__ P, uint32_t* __restrict__ meta, uint32_t cap, const DevMaterial* __restrict__ mats,
    BatchArgs b, const float4* __restrict__ vel, unsigned long long* __restrict__ acc_m, float* __restrict__ acc_p,
    uint8_t* __restrict__ fixed, uint8_t* __restrict__ pieces, uint32_t* __restrict__ status, Regather rg) {
  __shared__ __align__(16) unsigned char s_raw[Shape<FBK, false, CH>::LDS_BYTES];
  __shared__ int s_box[2 * 7 * Shape<FBK, false, CH>::WAVES];
  const uint32_t t = xcd_chunk();
  if (t >= b.nchunks) return;
  // The chunk and its scene's parameters are the same for every lane, but
  // the kernel writes memory, so the compiler reads them with vector loads;
  // moving them to scalar registers keeps ~20 uniform words and the
  // addresses built from them out of the VGPR budget.
  const uint2 ch = to_scalar(b.chunks[t]);
  const SceneDesc* sd = b.scenes + ch.x;
  const GridDims g = to_scalar(sd->g);
  const StepParams sp = to_scalar(sd->sp);
  const uint32_t node0 = __builtin_amdgcn_readfirstlane(sd->node0);
  const uint32_t end = __builtin_amdgcn_readfirstlane(sd->p0 + sd->n);
  fused_chunk<MODE == 3 ? 1 : MODE, false, FBK, EXT, CH, MODE == 3>(
      P, meta, cap, ch.y, end, mats, g, sp, b.node_stride, vel + node0, acc_m + node0, acc_p + node0, fixed + node0,
      pieces + (node0 >> PIECE_SHIFT), nullptr, nullptr, s_raw, s_box, status + ch.x, nullptr, nullptr, rg);
}

inline unsigned fblocks(uint32_t n) { return (n + FB - 1) / FB; }

Regather regather_of(const DevState& d, int mode) {
  Regather r;
  if (mode == 3) {
    r.order = d.vals;
    r.planes_out = d.planes_alt;
    r.meta_out = d.meta_alt;
    r.ids = d.ids;
    r.ids_out = d.ids_alt;
    r.history = d.history ? 1u : 0u;
  }
  return r;
}

}  // namespace

void swap_particle_buffers(DevState& d) {
  std::swap(d.planes, d.planes_alt);
  std::swap(d.ids, d.ids_alt);
  std::swap(d.meta, d.meta_alt);
}

void launch_fused(hipStream_t s, const DevState& d, uint32_t n, const GridDims& g, const StepParams& p, int mode,
                  LaunchEvents ev) {
  if (!n) return;
  const dim3 grid(fblocks(n)), block(FB);
#define MPM_FUSED(M, S, X, C)                                                                                      \
  hipExtLaunchKernelGGL((k_fused<M, S, X, C>), grid, block, 0, s, ev.start, ev.stop, 0, d.planes, d.meta, d.cap, n, \
                        d.materials, g, p, d.vel, d.acc_m, d.acc_p, d.fixed, d.pieces, d.mlp, d.sur_counts,         \
                        d.status, mode == 2 ? d.deltas : nullptr, d.delta_slot, rg)
#define MPM_FUSED_C(M, S, X)  \
  if (d.narrow)               \
    MPM_FUSED(M, S, X, 8);    \
  else                        \
    MPM_FUSED(M, S, X, 16)
#define MPM_FUSED_X(M, S)        \
  if (d.ext_laws) {              \
    MPM_FUSED_C(M, S, true);     \
  } else {                       \
    MPM_FUSED_C(M, S, false);    \
  }
  const bool sur = d.mlp != nullptr;
  const Regather rg = regather_of(d, mode);
  if (mode == 0) {
    MPM_FUSED_X(0, false);
  } else if (mode == 1 && sur) {
    MPM_FUSED_X(1, true);
  } else if (mode == 1) {
    MPM_FUSED_X(1, false);
  } else if (mode == 3 && sur) {
    MPM_FUSED_X(3, true);
  } else if (mode == 3) {
    MPM_FUSED_X(3, false);
  } else if (sur) {
    MPM_FUSED_X(2, true);
  } else {
    MPM_FUSED_X(2, false);
  }
#undef MPM_FUSED_X
#undef MPM_FUSED_C
#undef MPM_FUSED
}

void launch_fused_batch(hipStream_t s, const DevState& d, const BatchArgs& b, int mode, int fb, LaunchEvents ev) {
  if (!b.nchunks) return;
  const dim3 grid(b.nchunks);
#define MPM_FUSED_B(M, K, X, C)                                                                                     \
  hipExtLaunchKernelGGL((k_fused_batch<M, K, X, C>), grid, dim3(K), 0, s, ev.start, ev.stop, 0, d.planes, d.meta,   \
                        d.cap, d.materials, b, d.vel, d.acc_m, d.acc_p, d.fixed, d.pieces, d.status, rg)
#define MPM_FUSED_K(M, X, C)     \
  if (fb == 64)                  \
    MPM_FUSED_B(M, 64, X, C);    \
  else if (fb == 128)            \
    MPM_FUSED_B(M, 128, X, C);   \
  else                           \
    MPM_FUSED_B(M, 256, X, C)
#define MPM_FUSED_C(M, X)        \
  if (d.narrow) {                \
    MPM_FUSED_K(M, X, 8);        \
  } else {                       \
    MPM_FUSED_K(M, X, 16);       \
  }
#define MPM_FUSED_M(M)           \
  if (d.ext_laws) {              \
    MPM_FUSED_C(M, true);        \
  } else {                       \
    MPM_FUSED_C(M, false);       \
  }
  const Regather rg = regather_of(d, mode);
  if (mode == 0) {
    MPM_FUSED_M(0);
  } else if (mode == 1) {
    MPM_FUSED_M(1);
  } else if (mode == 3) {
    MPM_FUSED_M(3);
  } else {
    MPM_FUSED_M(2);
  }
#undef MPM_FUSED_M
#undef MPM_FUSED_C
#undef MPM_FUSED_K
#undef MPM_FUSED_B
}

}  // namespace mpm
