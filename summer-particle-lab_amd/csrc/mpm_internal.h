// Internal declarations shared by the engine's HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "mpm_device.h"
#include "mpm_engine.h"

namespace mpm {

struct MlpDev;
struct MlpNorm;

// Particle state planes (FP32, one plane of `cap` floats each, slot order).
enum Plane : int {
  P_X = 0, P_Y, P_Z,
  P_VX, P_VY, P_VZ,
  P_C0,                 // C00..C22: P_C0 + 0..8
  P_F0 = P_C0 + 9,      // F00..F22: P_F0 + 0..8
  P_T0 = P_F0 + 9,      // cached stress xx, yy, zz, xy, xz, yz: P_T0 + 0..5
  P_MASS = P_T0 + 6,
  P_VOL,
  P_TEMP,
  P_EP,
  P_JP,     // snow, pressure models: plastic volume ratio Jp (src/three/mpm.js:9)
  P_DIL,    // sand: dilation history
  P_DMG,    // tensile damage in [0, 1] (the scene's s.damage)
  NPLANES
};

// Integer planes: original particle id, and material | flags << 8.
enum : uint32_t { FLAG_PINNED = 1u, FLAG_CACHE = 2u };

struct GridDims {
  int N;              // nodes per axis (cubic grid, N^3 nodes)
  float O[3];         // origin
  float H, invH;
  float lower[3], upper[3];
  int wall[6];
  float gravity[3];
  int gravity_on_particles;
  float mass_eps;
};

struct StepParams {
  float dt;
  float mass_scale;      // fixed-point units per unit mass (2^k)
  float inv_mass_scale;
  // State guard (0: off): a particle whose wave speed exceeds wave_limit or
  // whose sum |C| exceeds grad_limit sets STATUS_CFL; an invalid state after
  // G2P sets STATUS_INVALID.
  uint32_t guard = 0;
  float wave_limit = 3.0e38f;
  float grad_limit = 3.0e38f;
  // Some material returns in G2P (models 2 and up, or fracture): the fused
  // kernel then loads ep and T with the particle's first loads.
  uint32_t returns = 1;
};

// Bits of the guard's status word (one word per scene).
enum : uint32_t { STATUS_CFL = 1u, STATUS_INVALID = 2u };

// Device pointers of one engine.
struct DevState {
  float* planes;      // NPLANES * cap
  bool ext_laws;      // some material uses damage or a pressure model (8, 9)
  bool narrow;        // sparse cells (under 12 particles each): 8-lane scatter chunks
  bool history;       // some material keeps Jp, dilation or damage (models 2, 3, 8, 9, fracture)
  bool returns;       // some material returns in G2P (models 2 and up, fracture)
  const float4* deltas;      // per-slot grid velocity deltas of a split step (N^3 each), or null
  const int32_t* delta_slot; // per material: delta slot or -1 [64]
  float* planes_alt;  // permutation target
  uint32_t* ids;      // original id per slot
  uint32_t* ids_alt;
  uint32_t* meta;     // material | flags << 8
  uint32_t* meta_alt;
  uint32_t cap;
  // grid
  unsigned long long* acc_m;  // fixed-point node mass [N^3]
  float* acc_p;               // node momentum 3 planes [3*N^3]
  float4* vel;                // node velocity (xyz) and mass (w) [N^3]
  uint8_t* fixed;             // [N^3]
  uint8_t* pieces;            // flags of the active 16-node pieces [ceil(N^3/16), padded to 16 bytes]
  // sort scratch
  uint32_t *keys, *keys_alt, *vals, *vals_alt, *hist, *partials;
  // grid sums of the last grid update, {fixed-point mass (int64 bits),
  // momentum x, y, z} per slot
  double* red;                 // grid sums: SUM_SLOTS slots of 4 per scene
  uint32_t* piece_list;        // active pieces of this step
  unsigned long long* piece_counters;  // {append position, step start, previous end}
  DevMaterial* materials;
  // learned constitutive surrogate (nullptr: off) and its per-branch
  // particle-step counters, SUR_SLOTS slots of {network, elastic, fallback, -}
  MlpDev* mlp;
  unsigned long long* sur_counts;
  // guard status words: one per scene (STATUS_* bits)
  uint32_t* status;
};

// Sort key of a base cell: cells are ordered by 4x4x4 block (blocks in
// z, y, x order), then x fastest inside the block. A run of consecutive
// cells then stays inside one block, so the node box of a chunk of
// particles is a few cells wide on every axis instead of a strip along x
// that wraps to the next row. Equal cells get equal keys (the runs of P2G
// only need that); keys are below blocked_key_count(N).
__host__ __device__ inline uint32_t blocked_cell_key(int bx, int by, int bz, int N) {
  const uint32_t nb = ((uint32_t)N + 3u) >> 2;
  const uint32_t block = (((uint32_t)bz >> 2) * nb + ((uint32_t)by >> 2)) * nb + ((uint32_t)bx >> 2);
  return (block << 6) | (((uint32_t)bz & 3u) << 4) | (((uint32_t)by & 3u) << 2) | ((uint32_t)bx & 3u);
}
__host__ __device__ inline uint64_t blocked_key_count(int N) {
  const uint64_t nb = ((uint64_t)N + 3u) >> 2;
  return nb * nb * nb * 64u;
}
// Base cell of a blocked key (inverse of blocked_cell_key).
__host__ __device__ inline void blocked_cell_decode(uint32_t key, int N, int& bx, int& by, int& bz) {
  const uint32_t nb = ((uint32_t)N + 3u) >> 2;
  const uint32_t block = key >> 6;
  bx = (int)(((block % nb) << 2) | (key & 3u));
  by = (int)((((block / nb) % nb) << 2) | ((key >> 2) & 3u));
  bz = (int)(((block / (nb * nb)) << 2) | ((key >> 4) & 3u));
}

constexpr int SUR_SLOTS = 64;
constexpr int SUM_SLOTS = 16;  // slots per scene of the grid sums (spreads the atomics)

// One scene of a batch: its grid and step parameters, its particle slot
// range [p0, p0 + n) and the first node / grid block of its grid in the
// concatenated grid arrays (node0 is a multiple of 256, one grid block).
struct SceneDesc {
  GridDims g;
  StepParams sp;
  uint32_t p0, n;
  uint32_t node0, block0;
  int lo_idx[3], hi_idx[3];
};

// Particle-kernel arguments of a batched launch: the scene table, the chunk
// list {scene, first slot} (chunks never straddle scenes) and the plane
// stride of the momentum accumulators (total nodes).
struct BatchArgs {
  const SceneDesc* scenes;
  const uint2* chunks;
  uint32_t nchunks;
  uint32_t node_stride;
};

// Settings of the frame wrapper as the kernels see them.
struct FrameDev {
  float lo[3], hi[3];
  float vlimit;
  int gmode;  // 0 none, 1 uniform, 2 attract, 3 repel
  float gstrength, gdir[3], gcenter[3], gsoft;
  float vol_unit;  // particle volume unit of the spacing rule (1/64 for the reference's V)
};

// Source of a fused step that re-gathers the particles after a re-bin sort:
// slot i of the step is slot order[i] of the current buffers, and the step
// writes slot i of the alternate buffers.
struct Regather {
  const uint32_t* order = nullptr;
  float* planes_out = nullptr;
  uint32_t* meta_out = nullptr;
  const uint32_t* ids = nullptr;
  uint32_t* ids_out = nullptr;
  uint32_t history = 0;  // Jp, dilation and damage carry over a step
};

// Optional events that a launch stamps with its own start and end (profiling
// without barrier packets between launches).
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};

// Kernel launchers (mpm_kernels.hip)
void launch_bin_keys(hipStream_t s, const DevState& d, uint32_t n, const GridDims& g);
void launch_permute(hipStream_t s, DevState& d, uint32_t n, const uint32_t* order);
// Grid update: launch_compact lists the pieces P2G flagged (and clears the
// grid sums), launch_grid / launch_grid_batch update the listed pieces with
// `blocks` workgroups, each taking a contiguous range of the list.
void launch_compact(hipStream_t s, const DevState& d, uint32_t npieces, uint32_t nscenes, LaunchEvents ev = {});
void launch_grid(hipStream_t s, const DevState& d, const GridDims& g, const StepParams& p, const int lo_idx[3],
                 const int hi_idx[3], uint32_t blocks, LaunchEvents ev = {});
// Fused particle kernel (mpm_fused.hip): mode 0 P2G from the stored state,
// 1 G2P of this step + P2G of the next, 2 G2P writing the full state, 3 the
// fused step of mode 1 right after a re-bin sort: it reads slot d.vals[i] of
// the current buffers and writes slot i of the alternate ones (every plane
// that carries over a step, all planes of a pinned particle, ids and meta),
// in place of a separate permute; swap_particle_buffers then makes the
// alternate buffers current. With d.mlp set, the G2P modes take the
// surrogate variant.
void launch_fused(hipStream_t s, const DevState& d, uint32_t n, const GridDims& g, const StepParams& p, int mode,
                  LaunchEvents ev = {});
void swap_particle_buffers(DevState& d);
uint32_t grid_pieces(const GridDims& g);
// Bytes of the piece flag array for `npieces` pieces: padded with zeros to a
// multiple of 16, since the compaction reads 16 flags per 16-byte load.
inline size_t piece_flag_bytes(size_t npieces) { return (npieces + 15) & ~(size_t)15; }
// The grid update works on pieces of 16 consecutive nodes (a quarter of a
// wave each): node n is in piece n >> PIECE_SHIFT.
constexpr int PIECE_SHIFT = 4;
constexpr uint32_t PIECE_NODES = 1u << PIECE_SHIFT;
inline uint32_t pieces_of_nodes(size_t nodes) { return (uint32_t)((nodes + PIECE_NODES - 1) >> PIECE_SHIFT); }
// Batched launchers (mpm_kernels.hip, mpm_fused.hip): all scenes of a batch
// in one launch; the particle launch runs one workgroup per chunk.
void launch_bin_keys_batch(hipStream_t s, const DevState& d, const BatchArgs& b);
void launch_grid_batch(hipStream_t s, const DevState& d, const SceneDesc* scenes, const uint32_t* block_scene,
                       uint32_t node_stride, uint32_t blocks, LaunchEvents ev = {});
void launch_fused_batch(hipStream_t s, const DevState& d, const BatchArgs& b, int mode, int fb, LaunchEvents ev = {});
// Frame wrapper kernels (mpm_kernels.hip). `ox` and `disp` are 3 planes of
// cap floats; cell_start is N^3 entries; maxbits one word.
void launch_frame_pre(hipStream_t s, const DevState& d, uint32_t n, float* ox, const FrameDev& f, float dt);
void launch_frame_post(hipStream_t s, const DevState& d, uint32_t n, const float* ox, const FrameDev& f, float dt);
void launch_regularize(hipStream_t s, const DevState& d, uint32_t n, const GridDims& g, int R, const FrameDev& f,
                       float dt, uint32_t* cell_start, float* disp, unsigned int* maxbits);
// Trained parameters packed into the device record (mpm_mlp.hip).
void pack_mlp(const float* params, const MlpNorm& norm, MlpDev& out);
void launch_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v);
// Zeroes n words with a kernel (no memset node, so the launch may be captured).
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n);
// Particle-side FP64 sums: particle_reduce_blocks() records of
// particle_reduce_k() doubles each (see k_particle_reduce).
int particle_reduce_k();
int particle_reduce_blocks();
void launch_particle_reduce(hipStream_t s, const DevState& d, uint32_t n, float mass_scale, float D, double* out);
// Per block {max wave speed, max sum |C|}; particle_reduce_blocks() records.
void launch_stable(hipStream_t s, const DevState& d, uint32_t n, float* out);
// Von Mises diagnostic per slot (out) and, when tau_out is set, the step's
// Kirchhoff stress (6 per slot); either may be null.
void launch_von_mises(hipStream_t s, const DevState& d, uint32_t n, float* out, float* tau_out = nullptr);
// Dense view of the last grid update (mass, 3 velocities per node, pinned
// flag) over the listed pieces, and velocities written back into them.
void launch_grid_export(hipStream_t s, const DevState& d, const GridDims& g, float* mass, float* v, uint8_t* fixed);
void launch_grid_import(hipStream_t s, const DevState& d, const GridDims& g, const float* v);
// Frame snapshot: rows of `data` [ceil(n / stride) x 8] (may be null), FP64
// block totals `partial` [snapshot_blocks(n) x 6] and counters `counts`
// [13] (per model, then non-finite), which the caller zeroes.
uint32_t snapshot_blocks(uint32_t n);
void launch_snapshot(hipStream_t s, const DevState& d, uint32_t n, const int32_t* codes, uint32_t stride, float* data,
                     double* partial, unsigned long long* counts);
void launch_return_batch(hipStream_t s, const DevMaterial& m, uint32_t n, float dt, const float* F, const float* ep,
                         const float* T, const float* jp, const float* dil, float* F_out, float* ep_out,
                         float* jp_out, float* dil_out, float* tau_out);
void launch_stress_batch(hipStream_t s, const DevMaterial& m, uint32_t n, const float* F, float* tau_out);
DevMaterial to_dev_material(const mpm_material& m);
// A material whose particles carry Jp (snow, the pressure models, and a
// fractured particle's sand state), dilation (sand) or damage (fracture).
inline bool keeps_history(const mpm_material& m) {
  return m.model == 2 || m.model == 3 || m.model == 8 || m.model == 9 || m.fracture != 0;
}
// A material whose particles take a plastic return (and read ep and T) in G2P.
inline bool returns_in_g2p(const mpm_material& m) { return m.model >= 2 || m.fracture != 0; }

// Radix sort (mpm_sort.hip): stable LSD sort of (keys, vals) pairs. On return
// the sorted pairs are in (keys, vals) (the buffers are swapped as needed).
size_t radix_hist_words(uint32_t n);
size_t radix_partial_words(uint32_t n);
// A sort of consecutive segments, each stably sorted on its own (the pairs of
// a segment stay in its range): tiles of at most RADIX_TILE keys that never
// straddle two segments, {first key, key count} each, and per segment
// {first tile, tile count} (segments with tiles only).
constexpr uint32_t RADIX_TILE = 4096;
struct SortSegments {
  const uint2* tiles = nullptr;
  const uint2* segs = nullptr;
  uint32_t ntiles = 0, nsegs = 0;
};
// The tiles and segments of segments of the given sizes, in order.
void radix_segment_tables(const uint32_t* sizes, uint32_t count, std::vector<uint2>& tiles, std::vector<uint2>& segs);
void radix_sort_pairs(hipStream_t s, uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt,
                      uint32_t* hist, uint32_t* partials, uint32_t n, uint32_t key_bits,
                      const SortSegments* seg = nullptr);

}  // namespace mpm
