// Batched multi-scene engine: many independent MLS-MPM scenes of uneven
// particle counts and grid sizes advanced together. No counterpart in the
// reference, which runs one scene per Worker or process.
//
// Layout: one particle pool (slots of scene s are [p0_s, p0_s + n_s), p0 the
// prefix sum of the counts) and one concatenated grid (scene s's dense N_s^3
// block starts at node0_s, a multiple of one 256-node grid block). One radix
// sort per re-bin, segmented by scene (its tiles never straddle two scenes),
// orders each scene's slot range by the scene's blocked cell key: cells in
// 4x4x4 blocks, ties by original id.
//
// Per step: one particle launch per size bucket (scenes grouped by grid
// edge; each bucket has its own particles-per-workgroup, chosen by the
// autotuner; one workgroup per chunk), or a single launch over the chunks of
// every bucket when the buckets share a particles-per-workgroup (one tail
// instead of one per bucket), and one grid launch over all scenes. A
// whole re-bin period (re-bin, P2G, P-1 fused steps, grid, G2P) can be
// captured once as a HIP graph and replayed; launches clear no memory with
// memset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpm_engine.h"
#include "mpm_host.h"
#include "mpm_internal.h"

using namespace mpm;

namespace {

constexpr int NBUCKETS = 3;
constexpr int BUCKET_EDGE[NBUCKETS - 1] = {48, 80};  // grid edge (nodes) limits of buckets 0 and 1
constexpr uint32_t GRID_BLOCK_NODES = 256;          // nodes per block of the grid kernel

int bucket_of(int N) {
  for (int b = 0; b < NBUCKETS - 1; ++b)
    if (N <= BUCKET_EDGE[b]) return b;
  return NBUCKETS - 1;
}

}  // namespace

struct mpm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t S = 0;
  std::vector<SceneDesc> sc;
  std::vector<mpm_config> cfg;  // per scene (the guard settings)
  std::vector<double> cell;
  std::vector<int64_t> mass_q;
  std::vector<uint64_t> occupied;  // distinct base cells per scene at upload
  std::vector<char> uploaded;
  std::vector<mpm_material> materials;
  DevState d{};
  SceneDesc* d_scenes = nullptr;
  uint2* d_bin_chunks = nullptr;
  uint32_t n_bin_chunks = 0;
  struct Bucket {
    int fb = 256;
    uint2* chunks = nullptr;
    uint32_t nchunks = 0;
    uint32_t scenes = 0;
  } bucket[NBUCKETS];
  // The chunks of all buckets in scene order, when every non-empty bucket
  // has the same particles-per-workgroup `all_fb` (else n_all_chunks = 0).
  uint2* d_all_chunks = nullptr;
  uint32_t n_all_chunks = 0;
  int all_fb = 0;
  uint32_t* d_block_scene = nullptr;
  // The re-bin sort runs per scene (each scene's slots hold its particles):
  // sort tiles that never straddle two scenes, and the scenes' tile ranges.
  uint2* d_sort_tiles = nullptr;
  uint2* d_sort_segs = nullptr;
  SortSegments sort_seg{};
  uint32_t total_n = 0, total_nodes = 0, total_blocks = 0, key_bits = 0;
  uint32_t grid_blocks = 1;
  int sort_every = 16;
  int steps_since_sort = 0;
  bool needs_rebin = true;
  bool use_graph = false;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  DevState graph_after[2];  // buffer assignment after each captured period
  bool graph_ready = false;
  int graph_next = 0;
  double* red_host = nullptr;
  uint32_t* status_host = nullptr;  // pinned copy of the per-scene guard words
  std::string err;
};

namespace {

int bfail(mpm_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

#define BHIP(b, call)                                                                           \
  do {                                                                                          \
    hipError_t _st = (call);                                                                    \
    if (_st != hipSuccess) return bfail((b), MPM_ERR_HIP, "%s: %s", #call, hipGetErrorString(_st)); \
  } while (0)

// The buffer assignment that a re-bin swaps.
bool same_buffers(const DevState& a, const DevState& c) {
  return a.planes == c.planes && a.ids == c.ids && a.meta == c.meta && a.keys == c.keys && a.vals == c.vals;
}

void drop_graphs(mpm_batch* b) {
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  b->graph_ready = false;
  b->graph_next = 0;
}

// Chunk lists: 256-slot chunks for binning, and per bucket chunks of the
// bucket's particles-per-workgroup.
int build_chunks(mpm_batch* b) {
  std::vector<uint2> bin;
  std::vector<uint2> per[NBUCKETS];
  std::vector<uint2> all;
  for (int32_t s = 0; s < b->S; ++s) {
    const SceneDesc& sd = b->sc[s];
    for (uint32_t f = 0; f < sd.n; f += 256) bin.push_back(make_uint2((uint32_t)s, sd.p0 + f));
    const int k = bucket_of(sd.g.N);
    const uint32_t fb = (uint32_t)b->bucket[k].fb;
    for (uint32_t f = 0; f < sd.n; f += fb) {
      per[k].push_back(make_uint2((uint32_t)s, sd.p0 + f));
      all.push_back(make_uint2((uint32_t)s, sd.p0 + f));
    }
  }
  int all_fb = 0;
  for (int k = 0; k < NBUCKETS; ++k) {
    if (per[k].empty()) continue;
    if (all_fb == 0) {
      all_fb = b->bucket[k].fb;
    } else if (all_fb != b->bucket[k].fb) {
      all_fb = -1;
    }
  }
  if (b->d_all_chunks) BHIP(b, hipFree(b->d_all_chunks));
  b->d_all_chunks = nullptr;
  b->n_all_chunks = 0;
  b->all_fb = 0;
  if (all_fb > 0) {
    BHIP(b, hipMalloc((void**)&b->d_all_chunks, sizeof(uint2) * all.size()));
    BHIP(b, hipMemcpy(b->d_all_chunks, all.data(), sizeof(uint2) * all.size(), hipMemcpyHostToDevice));
    b->n_all_chunks = (uint32_t)all.size();
    b->all_fb = all_fb;
  }
  if (b->d_bin_chunks) BHIP(b, hipFree(b->d_bin_chunks));
  b->d_bin_chunks = nullptr;
  b->n_bin_chunks = (uint32_t)bin.size();
  if (!bin.empty()) {
    BHIP(b, hipMalloc((void**)&b->d_bin_chunks, sizeof(uint2) * bin.size()));
    BHIP(b, hipMemcpy(b->d_bin_chunks, bin.data(), sizeof(uint2) * bin.size(), hipMemcpyHostToDevice));
  }
  for (int k = 0; k < NBUCKETS; ++k) {
    auto& bk = b->bucket[k];
    if (bk.chunks) BHIP(b, hipFree(bk.chunks));
    bk.chunks = nullptr;
    bk.nchunks = (uint32_t)per[k].size();
    if (!per[k].empty()) {
      BHIP(b, hipMalloc((void**)&bk.chunks, sizeof(uint2) * per[k].size()));
      BHIP(b, hipMemcpy(bk.chunks, per[k].data(), sizeof(uint2) * per[k].size(), hipMemcpyHostToDevice));
    }
  }
  drop_graphs(b);
  return MPM_OK;
}

BatchArgs bin_args(const mpm_batch* b) {
  return BatchArgs{b->d_scenes, b->d_bin_chunks, b->n_bin_chunks, b->total_nodes};
}

BatchArgs bucket_args(const mpm_batch* b, int k) {
  const auto& bk = b->bucket[k];
  return BatchArgs{b->d_scenes, bk.chunks, bk.nchunks, b->total_nodes};
}

// Bin and sort every scene's particles into cell order: d.vals then holds,
// for each new slot, the slot the particle lives in now. A re-bin at the
// start of a call also permutes every plane; a re-bin between two fused
// steps sorts only, and the fused step after it re-gathers the particles.
void sort_keys(mpm_batch* b) {
  launch_bin_keys_batch(b->stream, b->d, bin_args(b));
  radix_sort_pairs(b->stream, b->d.keys, b->d.vals, b->d.keys_alt, b->d.vals_alt, b->d.hist, b->d.partials,
                   b->total_n, b->key_bits, &b->sort_seg);
  b->steps_since_sort = 0;
}

void rebin(mpm_batch* b) {
  sort_keys(b);
  launch_permute(b->stream, b->d, b->total_n, b->d.vals);
}

// Mode 3 (the fused step re-gathering after a sort) writes the alternate
// buffers in every launch and swaps them after the last.
void particles(mpm_batch* b, int mode) {
  if (b->n_all_chunks) {
    launch_fused_batch(b->stream, b->d, BatchArgs{b->d_scenes, b->d_all_chunks, b->n_all_chunks, b->total_nodes}, mode,
                       b->all_fb);
  } else {
    for (int k = 0; k < NBUCKETS; ++k)
      if (b->bucket[k].nchunks) launch_fused_batch(b->stream, b->d, bucket_args(b, k), mode, b->bucket[k].fb);
  }
  if (mode == 3) swap_particle_buffers(b->d);
}

void grid(mpm_batch* b) {
  launch_compact(b->stream, b->d, pieces_of_nodes(b->total_nodes), (uint32_t)b->S);
  launch_grid_batch(b->stream, b->d, b->d_scenes, b->d_block_scene, b->total_nodes, b->grid_blocks);
}

// The steps of a call: the first P2G reads the stored state, every step but
// the last is a grid update and a fused step (G2P of this step, P2G of the
// next), and the last step's G2P writes the whole state back. A re-bin is
// due before the fused step once sort_every steps have passed since the
// last one.
void begin_steps(mpm_batch* b) {
  if (b->needs_rebin || b->steps_since_sort >= b->sort_every) rebin(b);
  particles(b, 0);
}

void fused_step(mpm_batch* b) {
  grid(b);
  if (++b->steps_since_sort >= b->sort_every) {
    sort_keys(b);
    particles(b, 3);
  } else {
    particles(b, 1);
  }
}

void end_steps(mpm_batch* b) {
  grid(b);
  particles(b, 2);
  b->steps_since_sort++;
}

// One re-bin period of P fused steps, starting with the re-bin that is due:
// the launch sequence a graph captures. The grid holds a P2G on entry and on
// exit, and the next re-bin is due P steps later.
void period(mpm_batch* b, int P) {
  grid(b);
  sort_keys(b);
  particles(b, 3);
  for (int s = 1; s < P; ++s) {
    grid(b);
    particles(b, 1);
  }
  b->steps_since_sort = P - 1;
}

int capture_period(mpm_batch* b, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  const int since = b->steps_since_sort;
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  period(b, b->sort_every);
  b->steps_since_sort = since;
  BHIP(b, hipStreamEndCapture(b->stream, &g));
  hipError_t st = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (st != hipSuccess) return bfail(b, MPM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(st));
  return MPM_OK;
}

// Two graphs, because a re-bin swaps the particle buffers: graph 0 starts
// from the current assignment, graph 1 from the swapped one, and the pair
// returns to the start.
int prepare_graphs(mpm_batch* b) {
  if (b->graph_ready) return MPM_OK;
  drop_graphs(b);
  const DevState before = b->d;
  int rc = capture_period(b, &b->graph[0]);
  b->graph_after[0] = b->d;
  if (!rc) rc = capture_period(b, &b->graph[1]);
  b->graph_after[1] = b->d;
  if (rc) return rc;
  if (!same_buffers(b->d, before) || same_buffers(b->graph_after[0], before))
    return bfail(b, MPM_ERR_HIP, "graph pair does not alternate between two buffer assignments");
  b->graph_ready = true;
  b->graph_next = 0;
  return MPM_OK;
}

int upload_scenes_table(mpm_batch* b) {
  BHIP(b, hipMemcpyAsync(b->d_scenes, b->sc.data(), sizeof(SceneDesc) * b->S, hipMemcpyHostToDevice, b->stream));
  return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_batch_create(const mpm_config* scenes, const uint32_t* counts, int32_t nscenes,
                     const mpm_batch_options* opt, mpm_batch** out) {
  if (!scenes || !counts || !out || nscenes < 1 || nscenes > 65536) return MPM_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* b = new mpm_batch();
  b->S = nscenes;
  b->sc.resize(nscenes);
  b->cfg.assign(scenes, scenes + nscenes);
  b->cell.resize(nscenes);
  b->mass_q.assign(nscenes, 0);
  b->occupied.assign(nscenes, 0);
  b->uploaded.assign(nscenes, 0);
  uint64_t p0 = 0, node0 = 0, block0 = 0, keys_max = 1;
  for (int32_t s = 0; s < nscenes; ++s) {
    SceneDesc& sd = b->sc[s];
    const int rc = grid_geometry(&scenes[s], sd.g, sd.lo_idx, sd.hi_idx);
    if (rc) {
      delete b;
      return rc;
    }
    sd.sp = StepParams{0.f, 1.f, 1.f};
    sd.p0 = (uint32_t)p0;
    sd.n = counts[s];
    sd.node0 = (uint32_t)node0;
    sd.block0 = (uint32_t)block0;
    keys_max = std::max<uint64_t>(keys_max, blocked_key_count(sd.g.N));
    const uint64_t Nn = (uint64_t)sd.g.N * sd.g.N * sd.g.N;
    const uint64_t blocks = (Nn + GRID_BLOCK_NODES - 1) / GRID_BLOCK_NODES;
    p0 += counts[s];
    node0 += blocks * GRID_BLOCK_NODES;
    block0 += blocks;
    b->cell[s] = scenes[s].cell_size;
  }
  if (p0 > (1ull << 31) || node0 > (1ull << 31)) {
    delete b;
    return MPM_ERR_CAPACITY;
  }
  b->total_n = (uint32_t)p0;
  b->total_nodes = (uint32_t)node0;
  b->total_blocks = (uint32_t)block0;
  // Keys are scene-local blocked cell keys, sorted per scene: the bits of
  // the largest scene's key range.
  b->key_bits = 1;
  while (b->key_bits < 32 && (1ull << b->key_bits) < keys_max) ++b->key_bits;
  if (opt) {
    if (opt->sort_every >= 1) b->sort_every = opt->sort_every;
    b->use_graph = opt->use_graph != 0;
  }
  hipError_t st = hipGetDevice(&b->device);
  int cus = 256;
  if (st == hipSuccess) st = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (st == hipSuccess) st = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  DevState& d = b->d;
  const uint32_t cap = std::max<uint32_t>(b->total_n, 1);
  d.cap = cap;
  auto alloc = [&](void** p, size_t bytes) {
    if (st == hipSuccess) st = hipMalloc(p, std::max<size_t>(bytes, 256));
  };
  const size_t Nt = b->total_nodes;
  alloc((void**)&d.planes, sizeof(float) * NPLANES * (size_t)cap);
  alloc((void**)&d.planes_alt, sizeof(float) * NPLANES * (size_t)cap);
  alloc((void**)&d.ids, sizeof(uint32_t) * cap);
  alloc((void**)&d.ids_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.meta, sizeof(uint32_t) * cap);
  alloc((void**)&d.meta_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.acc_m, sizeof(unsigned long long) * Nt);
  alloc((void**)&d.acc_p, sizeof(float) * 3 * Nt);
  alloc((void**)&d.vel, sizeof(float4) * Nt);
  alloc((void**)&d.fixed, Nt);
  alloc((void**)&d.pieces, piece_flag_bytes(pieces_of_nodes(Nt)));
  alloc((void**)&d.keys, sizeof(uint32_t) * cap);
  alloc((void**)&d.keys_alt, sizeof(uint32_t) * cap);
  alloc((void**)&d.vals, sizeof(uint32_t) * cap);
  alloc((void**)&d.vals_alt, sizeof(uint32_t) * cap);
  // the sort's tiles: at most one partial tile per scene beyond cap / tile
  const size_t hist_words = radix_hist_words((uint32_t)std::min<uint64_t>(
                                (uint64_t)cap + (uint64_t)nscenes * RADIX_TILE, 0xffffffffull)),
               part_words = radix_partial_words(cap);
  std::vector<uint2> sort_tiles, sort_segs;
  {
    std::vector<uint32_t> sizes(nscenes);
    for (int32_t s = 0; s < nscenes; ++s) sizes[s] = b->sc[s].n;
    radix_segment_tables(sizes.data(), (uint32_t)nscenes, sort_tiles, sort_segs);
  }
  alloc((void**)&b->d_sort_tiles, sizeof(uint2) * sort_tiles.size());
  alloc((void**)&b->d_sort_segs, sizeof(uint2) * sort_segs.size());
  if (st == hipSuccess && !sort_tiles.empty())
    st = hipMemcpyAsync(b->d_sort_tiles, sort_tiles.data(), sizeof(uint2) * sort_tiles.size(), hipMemcpyHostToDevice,
                        b->stream);
  if (st == hipSuccess && !sort_segs.empty())
    st = hipMemcpyAsync(b->d_sort_segs, sort_segs.data(), sizeof(uint2) * sort_segs.size(), hipMemcpyHostToDevice,
                        b->stream);
  b->sort_seg = SortSegments{b->d_sort_tiles, b->d_sort_segs, (uint32_t)sort_tiles.size(), (uint32_t)sort_segs.size()};
  alloc((void**)&d.hist, sizeof(uint32_t) * hist_words);
  alloc((void**)&d.partials, sizeof(uint32_t) * part_words);
  alloc((void**)&d.red, sizeof(double) * 4 * SUM_SLOTS * (size_t)nscenes);
  alloc((void**)&d.piece_list, sizeof(uint32_t) * pieces_of_nodes(Nt));
  alloc((void**)&d.piece_counters, sizeof(unsigned long long) * 3);
  alloc((void**)&d.materials, sizeof(DevMaterial) * 64);
  alloc((void**)&b->d_scenes, sizeof(SceneDesc) * nscenes);
  alloc((void**)&b->d_block_scene, sizeof(uint32_t) * b->total_blocks);
  alloc((void**)&d.status, sizeof(uint32_t) * (size_t)nscenes);
  if (st == hipSuccess)
    st = hipHostMalloc((void**)&b->red_host, sizeof(double) * 4 * SUM_SLOTS * (size_t)nscenes, hipHostMallocDefault);
  if (st == hipSuccess)
    st = hipHostMalloc((void**)&b->status_host, sizeof(uint32_t) * (size_t)nscenes, hipHostMallocDefault);
  if (st == hipSuccess) st = hipMemsetAsync(d.status, 0, sizeof(uint32_t) * (size_t)nscenes, b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.acc_m, 0, sizeof(unsigned long long) * Nt, b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.acc_p, 0, sizeof(float) * 3 * Nt, b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.vel, 0, sizeof(float4) * Nt, b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.fixed, 0, Nt, b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.pieces, 0, piece_flag_bytes(pieces_of_nodes(Nt)), b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.red, 0, sizeof(double) * 4 * SUM_SLOTS * (size_t)nscenes, b->stream);
  if (st == hipSuccess) st = hipMemsetAsync(d.piece_counters, 0, sizeof(unsigned long long) * 3, b->stream);
  // Eight grid update workgroups per CU: a batch touches pieces of every
  // scene each step.
  b->grid_blocks = (uint32_t)std::max(1, std::min<int>(8 * cus, (int)((pieces_of_nodes(Nt) + 15) / 16)));
  if (st == hipSuccess) st = hipMemsetAsync(d.planes, 0, sizeof(float) * NPLANES * (size_t)cap, b->stream);
  std::vector<uint32_t> block_scene(b->total_blocks);
  for (int32_t s = 0; s < nscenes; ++s) {
    const SceneDesc& sd = b->sc[s];
    const uint32_t end = s + 1 < nscenes ? b->sc[s + 1].block0 : b->total_blocks;
    for (uint32_t k = sd.block0; k < end; ++k) block_scene[k] = (uint32_t)s;
  }
  if (st == hipSuccess && b->total_blocks)
    st = hipMemcpyAsync(b->d_block_scene, block_scene.data(), sizeof(uint32_t) * b->total_blocks,
                        hipMemcpyHostToDevice, b->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(b->stream);
  if (st != hipSuccess) {
    mpm_batch_destroy(b);
    return MPM_ERR_HIP;
  }
  if (upload_scenes_table(b) || build_chunks(b)) {
    mpm_batch_destroy(b);
    return MPM_ERR_HIP;
  }
  mpm_material m0;
  mpm_material_defaults(MPM_MODEL_FIXED_COROTATED, 1.f, &m0);
  mpm_batch_set_materials(b, &m0, 1);
  *out = b;
  return MPM_OK;
}

int mpm_batch_destroy(mpm_batch* b) {
  if (!b) return MPM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  DevState& d = b->d;
  void* ptrs[] = {d.planes, d.planes_alt, d.ids,   d.ids_alt,  d.meta,     d.meta_alt, d.acc_m, d.acc_p,
                  d.vel,    d.fixed,      d.pieces, d.keys,    d.keys_alt, d.vals,     d.vals_alt, d.hist,
                  d.partials, d.red,      d.materials, b->d_scenes, b->d_bin_chunks, b->d_block_scene,
                  d.piece_list, d.piece_counters, d.status, b->d_sort_tiles, b->d_sort_segs};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& bk : b->bucket)
    if (bk.chunks) (void)hipFree(bk.chunks);
  if (b->d_all_chunks) (void)hipFree(b->d_all_chunks);
  if (b->red_host) (void)hipHostFree(b->red_host);
  if (b->status_host) (void)hipHostFree(b->status_host);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MPM_OK;
}

const char* mpm_batch_last_error(const mpm_batch* b) { return b ? b->err.c_str() : "null batch"; }

int mpm_batch_scene_count(const mpm_batch* b, int32_t* scenes, uint32_t* particles) {
  if (!b) return MPM_ERR_INVALID_ARGUMENT;
  if (scenes) *scenes = b->S;
  if (particles) *particles = b->total_n;
  return MPM_OK;
}

int mpm_batch_set_materials(mpm_batch* b, const mpm_material* mats, int32_t count) {
  if (!b || !mats || count < 1 || count > 64) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "1..64 materials required");
  std::vector<DevMaterial> dm(count);
  for (int i = 0; i < count; ++i) {
    const int rc = mpm_material_check(&mats[i]);
    if (rc == MPM_ERR_UNSUPPORTED)
      return bfail(b, rc, "material %d: model %d is not supported", i, mats[i].model);
    if (rc != MPM_OK)
      return bfail(b, rc, "material %d: E, mu, lambda or the model's parameters are out of range", i);
    dm[i] = to_dev_material(mats[i]);
  }
  b->materials.assign(mats, mats + count);
  bool ext = false, history = false, returns = false;
  for (int i = 0; i < count; ++i) {
    ext |= mats[i].model == 8 || mats[i].model == 9 || mats[i].fracture != 0;
    history |= keeps_history(mats[i]);
    returns |= returns_in_g2p(mats[i]);
  }
  b->d.returns = returns;
  for (auto& sd : b->sc) sd.sp.returns = returns ? 1u : 0u;
  // captured launches name the kernel variant and the planes a re-bin moves
  if (ext != b->d.ext_laws || history != b->d.history) drop_graphs(b);
  b->d.ext_laws = ext;
  b->d.history = history;
  BHIP(b, hipMemcpyAsync(b->d.materials, dm.data(), sizeof(DevMaterial) * count, hipMemcpyHostToDevice, b->stream));
  const int rc = upload_scenes_table(b);
  if (rc) return rc;
  BHIP(b, hipStreamSynchronize(b->stream));
  return MPM_OK;
}

int mpm_batch_upload(mpm_batch* b, int32_t scene, const mpm_particles* p) {
  if (!b || !p || scene < 0 || scene >= b->S) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bad scene or view");
  SceneDesc& sd = b->sc[scene];
  if (p->count != sd.n)
    return bfail(b, MPM_ERR_INVALID_ARGUMENT, "scene %d: count %u != declared %u", scene, p->count, sd.n);
  if (!particles_complete(p)) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bad dtype or missing particle array");
  const uint32_t n = sd.n;
  std::vector<float> planes((size_t)NPLANES * std::max<uint32_t>(n, 1), 0.f);
  std::vector<uint32_t> ids(n), meta(n);
  const std::string why = pack_particles(p, b->materials.size(), planes.data(), n, 0, sd.p0, ids.data(), meta.data());
  if (!why.empty()) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "scene %d: %s", scene, why.c_str());
  double mmax = 0;
  for (uint32_t i = 0; i < n; ++i) mmax = std::max(mmax, (double)planes[(size_t)P_MASS * n + i]);
  const double scale = std::ldexp(1.0, mass_exponent(mmax));
  int64_t q = 0;
  for (uint32_t i = 0; i < n; ++i) q += (int64_t)std::nearbyint(planes[(size_t)P_MASS * n + i] * (float)scale);
  b->mass_q[scene] = q;
  b->occupied[scene] = occupied_cells(planes.data(), n, n, sd.g);
  uint64_t cells = 0, particles = 0;  // over the scenes uploaded so far: one launch serves all
  for (int32_t s = 0; s < b->S; ++s) {
    if (s != scene && !b->uploaded[s]) continue;
    cells += b->occupied[s];
    particles += b->sc[s].n;
  }
  b->d.narrow = narrow_chunks(particles, cells);
  sd.sp.mass_scale = (float)scale;
  sd.sp.inv_mass_scale = (float)(1.0 / scale);
  if (n) {
    for (int k = 0; k < NPLANES; ++k)
      BHIP(b, hipMemcpyAsync(b->d.planes + (size_t)k * b->d.cap + sd.p0, planes.data() + (size_t)k * n,
                             sizeof(float) * n, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d.ids + sd.p0, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d.meta + sd.p0, meta.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, b->stream));
  }
  int rc = upload_scenes_table(b);
  if (rc) return rc;
  BHIP(b, hipStreamSynchronize(b->stream));
  b->uploaded[scene] = 1;
  b->needs_rebin = true;
  drop_graphs(b);
  return MPM_OK;
}

int mpm_batch_download(mpm_batch* b, int32_t scene, mpm_particles* p) {
  if (!b || !p || scene < 0 || scene >= b->S) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bad scene or view");
  const SceneDesc& sd = b->sc[scene];
  if (p->count != sd.n) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "count %u != scene's %u", p->count, sd.n);
  if (p->dtype != MPM_F64 && p->dtype != MPM_F32) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bad dtype");
  const uint32_t n = sd.n;
  if (!n) return MPM_OK;
  std::vector<float> planes((size_t)NPLANES * n);
  std::vector<uint32_t> ids(n), meta(n);
  for (int k = 0; k < NPLANES; ++k)
    BHIP(b, hipMemcpyAsync(planes.data() + (size_t)k * n, b->d.planes + (size_t)k * b->d.cap + sd.p0,
                           sizeof(float) * n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipMemcpyAsync(ids.data(), b->d.ids + sd.p0, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipMemcpyAsync(meta.data(), b->d.meta + sd.p0, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  unpack_particles(planes.data(), n, 0, ids.data(), meta.data(), n, sd.p0, p);
  return MPM_OK;
}

int mpm_batch_set_dt(mpm_batch* b, const double* dt) {
  if (!b || !dt) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "dt array required");
  for (int32_t s = 0; s < b->S; ++s)
    if (!(dt[s] > 0) || !std::isfinite(dt[s])) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "scene %d: bad dt", s);
  for (int32_t s = 0; s < b->S; ++s) {
    StepParams sp = b->sc[s].sp;
    if (guard_params(b->cfg[s], dt[s], sp))
      return bfail(b, MPM_ERR_CFL, "scene %d: MPM CFL limit exceeded: dt %g is above the stable step of any state", s,
                   dt[s]);
  }
  for (int32_t s = 0; s < b->S; ++s) {
    b->sc[s].sp.dt = (float)dt[s];
    guard_params(b->cfg[s], dt[s], b->sc[s].sp);
  }
  return upload_scenes_table(b);
}

int mpm_batch_step(mpm_batch* b, int32_t nsteps) {
  if (!b || nsteps < 0) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bad arguments");
  for (int32_t s = 0; s < b->S; ++s) {
    if (!b->uploaded[s] && b->sc[s].n) return bfail(b, MPM_ERR_NOT_READY, "scene %d has no particles uploaded", s);
    if (!(b->sc[s].sp.dt > 0)) return bfail(b, MPM_ERR_NOT_READY, "scene %d has no dt (mpm_batch_set_dt)", s);
  }
  if (nsteps == 0 || b->total_n == 0) return MPM_OK;
  const int P = b->sort_every;
  begin_steps(b);
  b->needs_rebin = false;
  int left = nsteps - 1;  // fused steps; the last step is G2P only
  while (left > 0) {
    if (b->use_graph && left >= P && b->steps_since_sort + 1 >= P) {
      // Re-bins outside the captured periods swap the buffers too: take the
      // graph of the pair that starts from the current assignment.
      if (b->graph_ready && !same_buffers(b->d, b->graph_after[b->graph_next ^ 1])) b->graph_next ^= 1;
      if (b->graph_ready && !same_buffers(b->d, b->graph_after[b->graph_next ^ 1])) drop_graphs(b);
      if (!b->graph_ready) {
        const int rc = prepare_graphs(b);
        if (rc) return rc;
      }
      BHIP(b, hipGraphLaunch(b->graph[b->graph_next], b->stream));
      b->d = b->graph_after[b->graph_next];
      b->steps_since_sort = P - 1;
      b->graph_next ^= 1;
      left -= P;
      continue;
    }
    fused_step(b);
    --left;
  }
  end_steps(b);
  BHIP(b, hipGetLastError());
  return MPM_OK;
}

int mpm_batch_debug_bin(mpm_batch* b, uint32_t* keys_in, uint32_t* keys_out, uint32_t* perm) {
  if (!b) return MPM_ERR_INVALID_ARGUMENT;
  const uint32_t n = b->total_n;
  if (!n) return MPM_OK;
  launch_bin_keys_batch(b->stream, b->d, bin_args(b));
  if (keys_in) BHIP(b, hipMemcpyAsync(keys_in, b->d.keys, 4ull * n, hipMemcpyDeviceToHost, b->stream));
  std::vector<uint32_t> ids(n), order(n);
  BHIP(b, hipMemcpyAsync(ids.data(), b->d.ids, 4ull * n, hipMemcpyDeviceToHost, b->stream));
  // local copies: the batch keeps its buffer assignment (graphs track it)
  uint32_t *k = b->d.keys, *v = b->d.vals, *k_alt = b->d.keys_alt, *v_alt = b->d.vals_alt;
  radix_sort_pairs(b->stream, k, v, k_alt, v_alt, b->d.hist, b->d.partials, n, b->key_bits, &b->sort_seg);
  if (keys_out) BHIP(b, hipMemcpyAsync(keys_out, k, 4ull * n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipMemcpyAsync(order.data(), v, 4ull * n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(b->stream));
  if (perm)
    for (uint32_t s = 0; s < n; ++s) perm[s] = ids[order[s]];
  return MPM_OK;
}

int mpm_batch_synchronize(mpm_batch* b) {
  if (!b) return MPM_ERR_INVALID_ARGUMENT;
  BHIP(b, hipMemcpyAsync(b->status_host, b->d.status, sizeof(uint32_t) * (size_t)b->S, hipMemcpyDeviceToHost,
                         b->stream));
  launch_zero_u32(b->stream, b->d.status, (uint32_t)b->S);
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(b->stream));
  int first_cfl = -1;
  for (int32_t s = 0; s < b->S; ++s) {
    if (b->status_host[s] & STATUS_INVALID)
      return bfail(b, MPM_ERR_INVALID_STATE,
                   "scene %d: GPU state became invalid: det F outside [1e-6, 1e6], a value that is not a number, "
                   "|v| above 1e5 or a particle outside the domain", s);
    if ((b->status_host[s] & STATUS_CFL) && first_cfl < 0) first_cfl = s;
  }
  if (first_cfl >= 0) return bfail(b, MPM_ERR_CFL, "scene %d: MPM CFL limit exceeded; reduce dt", first_cfl);
  return MPM_OK;
}

int mpm_batch_conservation(mpm_batch* b, int32_t scene, mpm_conservation* out) {
  if (!b || !out || scene < 0 || scene >= b->S) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bad scene");
  const SceneDesc& sd = b->sc[scene];
  const int K = particle_reduce_k(), nb = particle_reduce_blocks();
  double* dbuf = nullptr;
  BHIP(b, hipMallocAsync((void**)&dbuf, sizeof(double) * K * nb, b->stream));
  DevState dd = b->d;
  dd.planes = b->d.planes + sd.p0;
  const float D = (float)(b->cell[scene] * b->cell[scene] / 4);
  launch_particle_reduce(b->stream, dd, sd.n, sd.sp.mass_scale, D, dbuf);
  std::vector<double> h((size_t)K * nb);
  hipError_t st = hipMemcpyAsync(h.data(), dbuf, sizeof(double) * K * nb, hipMemcpyDeviceToHost, b->stream);
  const size_t slot0 = (size_t)scene * SUM_SLOTS * 4;
  if (st == hipSuccess)
    st = hipMemcpyAsync(b->red_host + slot0, b->d.red + slot0, sizeof(double) * 4 * SUM_SLOTS, hipMemcpyDeviceToHost,
                        b->stream);
  (void)hipFreeAsync(dbuf, b->stream);
  if (st == hipSuccess) st = hipStreamSynchronize(b->stream);
  if (st != hipSuccess) return bfail(b, MPM_ERR_HIP, "conservation: %s", hipGetErrorString(st));
  finish_conservation(h.data(), K, nb, b->red_host + slot0, SUM_SLOTS, b->mass_q[scene], sd.sp.mass_scale, sd.n,
                      out);
  return MPM_OK;
}

int mpm_batch_get_tuning(const mpm_batch* b, mpm_batch_tuning* out) {
  if (!b || !out) return MPM_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof *out);
  out->buckets = NBUCKETS;
  for (int k = 0; k < NBUCKETS; ++k) {
    out->bucket_edge[k] = k < NBUCKETS - 1 ? BUCKET_EDGE[k] : 0;
    out->particles_per_group[k] = b->bucket[k].fb;
    out->bucket_scenes[k] = 0;
  }
  for (int32_t s = 0; s < b->S; ++s) out->bucket_scenes[bucket_of(b->sc[s].g.N)]++;
  out->sort_every = b->sort_every;
  out->particle_blocks = 0;  // one workgroup per chunk: the chunk counts
  for (const auto& bk : b->bucket) out->particle_blocks += (int32_t)bk.nchunks;
  return MPM_OK;
}

int mpm_batch_set_tuning(mpm_batch* b, const mpm_batch_tuning* t) {
  if (!b || !t) return MPM_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < NBUCKETS; ++k) {
    const int fb = t->particles_per_group[k];
    if (fb != 64 && fb != 128 && fb != 256)
      return bfail(b, MPM_ERR_INVALID_ARGUMENT, "bucket %d: particles per group must be 64, 128 or 256", k);
  }
  if (t->sort_every < 1) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "sort_every must be >= 1");
  if (t->particle_blocks < 0) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "particle_blocks must be >= 0");
  for (int k = 0; k < NBUCKETS; ++k) b->bucket[k].fb = t->particles_per_group[k];
  b->sort_every = t->sort_every;
  return build_chunks(b);
}

// Autotuner: times each candidate on the resident state (which it advances)
// with events on the engine stream and keeps the fastest. Per size bucket,
// the particles per workgroup in {64, 128, 256} are timed on that bucket's
// particle launches alone (one G2P+P2G launch per trial step after a warm
// step); then the re-bin period in {4, 8, 16, 32, 64} is timed on whole steps.
int mpm_batch_tune(mpm_batch* b, int32_t trial_steps, mpm_batch_tuning* out) {
  if (!b || trial_steps < 1) return bfail(b, MPM_ERR_INVALID_ARGUMENT, "trial_steps must be >= 1");
  for (int32_t s = 0; s < b->S; ++s)
    if ((!b->uploaded[s] && b->sc[s].n) || !(b->sc[s].sp.dt > 0))
      return bfail(b, MPM_ERR_NOT_READY, "scene %d not ready for tuning", s);
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  std::vector<hipEvent_t> ev(2 * trial_steps);
  for (auto& x : ev) BHIP(b, hipEventCreate(&x));
  const bool graph = b->use_graph;
  b->use_graph = false;
  drop_graphs(b);
  double best_fb_ms[NBUCKETS];
  for (int k = 0; k < NBUCKETS; ++k) {
    best_fb_ms[k] = 0;
    if (!b->bucket[k].nchunks) continue;
    int best = b->bucket[k].fb;
    double best_ms = 1e300;
    for (int fb : {64, 128, 256}) {
      b->bucket[k].fb = fb;
      int rc = build_chunks(b);
      if (rc) return rc;
      rebin(b);
      particles(b, 0);
      grid(b);
      particles(b, 1);  // warm-up step
      for (int t = 0; t < trial_steps; ++t) {
        grid(b);
        for (int o = 0; o < NBUCKETS; ++o) {
          if (!b->bucket[o].nchunks) continue;
          const LaunchEvents le = o == k ? LaunchEvents{ev[2 * t], ev[2 * t + 1]} : LaunchEvents{};
          launch_fused_batch(b->stream, b->d, bucket_args(b, o), 1, b->bucket[o].fb, le);
        }
      }
      grid(b);
      particles(b, 2);
      BHIP(b, hipStreamSynchronize(b->stream));
      double total = 0;
      for (int t = 0; t < trial_steps; ++t) {
        float ms = 0;
        BHIP(b, hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]));
        total += ms;
      }
      if (total < best_ms) {
        best_ms = total;
        best = fb;
      }
    }
    b->bucket[k].fb = best;
    best_fb_ms[k] = best_ms / trial_steps;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  int rc = build_chunks(b);
  if (rc) return rc;
  int best_p = b->sort_every;
  double best_ms = 1e300;
  for (int P : {4, 8, 16, 32, 64}) {
    const int steps = std::max(P, (trial_steps + P - 1) / P * P);
    rebin(b);
    particles(b, 0);
    BHIP(b, hipEventRecord(e0, b->stream));
    for (int s = 0; s < steps; s += P) period(b, P);
    BHIP(b, hipEventRecord(e1, b->stream));
    end_steps(b);
    BHIP(b, hipEventSynchronize(e1));
    float ms = 0;
    BHIP(b, hipEventElapsedTime(&ms, e0, e1));
    if (ms / steps < best_ms) {
      best_ms = ms / steps;
      best_p = P;
    }
  }
  b->sort_every = best_p;
  b->use_graph = graph;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (out) {
    mpm_batch_get_tuning(b, out);
    out->ms_per_step = best_ms;
    for (int k = 0; k < NBUCKETS; ++k) out->bucket_ms[k] = best_fb_ms[k];
  }
  return MPM_OK;
}

}  // extern "C"
