// Stable LSD radix sort of (key, value) pairs for particle binning.
//
// Digits of up to 9 bits: a key of b bits takes ceil(b / 9) passes of equal
// width (25-bit cell keys: three 9-bit passes). Each pass: a per-tile digit
// histogram (a tile-major table, one contiguous row of digit counts per
// tile), a scan of each digit's column over the tiles plus the digits'
// bases (three small kernels over blocks of 64 tile rows), then a scatter. Tiles are 4096 keys; in the scatter each wave ranks
// its 1024 consecutive keys item by item (lanes with the same digit are
// matched with one ballot per digit bit, and a per-wave running count per
// digit in LDS gives earlier items' share), the tile's keys are placed in
// LDS in digit order, and the tile is written out from LDS in that order, so
// consecutive lanes write consecutive positions of a digit's run. Equal keys
// keep their order: the result equals std::stable_sort on the keys.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "mpm_internal.h"
#include "mpm_wave.h"

namespace mpm {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ITEMS = 16;                          // keys per lane
constexpr int RS_WAVE_KEYS = 64 * RS_ITEMS;           // consecutive keys of one wave
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;        // 4096 keys per tile
static_assert(RS_TILE == (int)RADIX_TILE, "tile size");
constexpr int RS_MAX_BITS = 9;
constexpr int RS_RADIX = 1 << RS_MAX_BITS;            // histogram table width per tile
constexpr int CS_TILES = 64;                          // tile rows per column-scan block

// Lanes of this wave whose digit equals `d`, among lanes with `valid` set
// (digits below 2^RS_MAX_BITS). Per digit bit: the lanes with the bit set
// (a ballot), and each lane keeps the lanes that agree with it, one 32-bit
// half at a time: peers & ~(ballot ^ own bit), one three-input bit operation
// per half (truth table 0x90 over operands 0xf0, 0xcc, 0xaa).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  const uint64_t v = __ballot(valid);
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
  for (int b = 0; b < RS_MAX_BITS; ++b) {
    const uint32_t own = (uint32_t)((int32_t)(d << (31 - b)) >> 31);  // all ones where this lane's bit is set
    const uint64_t m = __ballot(own != 0u);
    lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)m, own, 0x90);
    hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(m >> 32), own, 0x90);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Digit counts of one tile: hist[tile * radix + d] for d < radix. Counting
// needs no ranks: each run of equal digits in neighbouring lanes adds its
// length with one LDS atomic from its first lane (lanes of one atomic that
// share a digit are summed by the LDS).
// Keys [first, first + count) of tile t: consecutive tiles of RS_TILE keys,
// or the tile table of a segmented sort.
__device__ __forceinline__ uint2 tile_range(const uint2* __restrict__ tiles, uint32_t t, uint32_t n) {
  if (tiles) return make_uint2(__builtin_amdgcn_readfirstlane(tiles[t].x), __builtin_amdgcn_readfirstlane(tiles[t].y));
  return make_uint2(t * RS_TILE, min((uint32_t)RS_TILE, n - t * RS_TILE));
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                        uint32_t radix, const uint2* __restrict__ tiles,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[RS_RADIX];
  const uint32_t tid = threadIdx.x, tile = blockIdx.x;
  const uint32_t lane = tid & 63u;
  const uint2 tr = tile_range(tiles, tile, n);
  for (uint32_t d = tid; d < RS_RADIX; d += RS_THREADS) cnt[d] = 0;
  uint32_t k[RS_ITEMS];
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const uint32_t j = it * RS_THREADS + tid;
    k[it] = j < tr.y ? keys[tr.x + j] : 0u;
  }
  __syncthreads();
  const uint64_t above = ~1ull << lane;  // lanes after this one
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const bool valid = it * RS_THREADS + tid < tr.y;  // a prefix of the wave's lanes
    const uint32_t d = valid ? (k[it] >> shift) & (radix - 1u) : 0xffffffffu;
    // lane 0 reads a value that is no digit, so it starts a run
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)0xfffffffeu, (int)d, DPP_WAVE_SHR1, 0xf, 0xf,
                                                                false);
    const bool head = valid && d != prev;
    const uint64_t ends = __ballot(head) | ~__ballot(valid);  // a run ends before the next head or invalid lane
    if (head) {
      const uint64_t next = ends & above;
      const uint32_t len = (next ? (uint32_t)__builtin_ctzll(next) : 64u) - lane;
      atomicAdd(&cnt[d], len);
    }
  }
  __syncthreads();
  for (uint32_t d = tid; d < radix; d += RS_THREADS) hist[(size_t)tile * radix + d] = cnt[d];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out, uint32_t n,
                                                           uint32_t shift, uint32_t radix,
                                                           const uint2* __restrict__ tiles,
                                                           const uint32_t* __restrict__ offsets) {
  __shared__ uint32_t s_k[RS_TILE];
  __shared__ uint32_t s_v[RS_TILE];
  __shared__ uint32_t s_cnt[RS_WAVES][RS_RADIX];  // running counts per wave, then the wave's offset per digit
  __shared__ uint32_t s_start[RS_RADIX];          // first tile position of each digit
  __shared__ uint32_t s_glob[RS_RADIX];           // output position of tile position 0 of each digit's run
  __shared__ uint32_t s_w[RS_WAVES];
  const uint32_t tid = threadIdx.x, tile = blockIdx.x;
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  const uint32_t mask = radix - 1u;
  const uint2 tr = tile_range(tiles, tile, n);
  for (uint32_t d = tid; d < RS_RADIX; d += RS_THREADS) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) s_cnt[w][d] = 0;
    s_glob[d] = d < radix ? offsets[(size_t)tile * radix + d] : 0u;
  }
  // this wave's keys: tile positions [wave * RS_WAVE_KEYS, ...) below tr.y
  const uint32_t base = wave * RS_WAVE_KEYS;
  uint32_t k[RS_ITEMS], v[RS_ITEMS], r[RS_ITEMS];
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const uint32_t j = base + it * 64 + lane;
    k[it] = j < tr.y ? keys[tr.x + j] : 0u;
    v[it] = j < tr.y ? vals[tr.x + j] : 0u;
  }
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  uint32_t* cnt = s_cnt[wave];
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const bool valid = base + it * 64 + lane < tr.y;
    const uint32_t d = (k[it] >> shift) & mask;
    const uint64_t peers = match_digit(d, valid);
    const uint32_t before = (uint32_t)__popcll(peers & lt);
    const uint32_t run = valid ? cnt[d] : 0u;
    r[it] = run + before;
    // Only this wave touches its counts, and its LDS operations complete in
    // order: the next item's reads see this update.
    if (valid && before + 1u == (uint32_t)__popcll(peers)) cnt[d] = run + before + 1u;
  }
  __syncthreads();
  // Per digit: the waves' offsets within the digit's run, and the tile's
  // count; thread t holds digits 2t and 2t + 1.
  uint32_t tot[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t d = 2u * tid + h;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      const uint32_t c = s_cnt[w][d];
      s_cnt[w][d] = run;
      run += c;
    }
    tot[h] = run;
  }
  uint32_t inc = tot[0] + tot[1];
  const uint32_t pair = inc;
#pragma unroll
  for (int dd = 1; dd < 64; dd <<= 1) {
    const uint32_t t = __shfl_up(inc, dd, 64);
    if ((int)lane >= dd) inc += t;
  }
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t ex = inc - pair;
  for (uint32_t w = 0; w < wave; ++w) ex += s_w[w];
  s_start[2 * tid] = ex;
  s_start[2 * tid + 1] = ex + tot[0];
  s_glob[2 * tid] -= ex;
  s_glob[2 * tid + 1] -= ex + tot[0];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    if (base + it * 64 + lane < tr.y) {
      const uint32_t d = (k[it] >> shift) & mask;
      const uint32_t pos = s_start[d] + cnt[d] + r[it];
      s_k[pos] = k[it];
      s_v[pos] = v[it];
    }
  }
  __syncthreads();
  for (uint32_t j = tid; j < tr.y; j += RS_THREADS) {
    const uint32_t kk = s_k[j];
    const uint32_t pos = s_glob[(kk >> shift) & mask] + j;
    keys_out[pos] = kk;
    vals_out[pos] = s_v[j];
  }
}

// Inclusive scan across one wave of 64 lanes.
__device__ __forceinline__ uint32_t wave_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Column scan of the tile-major count table, thread d for digit d: the sum
// of each block of CS_TILES rows ...
__global__ __launch_bounds__(RS_RADIX) void k_col_reduce(const uint32_t* __restrict__ counts, uint32_t ntiles,
                                                         uint32_t radix, uint32_t* __restrict__ sums) {
  const uint32_t d = threadIdx.x, b = blockIdx.x;
  if (d >= radix) return;
  const uint32_t t0 = b * CS_TILES, t1 = min(ntiles, t0 + CS_TILES);
  uint32_t s = 0;
#pragma unroll 8
  for (uint32_t t = t0; t < t1; ++t) s += counts[(size_t)t * radix + d];
  sums[(size_t)b * radix + d] = s;
}

// ... each block's offset within its digit's column and each digit's total
// (one wave per digit, 64 blocks per round) ...
__global__ __launch_bounds__(RS_THREADS) void k_col_scan(uint32_t* __restrict__ sums, uint32_t nblk, uint32_t radix,
                                                         uint32_t* __restrict__ totals) {
  const uint32_t d = blockIdx.x * RS_WAVES + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (d >= radix) return;  // the same for the whole wave
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nblk; b0 += 64) {
    const uint32_t b = b0 + lane;
    const uint32_t x = b < nblk ? sums[(size_t)b * radix + d] : 0u;
    const uint32_t inc = wave_scan(x);
    if (b < nblk) sums[(size_t)b * radix + d] = carry + inc - x;
    carry += __shfl(inc, 63, 64);
  }
  if (lane == 0) totals[d] = carry;
}

// ... and each tile's start, in place of its counts: the totals of the
// smaller digits, plus the block's offset in the digit's column, plus the
// tile rows before it in the block.
__global__ __launch_bounds__(RS_RADIX) void k_col_down(uint32_t* __restrict__ counts, uint32_t ntiles, uint32_t radix,
                                                       const uint32_t* __restrict__ sums,
                                                       const uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_w[RS_RADIX / 64];
  const uint32_t d = threadIdx.x, b = blockIdx.x, lane = d & 63u, wave = d >> 6;
  const uint32_t tot = d < radix ? totals[d] : 0u;
  const uint32_t inc = wave_scan(tot);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t base = inc - tot;
  for (uint32_t w = 0; w < wave; ++w) base += s_w[w];
  if (d >= radix) return;
  const uint32_t t0 = b * CS_TILES, t1 = min(ntiles, t0 + CS_TILES);
  uint32_t run = base + sums[(size_t)b * radix + d];
#pragma unroll 8
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t x = counts[(size_t)t * radix + d];
    counts[(size_t)t * radix + d] = run;
    run += x;
  }
}

// Segmented sort: each tile's start per digit within its segment, in place
// of its counts (one workgroup per segment, thread d for digit d): the
// segment's first key position, plus the segment's counts of the smaller
// digits, plus the segment's tiles before this one. A thread walks its
// segment's tiles in order, which suits a batch's scenes (at most about 50
// tiles each in the benchmark's); a segment of thousands of tiles would be
// better served by the column scans over blocks of tiles.
__global__ __launch_bounds__(RS_RADIX) void k_seg_scan(uint32_t* __restrict__ counts, uint32_t radix,
                                                       const uint2* __restrict__ segs,
                                                       const uint2* __restrict__ tiles) {
  __shared__ uint32_t s_w[RS_RADIX / 64];
  const uint32_t d = threadIdx.x, lane = d & 63u, wave = d >> 6;
  const uint2 sg = segs[blockIdx.x];  // {first tile, tile count}
  uint32_t tot = 0;
  if (d < radix) {
#pragma unroll 8
    for (uint32_t t = sg.x; t < sg.x + sg.y; ++t) tot += counts[(size_t)t * radix + d];
  }
  const uint32_t inc = wave_scan(tot);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t run = inc - tot + tiles[sg.x].x;
  for (uint32_t w = 0; w < wave; ++w) run += s_w[w];
  if (d >= radix) return;
#pragma unroll 8
  for (uint32_t t = sg.x; t < sg.x + sg.y; ++t) {
    const uint32_t x = counts[(size_t)t * radix + d];
    counts[(size_t)t * radix + d] = run;
    run += x;
  }
}

}  // namespace

void radix_segment_tables(const uint32_t* sizes, uint32_t count, std::vector<uint2>& tiles, std::vector<uint2>& segs) {
  tiles.clear();
  segs.clear();
  uint32_t first = 0;
  for (uint32_t s = 0; s < count; ++s) {
    if (sizes[s]) {
      const uint32_t t0 = (uint32_t)tiles.size();
      for (uint32_t k = 0; k < sizes[s]; k += RS_TILE) tiles.push_back(make_uint2(first + k, std::min<uint32_t>(RS_TILE, sizes[s] - k)));
      segs.push_back(make_uint2(t0, (uint32_t)tiles.size() - t0));
    }
    first += sizes[s];
  }
}

size_t radix_hist_words(uint32_t n) {
  const size_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  return (ntiles > 0 ? ntiles : 1) * RS_RADIX;
}

size_t radix_partial_words(uint32_t n) {
  const size_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  return ((ntiles + CS_TILES - 1) / CS_TILES + 1) * RS_RADIX;
}

void radix_sort_pairs(hipStream_t s, uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt,
                      uint32_t* hist, uint32_t* partials, uint32_t n, uint32_t key_bits, const SortSegments* seg) {
  if (n == 0 || key_bits == 0) return;
  if (seg && !seg->ntiles) return;
  const uint32_t passes = (key_bits + RS_MAX_BITS - 1) / RS_MAX_BITS;
  const uint32_t width = (key_bits + passes - 1) / passes;
  const uint32_t radix = 1u << width;
  const uint32_t ntiles = seg ? seg->ntiles : (n + RS_TILE - 1) / RS_TILE;
  const uint2* tiles = seg ? seg->tiles : nullptr;
  const uint32_t nblk = (ntiles + CS_TILES - 1) / CS_TILES;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t shift = p * width;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, keys, n, shift, radix, tiles, hist);
    if (seg) {
      hipLaunchKernelGGL(k_seg_scan, dim3(seg->nsegs), dim3(RS_RADIX), 0, s, hist, radix, seg->segs, seg->tiles);
    } else {
      hipLaunchKernelGGL(k_col_reduce, dim3(nblk), dim3(RS_RADIX), 0, s, hist, ntiles, radix, partials);
      hipLaunchKernelGGL(k_col_scan, dim3((radix + RS_WAVES - 1) / RS_WAVES), dim3(RS_THREADS), 0, s, partials, nblk,
                         radix, partials + (size_t)nblk * radix);
      hipLaunchKernelGGL(k_col_down, dim3(nblk), dim3(RS_RADIX), 0, s, hist, ntiles, radix, partials,
                         partials + (size_t)nblk * radix);
    }
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, keys, vals, keys_alt, vals_alt, n, shift,
                       radix, tiles, hist);
    std::swap(keys, keys_alt);
    std::swap(vals, vals_alt);
  }
}

}  // namespace mpm

// ---- the sort without allocation (mpm_sort_pairs_noalloc) ----
//
// Scratch of the sort lives in one file-scope device array: the second key
// and value buffers, the digit histogram table and the scan partials, sized
// for MPM_SORT_NOALLOC_MAX keys (about 8.5 MiB, held on each device the
// library runs on until the process ends). Every call overwrites all the
// scratch it reads, so nothing has to be cleared between calls; calls that
// overlap on one device share the array and race.
namespace {

using namespace mpm;

constexpr size_t NA_MAX = MPM_SORT_NOALLOC_MAX;
constexpr size_t NA_TILES = (NA_MAX + RS_TILE - 1) / RS_TILE;
constexpr size_t NA_HIST = NA_TILES * RS_RADIX;
constexpr size_t NA_PARTIALS = ((NA_TILES + CS_TILES - 1) / CS_TILES + 1) * RS_RADIX;
constexpr size_t NA_WORDS = 2 * NA_MAX + NA_HIST + NA_PARTIALS;

__device__ __align__(256) uint32_t scratch[NA_WORDS];

// The array's device address per device. A host reference to `scratch` is a
// host-side stand-in, not a device address: only these pointers may reach a
// kernel or a copy.
constexpr int NA_DEVICES = 64;
uint32_t* g_scratch[NA_DEVICES] = {};
std::mutex g_scratch_lock;

// Resolves the array's address on the current device the first time the
// device calls; later calls read the cached pointer.
hipError_t scratch_on_device(uint32_t** out) {
  int dev = 0;
  hipError_t st = hipGetDevice(&dev);
  if (st != hipSuccess) return st;
  if (dev < 0 || dev >= NA_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_lock);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    st = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (st != hipSuccess) return st;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch[dev] = static_cast<uint32_t*>(ptr);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

// keys_out = keys_in, perm = 0..n-1 (keys_in may alias keys_out).
__global__ __launch_bounds__(256) void k_na_init(const uint32_t* keys_in, uint32_t* keys_out,
                                                 uint32_t* __restrict__ perm, uint32_t n) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    keys_out[i] = keys_in[i];
    perm[i] = i;
  }
}

__global__ __launch_bounds__(256) void k_na_copy(const uint32_t* __restrict__ k_src, const uint32_t* __restrict__ v_src,
                                                 uint32_t* __restrict__ k_dst, uint32_t* __restrict__ v_dst,
                                                 uint32_t n) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    k_dst[i] = k_src[i];
    v_dst[i] = v_src[i];
  }
}

}  // namespace

extern "C" int mpm_sort_pairs_noalloc(const uint32_t* keys, uint32_t n, uint32_t key_bits, uint32_t* keys_out,
                                      uint32_t* perm_out, void* stream) {
  if (key_bits > 32) return MPM_ERR_INVALID_ARGUMENT;
  if (n > NA_MAX) return MPM_ERR_CAPACITY;
  if (n == 0) return MPM_OK;
  if (!keys || !keys_out || !perm_out) return MPM_ERR_INVALID_ARGUMENT;
  uint32_t* base = nullptr;
  if (scratch_on_device(&base) != hipSuccess) return MPM_ERR_HIP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t* k = keys_out;
  uint32_t* v = perm_out;
  uint32_t* k_alt = base;
  uint32_t* v_alt = base + NA_MAX;
  uint32_t* hist = base + 2 * NA_MAX;
  uint32_t* partials = hist + NA_HIST;
  const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, 4096u);
  hipLaunchKernelGGL(k_na_init, dim3(blocks), dim3(256), 0, s, keys, keys_out, perm_out, n);
  radix_sort_pairs(s, k, v, k_alt, v_alt, hist, partials, n, key_bits);
  if (k != keys_out)  // an odd number of passes leaves the result in the scratch buffers
    hipLaunchKernelGGL(k_na_copy, dim3(blocks), dim3(256), 0, s, k, v, keys_out, perm_out, n);
  return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}
