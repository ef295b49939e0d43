// Stable LSD radix sort of (key, value) pairs for particle binning.
//
// 8-bit digits, one pass per digit: a per-tile digit histogram, a device-wide
// exclusive scan of the digit-major histogram table, then a scatter in which
// every key's rank among equal digits of its tile is computed from wave
// ballots (lanes with the same digit are matched bit by bit), so the order of
// equal keys is preserved and the result equals std::stable_sort on the keys.
// Tiles are 2048 keys (256 threads x 8 items), processed item by item in
// ascending key position.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "mpm_internal.h"

namespace mpm {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_RADIX = 256;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_ITEMS;

// Lanes of this wave whose digit equals `d`, among lanes with `valid` set.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                        uint32_t ntiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[RS_RADIX];
  const uint32_t tid = threadIdx.x, tile = blockIdx.x;
  const uint32_t lane = tid & 63u;
  cnt[tid] = 0;
  __syncthreads();
  for (int item = 0; item < RS_ITEMS; ++item) {
    const uint32_t idx = tile * RS_TILE + item * RS_THREADS + tid;
    const bool valid = idx < n;
    const uint32_t d = valid ? (keys[idx] >> shift) & 255u : 0u;
    const uint64_t peers = match_digit(d, valid);
    // The highest lane of each digit group adds the group's size.
    if (valid && (63u - (uint32_t)__clzll(peers)) == lane) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[(size_t)tid * ntiles + tile] = cnt[tid];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out, uint32_t n,
                                                           uint32_t shift, uint32_t ntiles,
                                                           const uint32_t* __restrict__ offsets) {
  __shared__ uint32_t s_base[RS_RADIX];
  __shared__ uint32_t s_cnt[4][RS_RADIX];
  const uint32_t tid = threadIdx.x, tile = blockIdx.x;
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  s_base[tid] = offsets[(size_t)tid * ntiles + tile];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  for (int item = 0; item < RS_ITEMS; ++item) {
    __syncthreads();
#pragma unroll
    for (int w = 0; w < 4; ++w) s_cnt[w][tid] = 0;
    const uint32_t idx = tile * RS_TILE + item * RS_THREADS + tid;
    const bool valid = idx < n;
    const uint32_t k = valid ? keys[idx] : 0u;
    const uint32_t v = valid ? vals[idx] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    const uint64_t peers = match_digit(d, valid);
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const uint32_t size = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid && rank == size - 1u) s_cnt[wave][d] = size;
    __syncthreads();
    {
      const uint32_t c0 = s_cnt[0][tid], c1 = s_cnt[1][tid], c2 = s_cnt[2][tid], c3 = s_cnt[3][tid];
      const uint32_t b = s_base[tid];
      s_cnt[0][tid] = b;
      s_cnt[1][tid] = b + c0;
      s_cnt[2][tid] = b + c0 + c1;
      s_cnt[3][tid] = b + c0 + c1 + c2;
      s_base[tid] = b + c0 + c1 + c2 + c3;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = s_cnt[wave][d] + rank;
      keys_out[pos] = k;
      vals_out[pos] = v;
    }
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

// Exclusive scan across a block of SCAN_THREADS threads; returns the block
// total in `total`.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_w, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_scan(v);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t before = 0, sum = 0;
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    if (w < wave) before += s_w[w];
    sum += s_w[w];
  }
  __syncthreads();
  total = sum;
  return before + inc - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t total,
                                                              uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_w[SCAN_THREADS / 64];
  const uint32_t base = blockIdx.x * SCAN_BLOCK;
  uint32_t sum = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const uint32_t idx = base + i * SCAN_THREADS + threadIdx.x;
    if (idx < total) sum += in[idx];
  }
  uint32_t t;
  block_exclusive(sum, s_w, t);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Single-block exclusive scan of the per-block sums.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partials(uint32_t* __restrict__ partials, uint32_t np) {
  __shared__ uint32_t s_w[SCAN_THREADS / 64];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < np; base += SCAN_THREADS) {
    const uint32_t idx = base + threadIdx.x;
    const uint32_t v = idx < np ? partials[idx] : 0u;
    uint32_t t;
    const uint32_t ex = block_exclusive(v, s_w, t);
    if (idx < np) partials[idx] = carry + ex;
    carry += t;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(uint32_t* __restrict__ data, uint32_t total,
                                                            const uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_data[SCAN_BLOCK];
  __shared__ uint32_t s_w[SCAN_THREADS / 64];
  const uint32_t base = blockIdx.x * SCAN_BLOCK;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const uint32_t idx = base + i * SCAN_THREADS + threadIdx.x;
    s_data[i * SCAN_THREADS + threadIdx.x] = idx < total ? data[idx] : 0u;
  }
  __syncthreads();
  uint32_t local[SCAN_ITEMS];
  uint32_t sum = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    local[i] = s_data[threadIdx.x * SCAN_ITEMS + i];
    sum += local[i];
  }
  uint32_t t;
  uint32_t run = block_exclusive(sum, s_w, t) + partials[blockIdx.x];
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    s_data[threadIdx.x * SCAN_ITEMS + i] = run;
    run += local[i];
  }
  __syncthreads();
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    const uint32_t idx = base + i * SCAN_THREADS + threadIdx.x;
    if (idx < total) data[idx] = s_data[i * SCAN_THREADS + threadIdx.x];
  }
}

}  // namespace

size_t radix_hist_words(uint32_t n) {
  const size_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  return (ntiles > 0 ? ntiles : 1) * RS_RADIX;
}

size_t radix_partial_words(uint32_t n) {
  return (radix_hist_words(n) + SCAN_BLOCK - 1) / SCAN_BLOCK + 1;
}

void radix_sort_pairs(hipStream_t s, uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt,
                      uint32_t* hist, uint32_t* partials, uint32_t n, uint32_t key_bits) {
  if (n == 0) return;
  const uint32_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  const uint32_t total = ntiles * RS_RADIX;
  const uint32_t nblocks = (total + SCAN_BLOCK - 1) / SCAN_BLOCK;
  for (uint32_t shift = 0; shift < key_bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, keys, n, shift, ntiles, hist);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nblocks), dim3(SCAN_THREADS), 0, s, hist, total, partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_THREADS), 0, s, partials, nblocks);
    hipLaunchKernelGGL(k_scan_down, dim3(nblocks), dim3(SCAN_THREADS), 0, s, hist, total, partials);
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, keys, vals, keys_alt, vals_alt, n, shift,
                       ntiles, hist);
    uint32_t* t = keys;
    keys = keys_alt;
    keys_alt = t;
    t = vals;
    vals = vals_alt;
    vals_alt = t;
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
constexpr size_t NA_PARTIALS = (NA_HIST + SCAN_BLOCK - 1) / SCAN_BLOCK + 1;
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
