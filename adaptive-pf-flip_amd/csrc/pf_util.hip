// pf_util.hip -- fills, a device-wide exclusive scan and flag compaction.
//
// The scan is the classic three-phase form sized for 64-wide wavefronts: each
// 256-thread workgroup owns a 4096-element tile (16 per thread, loaded coalesced
// through LDS), tile totals are scanned by one workgroup in a carried loop, and a
// down-sweep writes the exclusive prefix.  It is deterministic (no atomics).
#include "pf_internal.h"

namespace pf {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr int kScanTile = kScanThreads * kScanPerThread;

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename T>
hipError_t fill(hipStream_t s, T* p, T v, long long n) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// inclusive scan of one value per thread across a 256-thread workgroup
__device__ inline uint32_t block_inclusive_scan(uint32_t v, uint32_t* s_wave) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  if (lane == 63) s_wave[wave] = v;
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wave; ++w) off += s_wave[w];
  __syncthreads();
  return v + off;
}

template <typename In>
__device__ inline uint32_t load_in(const In* p, long long i, long long n) {
  return i < n ? (uint32_t)(p[i] != 0 ? (sizeof(In) == 1 ? 1u : (uint32_t)p[i]) : 0u) : 0u;
}

template <typename In>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const In* __restrict__ in, long long n,
                                                              uint32_t* __restrict__ tiles) {
  __shared__ uint32_t s_wave[4];
  long long base = (long long)blockIdx.x * kScanTile;
  uint32_t acc = 0;
#pragma unroll
  for (int q = 0; q < kScanPerThread; ++q) acc += load_in(in, base + q * kScanThreads + threadIdx.x, n);
  uint32_t tot = block_inclusive_scan(acc, s_wave);
  if (threadIdx.x == kScanThreads - 1) tiles[blockIdx.x] = tot;
}

// one workgroup: exclusive scan of tile totals in place; writes the grand total
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(uint32_t* __restrict__ tiles, long long nt,
                                                             uint32_t* __restrict__ total_out) {
  __shared__ uint32_t s_wave[4];
  uint32_t carry = 0;
  for (long long base = 0; base < nt; base += kScanThreads) {
    long long i = base + threadIdx.x;
    uint32_t v = i < nt ? tiles[i] : 0u;
    uint32_t inc = block_inclusive_scan(v, s_wave);
    if (i < nt) tiles[i] = carry + inc - v;
    __shared__ uint32_t s_last;
    if (threadIdx.x == kScanThreads - 1) s_last = inc;
    __syncthreads();
    carry += s_last;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total_out = carry;
}

template <typename In>
__global__ void __launch_bounds__(kScanThreads) k_scan_down(const In* __restrict__ in, long long n,
                                                            const uint32_t* __restrict__ tiles,
                                                            uint32_t* __restrict__ out) {
  __shared__ uint32_t s_data[kScanTile];
  __shared__ uint32_t s_wave[4];
  long long base = (long long)blockIdx.x * kScanTile;
#pragma unroll
  for (int q = 0; q < kScanPerThread; ++q) {
    int li = q * kScanThreads + threadIdx.x;
    s_data[li] = load_in(in, base + li, n);
  }
  __syncthreads();
  uint32_t loc[kScanPerThread];
  uint32_t acc = 0;
#pragma unroll
  for (int q = 0; q < kScanPerThread; ++q) {
    loc[q] = acc;
    acc += s_data[threadIdx.x * kScanPerThread + q];
  }
  uint32_t inc = block_inclusive_scan(acc, s_wave);
  uint32_t off = tiles[blockIdx.x] + inc - acc;
#pragma unroll
  for (int q = 0; q < kScanPerThread; ++q) s_data[threadIdx.x * kScanPerThread + q] = off + loc[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kScanPerThread; ++q) {
    int li = q * kScanThreads + threadIdx.x;
    if (base + li < n) out[base + li] = s_data[li];
  }
}

template <typename In>
hipError_t scan_impl(Ctx& c, const In* in, uint32_t* out, long long n) {
  long long nt = (n + kScanTile - 1) / kScanTile;
  if (nt < 1) nt = 1;
  if (nt > c.scan_tile_cap) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scan_reduce<In>, dim3((unsigned)nt), dim3(kScanThreads), 0, c.stream, in, n,
                     c.scan_tiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, c.stream, c.scan_tiles, nt, out + n);
  hipLaunchKernelGGL(k_scan_down<In>, dim3((unsigned)nt), dim3(kScanThreads), 0, c.stream, in, n,
                     c.scan_tiles, out);
  return hipGetLastError();
}

__global__ void k_compact(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ pos, long long n,
                          uint32_t* __restrict__ list, uint32_t* __restrict__ count_dev) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) list[pos[i]] = (uint32_t)i;
  if (i == 0) *count_dev = pos[n];
}

}  // namespace

// ---- Morton-ordered block lists.  A contiguous, aligned run of 2^k Morton codes
// is a compact box of blocks, so kernels that walk such a list keep every stencil
// neighbour of the blocks in flight close in time (and in one XCD's L2).

// every third bit of m, compacted (inverse of the 3D bit interleave)
__device__ inline uint32_t morton_part(uint32_t m) {
  m &= 0x09249249u;
  m = (m ^ (m >> 2)) & 0x030c30c3u;
  m = (m ^ (m >> 4)) & 0x0300f00fu;
  m = (m ^ (m >> 8)) & 0xff0000ffu;
  m = (m ^ (m >> 16)) & 0x000003ffu;
  return m;
}

__global__ void k_morton_flags(Dims d, const uint8_t* __restrict__ flag_by_bid, long long ncodes,
                               uint8_t* __restrict__ flag_by_code) {
  long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ncodes) return;
  const uint32_t bx = morton_part((uint32_t)m), by = morton_part((uint32_t)m >> 1), bz = morton_part((uint32_t)m >> 2);
  uint8_t f = 0;
  if ((int)bx < d.nbx && (int)by < d.nby && (int)bz < d.nbz) f = flag_by_bid[bid_of(d, bx, by, bz)] ? 1 : 0;
  flag_by_code[m] = f;
}

__global__ void k_compact_morton(Dims d, const uint8_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                                 long long n, uint32_t* __restrict__ list, uint32_t* __restrict__ count_dev) {
  long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n && flags[m]) {
    const uint32_t bx = morton_part((uint32_t)m), by = morton_part((uint32_t)m >> 1), bz = morton_part((uint32_t)m >> 2);
    list[pos[m]] = (uint32_t)bid_of(d, bx, by, bz);
  }
  if (m == 0) *count_dev = pos[n];
}

long long morton_codes(const Dims& d) {
  int e = 1;
  while (e < d.nbx || e < d.nby || e < d.nbz) e <<= 1;
  return (long long)e * e * e;
}

hipError_t compact_flags_morton(Ctx& c, const Dims& d, const uint8_t* flags, uint32_t* list, uint32_t* count_dev) {
  const long long n = morton_codes(d);
  if (n + 1 > c.morton_cap) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_morton_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, d, flags, n,
                     c.morton_flags);
  hipError_t e = scan_impl<uint8_t>(c, c.morton_flags, c.blk_scan, n);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_compact_morton, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, d, c.morton_flags,
                     c.blk_scan, n, list, count_dev);
  return hipGetLastError();
}

// plain 16-byte-per-lane copy: the known-size kernel for bandwidth calibration
__global__ void k_copy_f32x4(const float4* __restrict__ a, float4* __restrict__ b, long long n4) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n4; i += stride) b[i] = a[i];
}

hipError_t fill_u8(hipStream_t s, uint8_t* p, uint8_t v, long long n) { return fill(s, p, v, n); }
hipError_t fill_u32(hipStream_t s, uint32_t* p, uint32_t v, long long n) { return fill(s, p, v, n); }
hipError_t fill_f32(hipStream_t s, float* p, float v, long long n) { return fill(s, p, v, n); }
hipError_t fill_f64(hipStream_t s, double* p, double v, long long n) { return fill(s, p, v, n); }

hipError_t scan_u32(Ctx& c, const uint32_t* in, uint32_t* out, long long n) {
  return scan_impl<uint32_t>(c, in, out, n);
}

hipError_t compact_flags(Ctx& c, const uint8_t* flags, long long n, uint32_t* list, uint32_t* count_dev) {
  hipError_t e = scan_impl<uint8_t>(c, flags, c.blk_scan, n);
  if (e != hipSuccess) return e;
  long long blocks = (n + 255) / 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_compact, dim3((unsigned)blocks), dim3(256), 0, c.stream, flags, c.blk_scan, n, list,
                     count_dev);
  return hipGetLastError();
}

}  // namespace pf
