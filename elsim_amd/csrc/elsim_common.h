// elsim_common.h -- device/host infrastructure shared by the HIP translation
// units of libelsim_hip.so: error state, the library stream, ghosted-grid
// indexing, launch shapes and deterministic block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/elsim.h"

namespace elsim {

// ---------------------------------------------------------------- errors --
void set_error(const char* fmt, ...);
hipStream_t stream();
int ensure_init();

#define ELSIM_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      elsim::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return ELSIM_ERR_HIP;                                                          \
    }                                                                                \
  } while (0)

#define ELSIM_CHECK_LAUNCH()                                                         \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      elsim::set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return ELSIM_ERR_HIP;                                                          \
    }                                                                                \
  } while (0)

#define ELSIM_TRY(expr)      \
  do {                       \
    int r_ = (expr);         \
    if (r_ < 0) return r_;   \
  } while (0)

// Device scratch owned by the library: reduction partials and counters.
// Kernels of one stream run in order, so one partials buffer is shared.
constexpr int kMaxRedBlocks = 65536;
constexpr int kMaxRedVals = 4;
struct Workspace {
  double* partials = nullptr;   // kMaxRedVals * kMaxRedBlocks
  unsigned* counter = nullptr;  // 1
};
Workspace& workspace();

// ------------------------------------------------------------ grid shape --
struct Dims {
  int n1, n2, n3;
  __host__ __device__ int64_t g1() const { return n1 + 2; }
  __host__ __device__ int64_t g12() const { return (int64_t)(n1 + 2) * (n2 + 2); }
  __host__ __device__ int64_t G() const { return (int64_t)(n1 + 2) * (n2 + 2) * (n3 + 2); }
  __host__ __device__ int64_t N() const { return (int64_t)n1 * n2 * n3; }
  // ghosted index, i,j,k in 0..n+1
  __host__ __device__ int64_t gi(int i, int j, int k) const { return i + g1() * (j + (int64_t)(n2 + 2) * k); }
  // interior index, i,j,k in 1..n
  __host__ __device__ int64_t ni(int i, int j, int k) const {
    return (i - 1) + (int64_t)n1 * ((j - 1) + (int64_t)n2 * (k - 1));
  }
};

inline bool dims_ok(int n1, int n2, int n3) {
  return n1 > 0 && n2 > 0 && n3 > 0 && (int64_t)(n1 + 2) * (n2 + 2) * (n3 + 2) < (int64_t)1 << 40 &&
         (int64_t)n1 * n2 * n3 < ((int64_t)1 << 31);
}

// Column-marching launch: a block covers TX x TY columns of (d1,d2) and a
// chunk of KC planes of d3.
constexpr int TX = 64, TY = 4, KC = 16;
inline dim3 col_grid(const Dims& d, int kc = KC) {
  return dim3((d.n1 + TX - 1) / TX, (d.n2 + TY - 1) / TY, (d.n3 + kc - 1) / kc);
}
inline int blocks_of(const dim3& g) { return (int)(g.x * g.y * g.z); }

// Flat launch over n items.
constexpr int FLAT_T = 256;
inline int flat_blocks(int64_t n, int cap = 8192) {
  int64_t b = (n + FLAT_T - 1) / FLAT_T;
  return (int)(b < cap ? (b < 1 ? 1 : b) : cap);
}

// ------------------------------------------------------------ reductions --
__device__ inline int linear_tid() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }
__device__ inline int linear_bid() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }
__device__ inline int nblocks() { return gridDim.x * gridDim.y * gridDim.z; }
__device__ inline int nthreads() { return blockDim.x * blockDim.y * blockDim.z; }

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of NV values in a fixed order; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double sm[NV][16];
  const int tid = linear_tid(), lane = tid & 63, wid = tid >> 6, nw = (nthreads() + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) sm[q][wid] = v[q];
  __syncthreads();
  if (tid == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double s = 0.0;
      for (int w = 0; w < nw; w++) s += sm[q][w];
      v[q] = s;
    }
  __syncthreads();
}

// Store this block's partial sums; the block that arrives last sums all
// partials in a fixed order (so the result does not depend on scheduling),
// leaves the totals in `tot` (valid in thread 0) and returns true in every
// thread of that block.
template <int NV>
__device__ inline bool reduce_last_block(double (&v)[NV], double* partials, unsigned* counter, double (&tot)[NV]) {
  __shared__ bool last;
  const int tid = linear_tid(), bid = linear_bid(), nb = nblocks();
  block_sum<NV>(v);
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) partials[q * kMaxRedBlocks + bid] = v[q];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    last = (prev == (unsigned)nb - 1u);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = 0.0;
    for (int b = tid; b < nb; b += nthreads())
      s += __hip_atomic_load(&partials[q * kMaxRedBlocks + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc[q] = s;
  }
  block_sum<NV>(acc);
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) tot[q] = acc[q];
    atomicExch(counter, 0u);
  }
  return true;
}

// -------------------------------------------------------- device buffers --
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
int dev_alloc(DevBuf& b, size_t bytes);
void dev_free(DevBuf& b);

// Library-owned scratch reused by host-array entry points (grows, never shrinks).
void* scratch(int slot, size_t bytes);

}  // namespace elsim
