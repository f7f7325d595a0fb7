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
constexpr int kMaxRedBlocks = 262144;
constexpr int kMaxRedVals = 8;
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

// The ghost cells of a box, numbered 0 .. ghost_cells(d) - 1: the two d3
// planes (all i, j), the two d2 planes (all i, k in 1..n3), the two d1 lines
// (j, k interior).  Below 2^31 for every box the callers accept (checked on
// the host), so 32-bit arithmetic.
__host__ __device__ inline int64_t ghost_cells(const Dims& d) {
  return 2 * d.g12() + 2 * (int64_t)(d.n1 + 2) * d.n3 + 2 * (int64_t)d.n2 * d.n3;
}
__device__ inline void ghost_cell(const Dims& d, int q, int& i, int& j, int& k) {
  const int g1 = d.n1 + 2, g12 = g1 * (d.n2 + 2);
  const int A = 2 * g12, B = 2 * g1 * d.n3;
  if (q < A) {
    const int r = q < g12 ? q : q - g12;
    k = q < g12 ? 0 : d.n3 + 1;
    i = r % g1;
    j = r / g1;
  } else if (q < A + B) {
    const int r0 = q - A, half = B / 2, r = r0 < half ? r0 : r0 - half;
    j = r0 < half ? 0 : d.n2 + 1;
    i = r % g1;
    k = r / g1 + 1;
  } else {
    const int r0 = q - A - B, half = d.n2 * d.n3, r = r0 < half ? r0 : r0 - half;
    i = r0 < half ? 0 : d.n1 + 1;
    j = r % d.n2 + 1;
    k = r / d.n2 + 1;
  }
}

// Column-marching launch: a block of TX x TY threads owns TX x TY columns of
// (d1,d2) and marches KC planes of d3.  256 consecutive d1 voxels per block
// row give fully coalesced 1 KiB (FP32) / 2 KiB (FP64) row loads.  The grid is
// one-dimensional; blocks are renumbered so that each of the 8 XCDs (which
// receive blocks round-robin) owns a contiguous run of tiles, ordered d2
// fastest, and neighbouring rows' halo loads hit the same L2.
constexpr int TX = 256, TY = 1, KC = 16;
struct TileGrid {
  int gx, gy, gz;
};
__host__ __device__ inline TileGrid tile_grid(const Dims& d, int kc = KC) {
  return TileGrid{(d.n1 + TX - 1) / TX, (d.n2 + TY - 1) / TY, (d.n3 + kc - 1) / kc};
}
inline dim3 col_grid(const Dims& d, int kc = KC) {
  const TileGrid t = tile_grid(d, kc);
  return dim3((unsigned)t.gx * t.gy * t.gz);
}
inline int blocks_of(const dim3& g) { return (int)(g.x * g.y * g.z); }

struct ColIdx {
  int i, j, k0, k1;
  bool active;
};
__device__ inline ColIdx col_idx(const Dims& d, int kc = KC) {
  const TileGrid t = tile_grid(d, kc);
  const int nb = t.gx * t.gy * t.gz, per = nb / 8;
  int b = blockIdx.x;
  if (b < per * 8) b = (b % 8) * per + b / 8;
  const int tj = b % t.gy, ti = (b / t.gy) % t.gx, tk = b / (t.gx * t.gy);
  ColIdx c;
  c.i = ti * TX + threadIdx.x + 1;
  c.j = tj * TY + threadIdx.y + 1;
  c.k0 = tk * kc + 1;
  c.k1 = min(c.k0 + kc - 1, d.n3);
  c.active = c.i <= d.n1 && c.j <= d.n2;
  return c;
}

// Stencil launch: a block of 256 threads owns a tile of TW d1 voxels x
// (256 / TW) x SJ d2 rows x SKC d3 planes (TW = 256 unless the box is narrower
// in d1, see stencil_tw).  Each thread keeps its SJ rows in registers
// and marches the planes, so d2 and d3 neighbours come from registers and d1
// neighbours from the adjacent lanes (DPP); only one row above and below the
// tile and the wave-edge lanes go to memory: ~3 loads per voxel instead of 14.
// Tiles are renumbered XCD-contiguously as in col_idx.
constexpr int SKC = 16;
template <typename T>
struct Rows;
template <>
struct Rows<float> {
  static constexpr int SJ = 4;
};
template <>
struct Rows<double> {
  static constexpr int SJ = 2;
};
// Width of a tile in d1: the block's 256 threads cover TW d1 voxels times
// 256 / TW strips of SJ rows, so that boxes narrower than 256 in d1 leave
// fewer lanes idle (operator at 93 x 730 x 160: 68 us with 256-wide tiles,
// 48.5 us with 128-wide; at 186 and 256 the 256-wide tile is the fastest).
__host__ __device__ inline int stencil_tw(const Dims& d) { return d.n1 <= 64 ? 64 : (d.n1 <= 128 ? 128 : 256); }
__host__ __device__ inline TileGrid stencil_tiles(const Dims& d, int sj) {
  const int tw = stencil_tw(d), g = TX / tw;
  return TileGrid{(d.n1 + tw - 1) / tw, (d.n2 + sj * g - 1) / (sj * g), (d.n3 + SKC - 1) / SKC};
}
template <typename T>
__host__ __device__ inline int stencil_ntiles(const Dims& d) {
  const TileGrid t = stencil_tiles(d, Rows<T>::SJ);
  return t.gx * t.gy * t.gz;
}
// Tile coordinates of block `tile` after the XCD-contiguous renumbering.
struct TileXyz {
  int ti, tj, tk;
};
__host__ __device__ inline TileXyz stencil_tile(const Dims& d, int sj, int tile) {
  const TileGrid t = stencil_tiles(d, sj);
  const int nb = t.gx * t.gy * t.gz, per = nb / 8;
  int b = tile;
  if (b < per * 8) b = (b % 8) * per + b / 8;
  return TileXyz{(b / t.gy) % t.gx, b % t.gy, b / (t.gx * t.gy)};
}
struct StIdx {
  int i, ic, j0, k0, k1, lane;  // ic = min(i, n1 + 1): lanes past the edge read the ghost column
};
// This thread's column and first row in tile `tile` (blocks of TX threads).
// Rows of a strip past n2 are clamped by the callers; such strips store nothing.
template <int SJ>
__device__ inline StIdx stencil_idx(const Dims& d, int tile) {
  const TileXyz c = stencil_tile(d, SJ, tile);
  const int tw = stencil_tw(d), g = TX / tw;
  const int x = (int)threadIdx.x % tw, strip = (int)threadIdx.x / tw;
  StIdx s;
  s.i = c.ti * tw + x + 1;
  s.ic = min(s.i, d.n1 + 1);
  s.j0 = (c.tj * g + strip) * SJ + 1;
  s.k0 = c.tk * SKC + 1;
  s.k1 = min(s.k0 + SKC - 1, d.n3);
  s.lane = threadIdx.x & 63;
  return s;
}

// Lane l receives lane l-1's (FROM_LEFT) or lane l+1's value across the whole
// wave (DPP wave_shr:1 / wave_shl:1); the lane with no source keeps its own.
template <bool FROM_LEFT>
__device__ __forceinline__ float lane_shift(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), FROM_LEFT ? 0x138 : 0x130, 0xf, 0xf, false));
}
template <bool FROM_LEFT>
__device__ __forceinline__ double lane_shift(double v) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(lo, lo, FROM_LEFT ? 0x138 : 0x130, 0xf, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, FROM_LEFT ? 0x138 : 0x130, 0xf, 0xf, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// BF16 storage (upper half of an IEEE single, round to nearest even).  The
// mixed solver keeps its Jacobi inverse in BF16: a preconditioner only has to
// be a fixed symmetric definite approximation of A^-1, and BF16 keeps the
// exponent range of the inverse diagonal (it spans ~1e-1 .. 1e7 here).
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  union { uint32_t u; float f; } c;
  c.u = (uint32_t)h << 16;
  return c.f;
}
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  union { uint32_t u; float f; } c;
  c.f = f;
  return (uint16_t)((c.u + 0x7fffu + ((c.u >> 16) & 1u)) >> 16);
}

// Flat launch over n items: each block owns one contiguous tile of
// FLAT_T * FLAT_U items (one trip, no grid-stride loop: on MI355X this keeps
// the most bytes in flight, ~6.3 TB/s for a copy against ~4.5 TB/s with a
// capped grid-stride loop).
constexpr int FLAT_T = 256, FLAT_U = 4;
inline int flat_blocks(int64_t n) {
  int64_t b = (n + (int64_t)FLAT_T * FLAT_U - 1) / ((int64_t)FLAT_T * FLAT_U);
  return (int)(b < 1 ? 1 : b);
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

// Store this block's partial sums (plain stores: the kernel boundary makes
// them visible to the finalising kernel; no fences or atomics, which on a
// multi-XCD device would force L2 write-backs in every block).
template <int NV>
__device__ inline void block_partials(double (&v)[NV], double* partials) {
  block_sum<NV>(v);
  if (linear_tid() == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) partials[q * kMaxRedBlocks + linear_bid()] = v[q];
}

// One-block finalisation: the sum of nb partials in a fixed order (the result
// does not depend on scheduling); valid in thread 0.
template <int NV>
__device__ inline void sum_partials(const double* __restrict__ partials, int nb, double (&tot)[NV]) {
  const int tid = linear_tid(), nt = nthreads();
  constexpr int U = 8;  // loads in flight per thread and value
  double s[NV][U];
#pragma unroll
  for (int q = 0; q < NV; q++)
#pragma unroll
    for (int u = 0; u < U; u++) s[q][u] = 0.0;
  for (int b0 = tid; b0 < nb; b0 += U * nt) {
    double v[NV][U];
#pragma unroll
    for (int q = 0; q < NV; q++)
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int b = b0 + u * nt;
        v[q][u] = b < nb ? partials[(size_t)q * kMaxRedBlocks + b] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < NV; q++)
#pragma unroll
      for (int u = 0; u < U; u++) s[q][u] += v[q][u];
  }
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; q++)
    acc[q] = ((s[q][0] + s[q][1]) + (s[q][2] + s[q][3])) + ((s[q][4] + s[q][5]) + (s[q][6] + s[q][7]));
  block_sum<NV>(acc);
#pragma unroll
  for (int q = 0; q < NV; q++) tot[q] = acc[q];
}
constexpr int FIN_T = 1024;  // threads of a finalising kernel

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
