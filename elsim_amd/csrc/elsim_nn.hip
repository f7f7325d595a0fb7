// elsim_nn.hip -- the learned CG warm start: a 4-level 3-D convolutional
// encoder-decoder (channels 16/32/64/64, 3x3x3 "same" convolutions, 2x2x2
// average pooling, nearest upsampling, concatenated skips, ReLU, 1x1x1
// output), trained and run on the GPU by the kernels of this file.  It has
// no counterpart in the reference.
//
// Storage: activations, convolution weights and activation gradients in
// bfloat16, channels-last ([voxel][channel], voxel = i + n1 (j + n2 k)), all
// accumulation in FP32 on the MFMA units (v_mfma_f32_16x16x16_bf16).
// Master weights, weight gradients and Adam moments are FP32.
//
// Convolution as implicit GEMM: D[voxel, co] = sum_K A[voxel, K] B[K, co]
// with K = tap * cin + ci.  A block owns a 16 x 4 x MT voxel tile; the tile
// plus its one-voxel halo is staged in LDS 16 or 32 channels at a time, so
// each input value is read from HBM/L2 once per block and 27 times from LDS.
// Every wave owns MT 16-voxel rows of the tile and all output channels.
// The backward-data pass is the same kernel with the taps flipped and the
// channel roles exchanged (stride 1, zero padding), writing FP32.
//
// Channel counts are multiples of 16 (the input features are padded to 16).
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "elsim_common.h"
#include "elsim_nn.h"
#include "elsim_solver.h"

namespace elsim {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

// v_mfma_f32_16x16x32_bf16 (gfx950): lane l holds A[l % 16][8 (l / 16) .. + 7]
// and B[8 (l / 16) .. + 7][l % 16]; D[4 (l / 16) + i][l % 16] in acc[i] (the
// layout of the 16x16x16 form with eight k per lane; checked on the GPU
// against a CPU product before use).
__device__ __forceinline__ f4v mfma32(uint4 a, uint4 b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}

// ----------------------------------------------------------- convolution --

// compute units of the current device (persistent grids)
static int cu_count() {
  static int cus[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus[dev] = 256;
  return cus[dev] > 0 ? cus[dev] : 256;
}

constexpr int CONV_TX = 16, CONV_TY = 4;  // voxel tile in d1, d2 (d3: MT)
constexpr int LDS_PAD = 8;                // bf16 elements of padding per staged voxel
constexpr int CS_MAX = 32;                // channels staged per slab

struct View {
  int cs, co;  // channel stride and offset of a channel slice
};

// Channels [0, ch) of a convolution's input taken from a tensor at half
// resolution, nearest-upsampled while they are staged (the decoder's
// upsampled skip half, read where it lies instead of from a full-resolution
// copy); ch = 0: every channel from the full-resolution input.
struct Coarse {
  const uint16_t* p;
  int ch;
  View v;
};

// Stage one slab of the input tile plus its one-voxel halo (HN voxels, nq
// 16-byte chunks each, nq <= NQ) into LDS rows of rs elements, zero outside
// the box.  Each thread issues all of its global loads before its first LDS
// store, so they are in flight together (one load latency per slab instead
// of one per chunk).
template <int HN, int NQ>
__device__ __forceinline__ void stage_slab(int nq, int rs, int x0, int y0, int z0, int n1, int n2, int n3,
                                           const uint16_t* __restrict__ in, View iv, Coarse lo, int slab,
                                           uint16_t* __restrict__ lds) {
  constexpr int HX = CONV_TX + 2, HY = CONV_TY + 2;
  constexpr int NU = (HN * NQ + 255) / 256;
  const int tid = threadIdx.x, tot = HN * nq;
  uint4 v[NU];
#pragma unroll
  for (int u = 0; u < NU; u++) {
    const int it = tid + 256 * u;
    v[u] = make_uint4(0, 0, 0, 0);
    if (it < tot) {
      const int h = it / nq, q = it % nq;
      const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
      const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
      if (gx >= 0 && gx < n1 && gy >= 0 && gy < n2 && gz >= 0 && gz < n3) {
        if (slab < lo.ch) {
          const int64_t cv = (gx >> 1) + (int64_t)(n1 >> 1) * ((gy >> 1) + (int64_t)(n2 >> 1) * (gz >> 1));
          v[u] = *(const uint4*)(lo.p + cv * lo.v.cs + lo.v.co + slab + 8 * q);
        } else {
          const int64_t vox = gx + (int64_t)n1 * (gy + (int64_t)n2 * gz);
          v[u] = *(const uint4*)(in + vox * iv.cs + iv.co + slab + 8 * q);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NU; u++) {
    const int it = tid + 256 * u;
    if (it < tot) *(uint4*)(lds + (it / nq) * rs + 8 * (it % nq)) = v[u];
  }
}

// Output of a convolution tile.  D[row = 4 lq + i][col = lr] of a wave's
// MFMA tile is voxel x = 4 lq + i of its row, channel t*16 + lr.  MODE 0
// (bf16, bias, ReLU) goes through LDS when the tile fits the block's staging
// buffer (LDS_N elements): written as [voxel][channel], then stored as
// 16-byte rows of 8 channels instead of 2-byte scattered stores (E0 at 256^3:
// 1003 -> 844 us).  Every wave must have finished reading the staged input.
// pout (MODE 0 through LDS only): also the 2x2x2 average pooling of the tile
// (compact rows of COUT channels at half resolution), summed from the stored
// bf16 values in k_pool's order, so the pooled tensor is the same.
template <int COUT, int MT, int MODE, int LDS_N>
__device__ __forceinline__ void conv_epilogue(const f4v (&acc)[MT][COUT / 16], int x0, int y0, int z0, int n1,
                                              int n2, int n3, const float* __restrict__ bias, int relu,
                                              void* __restrict__ out, View ov, uint16_t* __restrict__ lds,
                                              uint16_t* __restrict__ pout) {
  constexpr int NT = COUT / 16;
  constexpr int OS = COUT + 8;  // LDS row stride of the output tile (elements)
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lq = lane >> 4;
  if constexpr (MODE == 0 && CONV_TX * CONV_TY * MT * OS <= LDS_N) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int row = wave * MT + m;
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const int co = t * 16 + lr;
          float v = acc[m][t][i] + bias[co];
          if (relu) v = fmaxf(v, 0.f);
          lds[(row * CONV_TX + 4 * lq + i) * OS + co] = f2bf(v);
        }
    }
    __syncthreads();
    constexpr int CH = COUT / 8;  // 16-byte chunks per voxel
    for (int e = tid; e < CONV_TX * CONV_TY * MT * CH; e += 256) {
      const int vt = e / CH, c8 = e % CH;
      const int x = x0 + vt % CONV_TX, row = vt / CONV_TX, y = y0 + row % CONV_TY, z = z0 + row / CONV_TY;
      if (x >= n1 || y >= n2 || z >= n3) continue;
      const int64_t vox = x + (int64_t)n1 * (y + (int64_t)n2 * z);
      *(uint4*)((uint16_t*)out + vox * ov.cs + ov.co + 8 * c8) = *(const uint4*)(lds + vt * OS + 8 * c8);
    }
    static_assert(CONV_TX % 2 == 0 && CONV_TY % 2 == 0, "pooling of whole 2x2x2 blocks");
    if (MT % 2 == 0 && pout) {  // the host launches an odd MT without pooling only
      constexpr int PX = CONV_TX / 2, PY = CONV_TY / 2, PN = PX * PY * (MT / 2 > 0 ? MT / 2 : 1);
      const int m1 = n1 / 2, m2 = n2 / 2, m3 = n3 / 2;
      for (int e = tid; e < PN * CH; e += 256) {
        const int pv = e / CH, c8 = e % CH;
        const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
        const int cx = x0 / 2 + px, cy = y0 / 2 + py, cz = z0 / 2 + pz;
        if (cx >= m1 || cy >= m2 || cz >= m3) continue;
        uint4 ch[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int vt = (2 * px + (q & 1)) + CONV_TX * ((2 * py + ((q >> 1) & 1)) + CONV_TY * (2 * pz + (q >> 2)));
          ch[q] = *(const uint4*)(lds + vt * OS + 8 * c8);
        }
        uint16_t r[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          float sum = 0.f;
#pragma unroll
          for (int q = 0; q < 8; q++) sum += bf2f(((const uint16_t*)&ch[q])[k]);
          r[k] = f2bf(sum * 0.125f);
        }
        *(uint4*)(pout + (cx + (int64_t)m1 * (cy + (int64_t)m2 * cz)) * COUT + 8 * c8) = *(const uint4*)r;
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int row = wave * MT + m, y = y0 + row % CONV_TY, z = z0 + row / CONV_TY;
      if (y >= n2 || z >= n3) continue;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int x = x0 + 4 * lq + i;
        if (x >= n1) continue;
        const int64_t vox = x + (int64_t)n1 * (y + (int64_t)n2 * z);
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const int co = t * 16 + lr;
          if (MODE == 0) {
            float v = acc[m][t][i] + bias[co];
            if (relu) v = fmaxf(v, 0.f);
            ((uint16_t*)out)[vox * ov.cs + ov.co + co] = f2bf(v);
          } else {
            ((float*)out)[vox * ov.cs + ov.co + co] = acc[m][t][i];
          }
        }
      }
    }
  }
}

// MODE 0: out bf16 = relu?(acc + bias); MODE 1: out FP32 = acc.
template <int COUT, int MT, int MODE>
__global__ __launch_bounds__(256) void k_conv3(int n1, int n2, int n3, const uint16_t* __restrict__ in, int cin,
                                               View iv, Coarse lo, const uint16_t* __restrict__ wt,
                                               const float* __restrict__ bias, int relu, void* __restrict__ out,
                                               View ov, uint16_t* __restrict__ pout) {
  constexpr int NT = COUT / 16;
  constexpr int HX = CONV_TX + 2, HY = CONV_TY + 2, HZ = MT + 2, HN = HX * HY * HZ;
  __shared__ __align__(16) uint16_t lds[HN * (CS_MAX + LDS_PAD)];
  const int tx = (n1 + CONV_TX - 1) / CONV_TX, ty = (n2 + CONV_TY - 1) / CONV_TY;
  const int b = blockIdx.x;
  const int x0 = (b % tx) * CONV_TX, y0 = ((b / tx) % ty) * CONV_TY, z0 = (b / (tx * ty)) * MT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lq = lane >> 4;
  const int cs = (cin % 32 == 0) ? 32 : 16;
  const int rs = cs + LDS_PAD;  // LDS row stride (elements)
  const int K = 27 * cin;
  f4v acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[m][t] = f4v{0.f, 0.f, 0.f, 0.f};

  for (int slab = 0; slab < cin; slab += cs) {
    __syncthreads();
    stage_slab<HN, CS_MAX / 8>(cs / 8, rs, x0, y0, z0, n1, n2, n3, in, iv, lo, slab, lds);
    __syncthreads();
    if (cs == 32) {
      // one 16x16x32 MFMA per output tile and tap: lane group lq holds
      // channels 8 lq .. 8 lq + 7 of the slab.  The next tap's weight
      // fragments (global) and input fragments (LDS) are loaded before this
      // tap's MFMAs are issued: with one or two waves per SIMD nothing else
      // would cover those latencies.
      auto lda = [&](int tap, uint4 (&a)[MT]) {
        const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
#pragma unroll
        for (int m = 0; m < MT; m++) {
          const int row = wave * MT + m, y = row % CONV_TY, z = row / CONV_TY;
          const int h = ((z + dz) * HY + (y + dy)) * HX + (lr + dx);
          a[m] = *(const uint4*)(lds + h * rs + 8 * lq);
        }
      };
      auto ldb = [&](int tap, uint4 (&bw)[NT]) {
        const int k = tap * cin + slab + 8 * lq;
#pragma unroll
        for (int t = 0; t < NT; t++) bw[t] = *(const uint4*)(wt + (int64_t)(t * 16 + lr) * K + k);
      };
      uint4 an[MT], bn[NT];
      lda(0, an);
      ldb(0, bn);
      for (int tap = 0; tap < 27; tap++) {
        uint4 a[MT], bw[NT];
#pragma unroll
        for (int m = 0; m < MT; m++) a[m] = an[m];
#pragma unroll
        for (int t = 0; t < NT; t++) bw[t] = bn[t];
        if (tap + 1 < 27) {
          lda(tap + 1, an);
          ldb(tap + 1, bn);
        }
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int m = 0; m < MT; m++) acc[m][t] = mfma32(a[m], bw[t], acc[m][t]);
      }
      continue;
    }
    // 16-channel slabs: one 16x16x16 MFMA per output tile and tap, the next
    // tap's fragments loaded before this tap's MFMAs (as above)
    auto lda16 = [&](int tap, s4v (&a)[MT]) {
      const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;  // halo offsets 0..2 (= -1..1 + 1)
#pragma unroll
      for (int m = 0; m < MT; m++) {
        const int row = wave * MT + m, y = row % CONV_TY, z = row / CONV_TY;
        const int h = ((z + dz) * HY + (y + dy)) * HX + (lr + dx);
        a[m] = *(const s4v*)(lds + h * rs + 4 * lq);
      }
    };
    auto ldb16 = [&](int tap, s4v (&bw)[NT]) {
      const int k = tap * cin + slab + 4 * lq;
#pragma unroll
      for (int t = 0; t < NT; t++) bw[t] = *(const s4v*)(wt + (int64_t)(t * 16 + lr) * K + k);
    };
    s4v an[MT], bn[NT];
    lda16(0, an);
    ldb16(0, bn);
    for (int tap = 0; tap < 27; tap++) {
      s4v a[MT], bw[NT];
#pragma unroll
      for (int m = 0; m < MT; m++) a[m] = an[m];
#pragma unroll
      for (int t = 0; t < NT; t++) bw[t] = bn[t];
      if (tap + 1 < 27) {
        lda16(tap + 1, an);
        ldb16(tap + 1, bn);
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int m = 0; m < MT; m++)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], bw[t], acc[m][t], 0, 0, 0);
    }
  }
  conv_epilogue<COUT, MT, MODE, HN*(CS_MAX + LDS_PAD)>(acc, x0, y0, z0, n1, n2, n3, bias, relu, out, ov, lds, pout);
}

// The same convolution for 16- or 8-channel slabs (CS; input channels not a
// multiple of 32, or a layer whose used input channels are 8 of its 16) and
// at most 32 output channels: each wave loads the weight fragments of all 27
// taps of a slab into registers at once (14 x NT loads in flight, issued
// before the slab is staged), instead of one load per tap whose latency the
// tap's few MFMAs cannot hide.  One 16x16x32 MFMA covers 32 / CS taps: lane
// group lq holds tap (32 / CS) pair + lq / (CS / 8), channels
// 8 (lq % (CS / 8)) .. + 7; the taps past the 27th have zero weights.  The
// slabs cover channels [0, cused) of cin (the weight rows' length); the LDS
// row stride is 24 elements in both cases, which keeps a 16-lane ds_read_b128
// free of bank conflicts.
template <int COUT, int MT, int MODE, int CS>
__global__ __launch_bounds__(256) void k_conv3_w27(int n1, int n2, int n3, const uint16_t* __restrict__ in, int cin,
                                                   int cused, View iv, Coarse lo, const uint16_t* __restrict__ wt,
                                                   const float* __restrict__ bias, int relu, void* __restrict__ out,
                                                   View ov, uint16_t* __restrict__ pout) {
  constexpr int NT = COUT / 16;
  static_assert(NT <= 2, "weight fragments of 27 taps for at most two output tiles");
  static_assert(CS == 16 || CS == 8, "slabs of 16 or 8 channels");
  constexpr int TPM = 32 / CS, NP = (27 + TPM - 1) / TPM;  // taps per MFMA, MFMAs per slab
  constexpr int HX = CONV_TX + 2, HY = CONV_TY + 2, HZ = MT + 2, HN = HX * HY * HZ;
  constexpr int RS = 24;
  __shared__ __align__(16) uint16_t lds[HN * RS];
  const int tx = (n1 + CONV_TX - 1) / CONV_TX, ty = (n2 + CONV_TY - 1) / CONV_TY;
  const int b = blockIdx.x;
  const int x0 = (b % tx) * CONV_TX, y0 = ((b / tx) % ty) * CONV_TY, z0 = (b / (tx * ty)) * MT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lq = lane >> 4;
  const int K = 27 * cin;
  f4v acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[m][t] = f4v{0.f, 0.f, 0.f, 0.f};

  const int tsub = lq / (CS / 8), csub = 8 * (lq % (CS / 8));
  for (int slab = 0; slab < cused; slab += CS) {
    uint4 bw[NP][NT];
#pragma unroll
    for (int pr = 0; pr < NP; pr++) {
      const int tap = TPM * pr + tsub;
#pragma unroll
      for (int t = 0; t < NT; t++)
        bw[pr][t] = tap < 27 ? *(const uint4*)(wt + (int64_t)(t * 16 + lr) * K + tap * cin + slab + csub)
                             : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    stage_slab<HN, CS / 8>(CS / 8, RS, x0, y0, z0, n1, n2, n3, in, iv, lo, slab, lds);
    __syncthreads();
#pragma unroll
    for (int pr = 0; pr < NP; pr++) {
      const int tap = min(TPM * pr + tsub, 26);  // taps past the 27th read tap 26's values against zero weights
      const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
      uint4 a[MT];
#pragma unroll
      for (int m = 0; m < MT; m++) {
        const int row = wave * MT + m, y = row % CONV_TY, z = row / CONV_TY;
        const int h = ((z + dz) * HY + (y + dy)) * HX + (lr + dx);
        a[m] = *(const uint4*)(lds + h * RS + csub);
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int m = 0; m < MT; m++) acc[m][t] = mfma32(a[m], bw[pr][t], acc[m][t]);
    }
  }
  conv_epilogue<COUT, MT, MODE, HN * RS>(acc, x0, y0, z0, n1, n2, n3, bias, relu, out, ov, lds, pout);
}

// The 8-channel form for 16 bf16 outputs (the first layer) as a persistent
// kernel: a block keeps the 7 weight fragments in registers for all of its
// tiles and loads the next tile's input while it computes the current one.
template <int MT>
__global__ __launch_bounds__(256) void k_conv3_p8(int n1, int n2, int n3, const uint16_t* __restrict__ in, int cin,
                                                  View iv, const uint16_t* __restrict__ wt,
                                                  const float* __restrict__ bias, int relu, void* __restrict__ out,
                                                  View ov, uint16_t* __restrict__ pout) {
  constexpr int HX = CONV_TX + 2, HY = CONV_TY + 2, HZ = MT + 2, HN = HX * HY * HZ;
  constexpr int RS = 24, NU = (HN + 255) / 256;
  __shared__ __align__(16) uint16_t lds[HN * RS];
  const int tx = (n1 + CONV_TX - 1) / CONV_TX, ty = (n2 + CONV_TY - 1) / CONV_TY, tz = (n3 + MT - 1) / MT;
  const int ntiles = tx * ty * tz;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lq = lane >> 4;
  uint4 bw[7];
#pragma unroll
  for (int pr = 0; pr < 7; pr++) {
    const int tap = 4 * pr + lq;
    bw[pr] = tap < 27 ? *(const uint4*)(wt + (int64_t)lr * 27 * cin + tap * cin) : make_uint4(0, 0, 0, 0);
  }
  uint4 v[NU];
  auto load = [&](int t) {
    const int x0 = (t % tx) * CONV_TX, y0 = ((t / tx) % ty) * CONV_TY, z0 = (t / (tx * ty)) * MT;
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int h = tid + 256 * u;
      const int gx = x0 - 1 + h % HX, gy = y0 - 1 + (h / HX) % HY, gz = z0 - 1 + h / (HX * HY);
      v[u] = make_uint4(0, 0, 0, 0);
      if (h < HN && gx >= 0 && gx < n1 && gy >= 0 && gy < n2 && gz >= 0 && gz < n3)
        v[u] = *(const uint4*)(in + (gx + (int64_t)n1 * (gy + (int64_t)n2 * gz)) * iv.cs + iv.co);
    }
  };
  int t = blockIdx.x;
  if (t < ntiles) load(t);
  for (; t < ntiles; t += gridDim.x) {
    const int x0 = (t % tx) * CONV_TX, y0 = ((t / tx) % ty) * CONV_TY, z0 = (t / (tx * ty)) * MT;
    __syncthreads();  // the previous tile's output is stored
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int h = tid + 256 * u;
      if (h < HN) *(uint4*)(lds + h * RS) = v[u];
    }
    if (t + (int)gridDim.x < ntiles) load(t + gridDim.x);
    __syncthreads();
    f4v acc[MT][1];
#pragma unroll
    for (int m = 0; m < MT; m++) acc[m][0] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int pr = 0; pr < 7; pr++) {
      const int tap = min(4 * pr + lq, 26);  // taps past the 27th read tap 26 against zero weights
      const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
#pragma unroll
      for (int m = 0; m < MT; m++) {
        const int row = wave * MT + m, y = row % CONV_TY, z = row / CONV_TY;
        const int h = ((z + dz) * HY + (y + dy)) * HX + (lr + dx);
        acc[m][0] = mfma32(*(const uint4*)(lds + h * RS), bw[pr], acc[m][0]);
      }
    }
    conv_epilogue<16, MT, 0, HN * RS>(acc, x0, y0, z0, n1, n2, n3, bias, relu, out, ov, lds, pout);
  }
}

template <int MODE, int COUT, int MT>
static void conv_launch(int n1, int n2, int n3, const uint16_t* in, int cin, int cused, View iv, Coarse lo,
                        const uint16_t* w, const float* bias, int relu, void* out, View ov, hipStream_t s,
                        uint16_t* pout) {
  const int nb = ((n1 + CONV_TX - 1) / CONV_TX) * ((n2 + CONV_TY - 1) / CONV_TY) * ((n3 + MT - 1) / MT);
  if constexpr (COUT == 16 && MODE == 0) {
    if (cused == 8 && lo.ch == 0) {
      const int nt = ((n1 + CONV_TX - 1) / CONV_TX) * ((n2 + CONV_TY - 1) / CONV_TY) * ((n3 + MT - 1) / MT);
      static int per_cu = 0;  // resident blocks per CU
      if (!per_cu && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3_p8<MT>, 256, 0) != hipSuccess ||
                      per_cu < 1))
        per_cu = 1;
      k_conv3_p8<MT><<<std::min(nt, per_cu * cu_count()), 256, 0, s>>>(n1, n2, n3, in, cin, iv, w, bias, relu, out,
                                                                      ov, pout);
      return;
    }
  }
  if constexpr (COUT <= 32) {
    if (cused % 16 != 0) {
      k_conv3_w27<COUT, MT, MODE, 8><<<nb, 256, 0, s>>>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, pout);
      return;
    }
    // 16-channel slabs with all 27 taps' weights in registers pay for 16
    // output channels only: for 32, the 112 registers of weights leave one
    // wave per SIMD (E1 at 128^3: 278 us against 213 us for k_conv3)
    if (cin % 32 != 0 && COUT == 16) {
      k_conv3_w27<COUT, MT, MODE, 16><<<nb, 256, 0, s>>>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, pout);
      return;
    }
  }
  k_conv3<COUT, MT, MODE><<<nb, 256, 0, s>>>(n1, n2, n3, in, cin, iv, lo, w, bias, relu, out, ov, pout);
}

// Input channels: a multiple of 16, or 8.  cused < cin: only input channels
// [0, cused) are read (the others are known to be zero, their weights
// contribute nothing).  8 used channels (cin 8, or 8 of 16) run the 8-channel
// slab kernel, which takes at most 32 output channels.  pout: also the
// output's 2x2x2 average pooling (MODE 0, at most 64 output channels, even
// grid), as k_pool computes it.
template <int MODE>
static int conv(int n1, int n2, int n3, const uint16_t* in, int cin, View iv, const uint16_t* w, const float* bias,
                int relu, int cout, void* out, View ov, hipStream_t s, Coarse lo = Coarse{nullptr, 0, View{0, 0}},
                int cused = 0, uint16_t* pout = nullptr) {
  if (pout && (MODE != 0 || cout > 64 || n1 % 2 || n2 % 2 || n3 % 2)) {
    set_error("convolution: pooled output needs bf16 output, at most 64 channels and an even grid");
    return ELSIM_ERR_ARG;
  }
  if ((cin % 16 && cin != 8) || cin <= 0 || iv.cs % 8 || iv.co % 8) {
    set_error("convolution: input channels %d (stride %d, offset %d) must be 8 or multiples of 16 (8)", cin, iv.cs,
              iv.co);
    return ELSIM_ERR_ARG;
  }
  if (cused == 0) cused = cin;
  if ((cused != cin && !(cused == 8 && cin == 16)) || (cused == 8 && (cout > 32 || lo.ch))) {
    set_error("convolution: %d of %d input channels: only 8 of 16, and 8 with at most 32 outputs", cused, cin);
    return ELSIM_ERR_ARG;
  }
  const int cs = cin % 32 == 0 ? 32 : 16;
  if (lo.ch && (lo.ch % cs || lo.ch > cin || lo.v.cs % 8 || lo.v.co % 8 || n1 % 2 || n2 % 2 || n3 % 2)) {
    set_error("convolution: upsampled channels %d do not fit slabs of %d", lo.ch, cs);
    return ELSIM_ERR_ARG;
  }
  switch (cout) {
    case 16: conv_launch<MODE, 16, 4>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, s, pout); break;
    case 32: conv_launch<MODE, 32, 4>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, s, pout); break;
    case 48: conv_launch<MODE, 48, 4>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, s, pout); break;
    case 64: conv_launch<MODE, 64, 4>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, s, pout); break;
    case 96: conv_launch<MODE, 96, 2>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, s, pout); break;
    case 128: conv_launch<MODE, 128, 2>(n1, n2, n3, in, cin, cused, iv, lo, w, bias, relu, out, ov, s, pout); break;
    default: set_error("convolution: %d output channels not supported", cout); return ELSIM_ERR_ARG;
  }
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

// ------------------------------------- last decoder layer at inference --
// The d0 convolution (48 -> 16 channels: 0..31 nearest-upsampled from d1,
// 32..47 the skip tensor) with its ReLU and the 1x1x1 output layer, for the
// inference forward.  The upsampled half is convolved where it lies, at d1's
// resolution: for a fine voxel of parity (px, py, pz) the 27 taps fall on
// 2 x 2 x 2 coarse voxels (along one axis, parity 0 sends tap 0 to coarse
// offset -1 and taps 1, 2 to 0; parity 1 sends taps 0, 1 to 0 and tap 2 to
// +1), so its 27 weight matrices fold into 8 per parity (wf, summed in FP32
// from the bf16 weights and rounded to bf16): 8 MFMAs of 32 channels in
// place of 27.  Each of the 8 waves of a block owns one parity class: its
// MFMA rows are 16 fine voxels two apart in x, and the staged skip tile is
// split by x parity so that these rows read consecutive LDS rows.  Blocks are
// persistent (one per CU): both weight sets stay in registers, and the next
// tile's loads are in flight while the current one is computed.  The output
// layer reads the bf16-rounded ReLU output as the unfused path does.
constexpr int D0_CC = 32, D0_SC = 16, D0_CIN = D0_CC + D0_SC;
constexpr int D0_TX = 16, D0_TY = 2, D0_TZ = 2;  // coarse tile (fine: 32 x 4 x 4)
constexpr int D0_CHX = D0_TX + 2, D0_CHY = D0_TY + 2, D0_CHZ = D0_TZ + 2;
constexpr int D0_CHN = D0_CHX * D0_CHY * D0_CHZ;  // staged coarse voxels
constexpr int D0_CRS = D0_CC + 8;                 // their LDS row (elements)
constexpr int D0_SHX = 2 * D0_TX + 2, D0_SHY = 2 * D0_TY + 2, D0_SHZ = 2 * D0_TZ + 2;
constexpr int D0_SHN = D0_SHX * D0_SHY * D0_SHZ;  // staged fine skip voxels
constexpr int D0_SPR = D0_SHX / 2;                // voxels per x-parity row
constexpr int D0_SRS = D0_SC + 8;
constexpr int D0_FN = 8 * D0_TX * D0_TY * D0_TZ;  // fine voxels per tile
constexpr int D0_ORS = 24;
constexpr int D0_NUC = (D0_CHN * (D0_CC / 8) + 511) / 512, D0_NUS = (D0_SHN * (D0_SC / 8) + 511) / 512;
static_assert(D0_FN == 512 && D0_CC == 32 && D0_SC == 16, "one fine voxel per thread; 4 + 2 chunks per staged voxel");

// wf[p][co][o][ci] = sum over the taps that parity p sends to coarse offset
// o of wt[co][tap][ci], ci < 32 (p = px + 2 py + 4 pz, o likewise).
__global__ void k_fold_d0(const uint16_t* __restrict__ wt, uint16_t* __restrict__ wf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 8 * 16 * 8 * D0_CC) return;
  const int ci = e % D0_CC, o = (e / D0_CC) % 8, co = (e / (D0_CC * 8)) % 16, p = e / (D0_CC * 8 * 16);
  // taps [lo, hi] of parity pp that land on coarse offset oo
  auto lo = [](int pp, int oo) { return pp == 0 ? (oo == 0 ? 0 : 1) : (oo == 0 ? 0 : 2); };
  auto hi = [](int pp, int oo) { return pp == 0 ? (oo == 0 ? 0 : 2) : (oo == 0 ? 1 : 2); };
  const int px = p & 1, py = (p >> 1) & 1, pz = p >> 2, ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
  float sum = 0.f;
  for (int dz = lo(pz, oz); dz <= hi(pz, oz); dz++)
    for (int dy = lo(py, oy); dy <= hi(py, oy); dy++)
      for (int dx = lo(px, ox); dx <= hi(px, ox); dx++)
        sum += bf2f(wt[(co * 27 + dx + 3 * (dy + 3 * dz)) * D0_CIN + ci]);
  wf[e] = f2bf(sum);
}

__global__ __launch_bounds__(512) void k_dec0_out(int n1, int n2, int n3, const uint16_t* __restrict__ skip, View sv,
                                                  const uint16_t* __restrict__ lo, View lv,
                                                  const uint16_t* __restrict__ wt, const uint16_t* __restrict__ wf,
                                                  const float* __restrict__ bias, const float* __restrict__ ow,
                                                  float* __restrict__ pred, double* __restrict__ pred_g) {
  __shared__ __align__(16) uint16_t lds_c[D0_CHN * D0_CRS];
  __shared__ __align__(16) uint16_t lds_s[D0_SHN * D0_SRS];
  __shared__ __align__(16) uint16_t lds_o[D0_FN * D0_ORS];
  const int m1 = n1 / 2, m2 = n2 / 2, m3 = n3 / 2;
  const int tx = (m1 + D0_TX - 1) / D0_TX, ty = (m2 + D0_TY - 1) / D0_TY, tz = (m3 + D0_TZ - 1) / D0_TZ;
  const int ntiles = tx * ty * tz;
  const int tid = threadIdx.x, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int px = wave & 1, py = (wave >> 1) & 1, pz = wave >> 2;
  const int tsub = lq >> 1, csub = 8 * (lq & 1);
  // Skip-half tap pairs of one MFMA (lanes 0..31: the first tap, 32..63 the
  // second, one step along z, y or x): pairs 0..8 (dx, dy, 0) + z, 9..11
  // (dx, 0, 2) + y, 12 (0, 2, 2) + x, 13 (2, 2, 2) alone.  The second tap's
  // staged address is the first's plus a per-lane constant of its direction.
  auto pair_tap = [](int k, int& dx, int& dy, int& dz, int& dir) {
    if (k < 9) { dx = k % 3; dy = k / 3; dz = 0; dir = 2; }
    else if (k < 12) { dx = k - 9; dy = 0; dz = 2; dir = 1; }
    else if (k == 12) { dx = 0; dy = 2; dz = 2; dir = 0; }
    else { dx = 2; dy = 2; dz = 2; dir = 3; }
  };
  uint4 bc[8], bs[14];
#pragma unroll
  for (int o = 0; o < 8; o++) bc[o] = *(const uint4*)(wf + ((wave * 16 + lr) * 8 + o) * D0_CC + 8 * lq);
#pragma unroll
  for (int k = 0; k < 14; k++) {
    int dx, dy, dz, dir;
    pair_tap(k, dx, dy, dz, dir);
    if (tsub) {
      dx += dir == 0;
      dy += dir == 1;
      dz += dir == 2;
    }
    const int tap = dx + 3 * (dy + 3 * dz);
    bs[k] = (tsub && dir == 3) ? make_uint4(0, 0, 0, 0)
                               : *(const uint4*)(wt + (lr * 27 + tap) * D0_CIN + D0_CC + csub);
  }
  const float bco = bias[lr];
  // LDS element offsets: coarse rows of this lane and wave parity; skip rows
  // (plus the second tap's step along z, y and x for lanes 32..63)
  const int cbase = ((pz * D0_CHY + py) * D0_CHX + lr + px) * D0_CRS + 8 * lq;
  const int sbase = (((py + D0_SHY * pz) * 2 + (px & 1)) * D0_SPR + lr) * D0_SRS + csub;
  const int sdz = tsub * D0_SHY * 2 * D0_SPR * D0_SRS, sdy = tsub * 2 * D0_SPR * D0_SRS;
  const int sdx = tsub * (px == 0 ? D0_SPR : 1 - D0_SPR) * D0_SRS;  // pair 12 has dx = 0

  uint4 vc[D0_NUC], vs[D0_NUS];
  auto load = [&](int t) {
    const int X0 = (t % tx) * D0_TX, Y0 = ((t / tx) % ty) * D0_TY, Z0 = (t / (tx * ty)) * D0_TZ;
#pragma unroll
    for (int u = 0; u < D0_NUC; u++) {
      const int e = tid + 512 * u, h = e >> 2, q = e & 3;
      const int gx = X0 - 1 + h % D0_CHX, gy = Y0 - 1 + (h / D0_CHX) % D0_CHY, gz = Z0 - 1 + h / (D0_CHX * D0_CHY);
      vc[u] = make_uint4(0, 0, 0, 0);
      if (e < D0_CHN * 4 && gx >= 0 && gx < m1 && gy >= 0 && gy < m2 && gz >= 0 && gz < m3)
        vc[u] = *(const uint4*)(lo + (gx + (int64_t)m1 * (gy + (int64_t)m2 * gz)) * lv.cs + lv.co + 8 * q);
    }
#pragma unroll
    for (int u = 0; u < D0_NUS; u++) {
      const int e = tid + 512 * u, h = e >> 1, q = e & 1;
      const int gx = 2 * X0 - 1 + h % D0_SHX, gy = 2 * Y0 - 1 + (h / D0_SHX) % D0_SHY,
                gz = 2 * Z0 - 1 + h / (D0_SHX * D0_SHY);
      vs[u] = make_uint4(0, 0, 0, 0);
      if (e < D0_SHN * 2 && gx >= 0 && gx < n1 && gy >= 0 && gy < n2 && gz >= 0 && gz < n3)
        vs[u] = *(const uint4*)(skip + (gx + (int64_t)n1 * (gy + (int64_t)n2 * gz)) * sv.cs + sv.co + 8 * q);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int u = 0; u < D0_NUC; u++) {
      const int e = tid + 512 * u;
      if (e < D0_CHN * 4) *(uint4*)(lds_c + (e >> 2) * D0_CRS + 8 * (e & 3)) = vc[u];
    }
#pragma unroll
    for (int u = 0; u < D0_NUS; u++) {
      const int e = tid + 512 * u, h = e >> 1;
      if (e < D0_SHN * 2) {
        const int hx = h % D0_SHX, hyz = h / D0_SHX;  // hyz = hy + D0_SHY hz
        *(uint4*)(lds_s + ((hyz * 2 + (hx & 1)) * D0_SPR + (hx >> 1)) * D0_SRS + 8 * (e & 1)) = vs[u];
      }
    }
  };

  int t = blockIdx.x;
  if (t < ntiles) load(t);
  for (; t < ntiles; t += gridDim.x) {
    __syncthreads();  // every wave is done with the previous tile's LDS
    store();
    if (t + (int)gridDim.x < ntiles) load(t + gridDim.x);
    __syncthreads();
    f4v acc[4];
#pragma unroll
    for (int m = 0; m < 4; m++) acc[m] = f4v{0.f, 0.f, 0.f, 0.f};
    // 22 steps of 4 MFMAs (one per coarse row): 8 coarse offsets o of the
    // upsampled half, then 14 tap pairs of the skip half.  The A fragments of
    // step k + 1 are read from LDS before the MFMAs of step k are issued.
    // Upsampled half: coarse voxel (X + px - 1 + ox, ...) is staged at
    // (lr + px + ox, ...); skip half: fine voxel (2 X + px + dx - 1, ...) at
    // staged x 2 lr + px + dx.
    auto lda = [&](int k, uint4 (&a)[4]) {
      if (k < 8) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const int h = (((m >> 1) + oz) * D0_CHY + ((m & 1) + oy)) * D0_CHX + ox;
          a[m] = *(const uint4*)(lds_c + cbase + h * D0_CRS);
        }
      } else {
        int dx, dy, dz, dir;
        pair_tap(k - 8, dx, dy, dz, dir);
        const int step = dir == 2 ? sdz : dir == 1 ? sdy : dir == 0 ? sdx : 0;
        // staged x of the first tap: 2 lr + px + dx (px's parity is in sbase)
        const int xo = (px + dx) / 2 * D0_SRS + ((px + dx) % 2 - px) * D0_SPR * D0_SRS;
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const int hyz = (2 * (m & 1) + dy) + D0_SHY * (2 * (m >> 1) + dz);
          a[m] = *(const uint4*)(lds_s + sbase + step + xo + hyz * 2 * D0_SPR * D0_SRS);
        }
      }
    };
    uint4 a0[4], a1[4];
    lda(0, a0);
#pragma unroll
    for (int k = 0; k < 22; k += 2) {
      lda(k + 1, a1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < 4; m++) acc[m] = mfma32(a0[m], k < 8 ? bc[k] : bs[k - 8], acc[m]);
      __builtin_amdgcn_sched_barrier(0);
      if (k + 2 < 22) lda(k + 2, a0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < 4; m++) acc[m] = mfma32(a1[m], k + 1 < 8 ? bc[k + 1] : bs[k - 7], acc[m]);
      __builtin_amdgcn_sched_barrier(0);
    }
    // ReLU output of voxel X = 4 lq + i, channel lr, as bf16 into the tile
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int fy = 2 * (m & 1) + py, fz = 2 * (m >> 1) + pz;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int fx = 2 * (4 * lq + i) + px;
        lds_o[((fz * 2 * D0_TY + fy) * 2 * D0_TX + fx) * D0_ORS + lr] = f2bf(fmaxf(acc[m][i] + bco, 0.f));
      }
    }
    __syncthreads();
    {
      const int X0 = (t % tx) * D0_TX, Y0 = ((t / tx) % ty) * D0_TY, Z0 = (t / (tx * ty)) * D0_TZ;
      const int fx = tid % (2 * D0_TX), fy = (tid / (2 * D0_TX)) % (2 * D0_TY), fz = tid / (4 * D0_TX * D0_TY);
      const int gx = 2 * X0 + fx, gy = 2 * Y0 + fy, gz = 2 * Z0 + fz;
      if (gx < n1 && gy < n2 && gz < n3) {
        float sum = 0.f;
#pragma unroll
        for (int c = 0; c < 16; c++) sum += ow[c] * bf2f(lds_o[tid * D0_ORS + c]);
        sum += ow[16];
        if (pred_g)  // FP64 into the interior of a ghosted box
          pred_g[(gx + 1) + (int64_t)(n1 + 2) * ((gy + 1) + (int64_t)(n2 + 2) * (gz + 1))] = (double)sum;
        else
          pred[gx + (int64_t)n1 * (gy + (int64_t)n2 * gz)] = sum;
      }
    }
  }
}

// ------------------------------------------------------- weight gradient --

constexpr int WG_VOX = 4096;  // voxels per block of the weight-gradient kernel
constexpr int WG_MAXP = 8;    // (co tile, ci tile) pairs per wave

// part[g][tap][co][ci] = sum_{v in group g} dz[v][co] x[v + o_tap][ci]
// (block: one group, one tap; MFMA with the voxels as the reduction index).
__global__ __launch_bounds__(256) void k_wgrad(int n1, int n2, int n3, const uint16_t* __restrict__ x, int cin,
                                               View xv, const uint16_t* __restrict__ dz, int cout,
                                               float* __restrict__ part) {
  const int g = blockIdx.x, tap = blockIdx.y;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
  const int ox = tap % 3 - 1, oy = (tap / 3) % 3 - 1, oz = tap / 9 - 1;
  const int nct = cout / 16, nkt = cin / 16, np = nct * nkt;
  const int64_t N = (int64_t)n1 * n2 * n3;
  f4v acc[WG_MAXP];
#pragma unroll
  for (int q = 0; q < WG_MAXP; q++) acc[q] = f4v{0.f, 0.f, 0.f, 0.f};
  const int64_t vbeg = (int64_t)g * WG_VOX;
  for (int kc = 0; kc < WG_VOX; kc += 16) {
    int64_t vv[4], nb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t v = vbeg + kc + 4 * lq + i;
      vv[i] = v < N ? v : -1;
      nb[i] = -1;
      if (v < N) {
        const int xi = (int)(v % n1), yi = (int)((v / n1) % n2), zi = (int)(v / ((int64_t)n1 * n2));
        const int xa = xi + ox, ya = yi + oy, za = zi + oz;
        if (xa >= 0 && xa < n1 && ya >= 0 && ya < n2 && za >= 0 && za < n3)
          nb[i] = xa + (int64_t)n1 * (ya + (int64_t)n2 * za);
      }
    }
    if (vbeg + kc >= N) break;
#pragma unroll
    for (int q = 0; q < WG_MAXP; q++) {
      const int p = wave + 4 * q;
      if (p >= np) break;
      const int ct = p / nkt, kt = p % nkt;
      s4v a, bb;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[i] = vv[i] >= 0 ? (short)dz[vv[i] * cout + ct * 16 + lr] : (short)0;
        bb[i] = nb[i] >= 0 ? (short)x[nb[i] * xv.cs + xv.co + kt * 16 + lr] : (short)0;
      }
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[q], 0, 0, 0);
    }
  }
  // D[row = co 4 lq + i][col = ci lr]
#pragma unroll
  for (int q = 0; q < WG_MAXP; q++) {
    const int p = wave + 4 * q;
    if (p >= np) break;
    const int ct = p / nkt, kt = p % nkt;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = ct * 16 + 4 * lq + i, ci = kt * 16 + lr;
      part[(((int64_t)g * 27 + tap) * cout + co) * cin + ci] = acc[q][i];
    }
  }
}

// dw[co][tap][ci] = sum_g part[g][tap][co][ci] (fixed order).
__global__ void k_wgrad_reduce(int ng, int cin, int cout, const float* __restrict__ part, float* __restrict__ dw) {
  const int64_t n = (int64_t)27 * cout * cin;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int ci = (int)(e % cin), tap = (int)((e / cin) % 27), co = (int)(e / ((int64_t)27 * cin));
  const int64_t pe = ((int64_t)tap * cout + co) * cin + ci;
  float s = 0.f;
  for (int g = 0; g < ng; g++) s += part[(int64_t)g * n + pe];
  dw[e] = s;
}

// ---------------------------------------------- column sums (bias grads) --

constexpr int CSUM_VOX = 4096;

// part[g][c] = sum over the group's voxels of a[v][c] (bf16), c < C <= 128.
__global__ __launch_bounds__(256) void k_colsum(int64_t N, int C, const uint16_t* __restrict__ a,
                                                float* __restrict__ part) {
  __shared__ float sm[256];
  const int t = threadIdx.x, S = 256 / C;
  const int c = t % C, s = t / C;
  float acc = 0.f;
  const int64_t v0 = (int64_t)blockIdx.x * CSUM_VOX;
  if (s < S)
    for (int64_t v = v0 + s; v < v0 + CSUM_VOX && v < N; v += S) acc += bf2f(a[v * C + c]);
  sm[t] = acc;
  __syncthreads();
  if (t < C) {
    float tot = 0.f;
    for (int q = 0; q < S; q++) tot += sm[q * C + t];
    part[(int64_t)blockIdx.x * C + t] = tot;
  }
}

__global__ void k_colsum_reduce(int ng, int C, const float* __restrict__ part, float* __restrict__ out) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int g = 0; g < ng; g++) s += part[(int64_t)g * C + c];
  out[c] = s;
}

// ------------------------------------------------- pooling and upsampling --

// coarse[v][c] = bf16(mean of the 8 children), children from a channel view.
// One thread per coarse voxel and 8-channel chunk: eight 16-byte child loads
// and one 16-byte store (C, the view's stride and offset are multiples of 8);
// the children are summed in the same order for every channel.
__global__ void k_pool(int n1, int n2, int n3, int C, const uint16_t* __restrict__ fine, View fv,
                       uint16_t* __restrict__ coarse) {
  const int m1 = n1 / 2, m2 = n2 / 2, m3 = n3 / 2, C8 = C / 8;
  const int64_t n = (int64_t)m1 * m2 * m3 * C8;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c8 = (int)(e % C8);
    const int64_t vc = e / C8;
    const int x = (int)(vc % m1), y = (int)((vc / m1) % m2), z = (int)(vc / ((int64_t)m1 * m2));
    uint4 ch[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int fx = 2 * x + (q & 1), fy = 2 * y + ((q >> 1) & 1), fz = 2 * z + (q >> 2);
      const int64_t vf = fx + (int64_t)n1 * (fy + (int64_t)n2 * fz);
      ch[q] = *(const uint4*)(fine + vf * fv.cs + fv.co + 8 * c8);
    }
    uint16_t r[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      float s = 0.f;
#pragma unroll
      for (int q = 0; q < 8; q++) s += bf2f(((const uint16_t*)&ch[q])[k]);
      r[k] = f2bf(s * 0.125f);
    }
    *(uint4*)(coarse + vc * C + 8 * c8) = *(const uint4*)r;
  }
}

// dfine[child] += dcoarse / 8 (FP32, into a channel view).
__global__ void k_pool_bwd(int n1, int n2, int n3, int C, const float* __restrict__ dcoarse, float* __restrict__ dfine,
                           View fv) {
  const int64_t n = (int64_t)n1 * n2 * n3 * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t vf = e / C;
    const int x = (int)(vf % n1), y = (int)((vf / n1) % n2), z = (int)(vf / ((int64_t)n1 * n2));
    const int m1 = n1 / 2, m2 = n2 / 2;
    const int64_t vc = x / 2 + (int64_t)m1 * (y / 2 + (int64_t)m2 * (z / 2));
    dfine[vf * fv.cs + fv.co + c] += dcoarse[vc * C + c] * 0.125f;
  }
}

// fine view [v][c] = coarse[parent][c] (bf16 copy); n* are the fine dims.
// One thread per fine voxel and 8-channel chunk (16-byte copies; channel
// counts, strides and offsets are multiples of 8).
__global__ void k_up(int n1, int n2, int n3, int C, const uint16_t* __restrict__ coarse, uint16_t* __restrict__ fine,
                     View fv) {
  const int nq = C / 8;
  const int64_t n = (int64_t)n1 * n2 * n3 * nq;
  const int m1 = n1 / 2, m2 = n2 / 2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(e % nq);
    const int64_t vf = e / nq;
    const int x = (int)(vf % n1), y = (int)((vf / n1) % n2), z = (int)(vf / ((int64_t)n1 * n2));
    const int64_t vc = x / 2 + (int64_t)m1 * (y / 2 + (int64_t)m2 * (z / 2));
    *(uint4*)(fine + vf * fv.cs + fv.co + 8 * q) = *(const uint4*)(coarse + vc * C + 8 * q);
  }
}

// dcoarse[v][c] = sum of the 8 children of a fine FP32 view (fixed order).
__global__ void k_up_bwd(int n1, int n2, int n3, int C, const float* __restrict__ dfine, View fv,
                         float* __restrict__ dcoarse) {
  const int m1 = n1 / 2, m2 = n2 / 2, m3 = n3 / 2;
  const int64_t n = (int64_t)m1 * m2 * m3 * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t vc = e / C;
    const int x = (int)(vc % m1), y = (int)((vc / m1) % m2), z = (int)(vc / ((int64_t)m1 * m2));
    float s = 0.f;
    for (int q = 0; q < 8; q++) {
      const int fx = 2 * x + (q & 1), fy = 2 * y + ((q >> 1) & 1), fz = 2 * z + (q >> 2);
      const int64_t vf = fx + (int64_t)n1 * (fy + (int64_t)n2 * fz);
      s += dfine[vf * fv.cs + fv.co + c];
    }
    dcoarse[vc * C + c] = s;
  }
}

// dz[v][c] = bf16(dact[v][c] * (act[v][c] > 0)).
__global__ void k_relu_bwd(int64_t N, int C, const float* __restrict__ dact, View dv, const uint16_t* __restrict__ act,
                           View av, uint16_t* __restrict__ dz) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N * C; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const int64_t v = e / C;
    const uint16_t a = act[v * av.cs + av.co + c];
    const bool pos = a != 0 && !(a & 0x8000u);
    dz[e] = f2bf(pos ? dact[v * dv.cs + dv.co + c] : 0.f);
  }
}

// ----------------------------------------------------------- output layer --

// pred = d0 . w + b; with a target also dpred = 2 (pred - t) / N and the
// per-block partial sum of (pred - t)^2.
__global__ __launch_bounds__(256) void k_out_fwd(int64_t N, const uint16_t* __restrict__ d0, const float* __restrict__ w,
                                                 const float* __restrict__ target, float* __restrict__ pred,
                                                 float* __restrict__ dpred, double* partials) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (v < N) {
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 16; c++) s += w[c] * bf2f(d0[v * 16 + c]);
    s += w[16];
    pred[v] = s;
    if (target) {
      const float r = s - target[v];
      dpred[v] = 2.f * r / (float)N;
      acc[0] = (double)r * r;
    }
  }
  if (target) block_partials<1>(acc, partials);
}

// dd0[v][c] = dpred[v] w[c]; partial sums of dpred d0[v][c] (16) and dpred (1).
__global__ __launch_bounds__(256) void k_out_bwd(int64_t N, const uint16_t* __restrict__ d0, const float* __restrict__ w,
                                                 const float* __restrict__ dpred, float* __restrict__ dd0,
                                                 float* __restrict__ part) {
  __shared__ float sm[17][256];
  const int t = threadIdx.x;
  float acc[17];
#pragma unroll
  for (int c = 0; c < 17; c++) acc[c] = 0.f;
  const int64_t v0 = (int64_t)blockIdx.x * CSUM_VOX;
  for (int64_t v = v0 + t; v < v0 + CSUM_VOX && v < N; v += 256) {
    const float g = dpred[v];
#pragma unroll
    for (int c = 0; c < 16; c++) {
      dd0[v * 16 + c] = g * w[c];
      acc[c] += g * bf2f(d0[v * 16 + c]);
    }
    acc[16] += g;
  }
#pragma unroll
  for (int c = 0; c < 17; c++) sm[c][t] = acc[c];
  __syncthreads();
  if (t < 17) {
    float s = 0.f;
    for (int q = 0; q < 256; q++) s += sm[t][q];
    part[(int64_t)blockIdx.x * 17 + t] = s;
  }
}

// ------------------------------------------------- energy-norm training --

// dg (ghosted (n+2)^3, zero ghosts) interior = pred - target.
__global__ void k_delta_g(int n1, int n2, int n3, const float* __restrict__ pred, const float* __restrict__ target,
                          float* __restrict__ dg) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, N = (int64_t)n1 * n2 * n3;
  if (v >= N) return;
  const int i = (int)(v % n1), j = (int)((v / n1) % n2), k = (int)(v / ((int64_t)n1 * n2));
  dg[(i + 1) + (int64_t)(n1 + 2) * ((j + 1) + (int64_t)(n2 + 2) * (k + 1))] = pred[v] - target[v];
}

__global__ void k_scale_f(int64_t n, const float* __restrict__ a, float s, float* __restrict__ b) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) b[e] = s * a[e];
}

// ------------------------------------------------------------- optimiser --

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one layer's weights: wt[co][tap*cin + ci] = W[co][tap][ci]
// (forward) and wb[ci][tap*cout + co] = W[co][26 - tap][ci] (backward data).
__global__ void k_pack(int cin, int cout, const float* __restrict__ W, uint16_t* __restrict__ wt,
                       uint16_t* __restrict__ wb) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, n = (int64_t)27 * cin * cout;
  if (e >= n) return;
  const int ci = (int)(e % cin), tap = (int)((e / cin) % 27), co = (int)(e / ((int64_t)27 * cin));
  const uint16_t h = f2bf(W[e]);
  wt[e] = h;
  wb[(int64_t)ci * 27 * cout + (int64_t)(26 - tap) * cout + co] = h;
}

// ------------------------------------------------------------------ host --

namespace {

inline int nblk(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }
// Element kernels over voxel x channel index spaces loop over a capped grid:
// a dispatch holds at most 2^32 - 1 work-items per dimension, which one
// thread per element would exceed at 512^3 x 32 channels.
inline int nblk_cap(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 1 << 20); }

struct LayerDef {
  int cin, cout;
};
constexpr LayerDef kLayers[NN_NLAYERS] = {{16, 16}, {16, 32}, {32, 64}, {64, 64}, {128, 64}, {96, 32}, {48, 16}};
enum { L_E0 = 0, L_E1, L_E2, L_E3, L_D2, L_D1, L_D0 };

}  // namespace

int64_t nn_param_count() {
  int64_t n = 0;
  for (const LayerDef& l : kLayers) n += (int64_t)27 * l.cin * l.cout + l.cout;
  return n + 17;
}

struct NnNet {
  DevBuf params, grads, m, v;  // FP32, nn_param_count()
  DevBuf wt, wb;               // bf16 packed convolution weights
  DevBuf wf;                   // d0's upsampled half folded per parity (inference)
  int64_t w_off[NN_NLAYERS], b_off[NN_NLAYERS], wt_off[NN_NLAYERS], out_off = 0;
  int t = 0;
  bool packed = false;
  int n1 = 0, n2 = 0, n3 = 0;
  // activations (bf16) and gradients (FP32) for the current grid
  DevBuf cat0, p1, cat1, p2, cat2, p3, e3, d2, d1, d0, dpred;
  DevBuf dd0, dcat0, dd1, dcat1, dd2, dcat2, de3, dp3, dp2, dp1, dz;
  DevBuf wpart, cpart, opart;
  DevBuf loss;            // device double
  DevBuf eg_dg, eg_w;     // energy loss: ghosted error, operator output

  NnNet() {
    int64_t o = 0, ow = 0;
    for (int l = 0; l < NN_NLAYERS; l++) {
      w_off[l] = o;
      o += (int64_t)27 * kLayers[l].cin * kLayers[l].cout;
      b_off[l] = o;
      o += kLayers[l].cout;
      wt_off[l] = ow;
      ow += (int64_t)27 * kLayers[l].cin * kLayers[l].cout;
    }
    out_off = o;
  }
  int64_t nw() const { return wt_off[NN_NLAYERS - 1] + (int64_t)27 * kLayers[NN_NLAYERS - 1].cin * kLayers[NN_NLAYERS - 1].cout; }
  float* P() { return (float*)params.p; }
  float* Gr() { return (float*)grads.p; }
  uint16_t* WT(int l) { return (uint16_t*)wt.p + wt_off[l]; }
  uint16_t* WB(int l) { return (uint16_t*)wb.p + wt_off[l]; }

  int init() {
    const int64_t n = nn_param_count();
    for (DevBuf* b : {&params, &grads, &m, &v}) {
      ELSIM_TRY(dev_alloc(*b, n * sizeof(float)));
      ELSIM_HIP(hipMemsetAsync(b->p, 0, n * sizeof(float), stream()));
    }
    ELSIM_TRY(dev_alloc(wt, nw() * 2));
    ELSIM_TRY(dev_alloc(wb, nw() * 2));
    ELSIM_TRY(dev_alloc(wf, 8 * 16 * 8 * D0_CC * 2));
    ELSIM_TRY(dev_alloc(loss, sizeof(double)));
    return ELSIM_OK;
  }
  int pack() {
    if (packed) return ELSIM_OK;
    for (int l = 0; l < NN_NLAYERS; l++) {
      const int64_t n = (int64_t)27 * kLayers[l].cin * kLayers[l].cout;
      k_pack<<<nblk(n), 256, 0, stream()>>>(kLayers[l].cin, kLayers[l].cout, P() + w_off[l], WT(l), WB(l));
    }
    k_fold_d0<<<nblk(8 * 16 * 8 * D0_CC), 256, 0, stream()>>>(WT(L_D0), (uint16_t*)wf.p);
    ELSIM_CHECK_LAUNCH();
    packed = true;
    return ELSIM_OK;
  }
  int ensure(int a1, int a2, int a3, bool train) {
    if (a1 % 8 || a2 % 8 || a3 % 8 || a1 <= 0 || a2 <= 0 || a3 <= 0) {
      set_error("warm-start network: grid %dx%dx%d must be a positive multiple of 8 in every direction", a1, a2, a3);
      return ELSIM_ERR_ARG;
    }
    const int64_t N0 = (int64_t)a1 * a2 * a3, N1 = N0 / 8, N2 = N1 / 8, N3 = N2 / 8;
    ELSIM_TRY(dev_alloc(cat0, N0 * 48 * 2));
    ELSIM_TRY(dev_alloc(p1, N1 * 16 * 2));
    ELSIM_TRY(dev_alloc(cat1, N1 * 96 * 2));
    ELSIM_TRY(dev_alloc(p2, N2 * 32 * 2));
    ELSIM_TRY(dev_alloc(cat2, N2 * 128 * 2));
    ELSIM_TRY(dev_alloc(p3, N3 * 64 * 2));
    ELSIM_TRY(dev_alloc(e3, N3 * 64 * 2));
    ELSIM_TRY(dev_alloc(d2, N2 * 64 * 2));
    ELSIM_TRY(dev_alloc(d1, N1 * 32 * 2));
    ELSIM_TRY(dev_alloc(d0, N0 * 16 * 2));
    if (train) {
      ELSIM_TRY(dev_alloc(dpred, N0 * 4));
      ELSIM_TRY(dev_alloc(dd0, N0 * 16 * 4));
      ELSIM_TRY(dev_alloc(dcat0, N0 * 48 * 4));
      ELSIM_TRY(dev_alloc(dd1, N1 * 32 * 4));
      ELSIM_TRY(dev_alloc(dcat1, N1 * 96 * 4));
      ELSIM_TRY(dev_alloc(dd2, N2 * 64 * 4));
      ELSIM_TRY(dev_alloc(dcat2, N2 * 128 * 4));
      ELSIM_TRY(dev_alloc(de3, N3 * 64 * 4));
      ELSIM_TRY(dev_alloc(dp3, N3 * 64 * 4));
      ELSIM_TRY(dev_alloc(dp2, N2 * 32 * 4));
      ELSIM_TRY(dev_alloc(dp1, N1 * 16 * 4));
      ELSIM_TRY(dev_alloc(dz, N0 * 16 * 2));
      int64_t wp = 0;
      const int64_t Nl[4] = {N0, N1, N2, N3};
      const int lvl[NN_NLAYERS] = {0, 1, 2, 3, 2, 1, 0};
      for (int l = 0; l < NN_NLAYERS; l++) {
        const int64_t ng = (Nl[lvl[l]] + WG_VOX - 1) / WG_VOX;
        wp = std::max(wp, ng * 27 * kLayers[l].cin * kLayers[l].cout);
      }
      ELSIM_TRY(dev_alloc(wpart, wp * 4));
      ELSIM_TRY(dev_alloc(cpart, ((N0 + CSUM_VOX - 1) / CSUM_VOX) * 128 * 4));
      ELSIM_TRY(dev_alloc(opart, ((N0 + CSUM_VOX - 1) / CSUM_VOX) * 17 * 4));
    }
    n1 = a1;
    n2 = a2;
    n3 = a3;
    return ELSIM_OK;
  }

  // forward into pred (FP32); with a target also dpred and the loss partials
  // infer: the activations are not kept for a backward pass, so the
  // decoder reads its upsampled halves at half resolution while staging
  // (no full-resolution copies; the same values, so the same results).
  // in_used: x holds rows of in_used channels, the network's input
  // channels [in_used, 16) being zero (8: the first layer runs on 8-channel
  // slabs).
  int forward(const uint16_t* x0, float* pred, const float* target, double* loss_host, bool infer = false,
              int in_used = NN_IN_CH, double* pred_g = nullptr) {
    ELSIM_TRY(pack());
    hipStream_t s = stream();
    const int a1 = n1, a2 = n2, a3 = n3;
    const int m1 = a1 / 2, m2 = a2 / 2, m3 = a3 / 2, q1 = a1 / 4, q2 = a2 / 4, q3 = a3 / 4, r1 = a1 / 8, r2 = a2 / 8,
              r3 = a3 / 8;
    const int64_t N0 = (int64_t)a1 * a2 * a3, N1 = N0 / 8, N2 = N1 / 8, N3 = N2 / 8;
    uint16_t *C0 = (uint16_t*)cat0.p, *C1 = (uint16_t*)cat1.p, *C2 = (uint16_t*)cat2.p;
    const float* B = P();
    // encoder.  The skip halves of the decoder inputs live in cat0/1/2 after
    // the upsampled halves (channels-last rows of 48, 96, 128); at inference
    // the upsampled halves are read at half resolution and never stored, so
    // the skip tensors are kept compact (rows of 16, 32, 64: every cache line
    // the encoder writes and the pooling and the decoder read is all skip
    // data) and the decoder addresses them with a negative channel offset.
    const View s0 = infer ? View{16, 0} : View{48, 32}, s1 = infer ? View{32, 0} : View{96, 64},
               s2 = infer ? View{64, 0} : View{128, 64};
    const View i0 = infer ? View{16, -32} : View{48, 0}, i1 = infer ? View{32, -64} : View{96, 0},
               i2 = infer ? View{64, -64} : View{128, 0};
    // each encoder convolution also writes its 2x2x2-pooled output
    const Coarse none{nullptr, 0, View{0, 0}};
    ELSIM_TRY(conv<0>(a1, a2, a3, x0, 16, View{in_used, 0}, WT(L_E0), B + b_off[L_E0], 1, 16, C0, s0, s, none,
                      in_used, (uint16_t*)p1.p));
    ELSIM_TRY(conv<0>(m1, m2, m3, (uint16_t*)p1.p, 16, View{16, 0}, WT(L_E1), B + b_off[L_E1], 1, 32, C1, s1, s, none,
                      0, (uint16_t*)p2.p));
    ELSIM_TRY(conv<0>(q1, q2, q3, (uint16_t*)p2.p, 32, View{32, 0}, WT(L_E2), B + b_off[L_E2], 1, 64, C2, s2, s, none,
                      0, (uint16_t*)p3.p));
    ELSIM_TRY(conv<0>(r1, r2, r3, (uint16_t*)p3.p, 64, View{64, 0}, WT(L_E3), B + b_off[L_E3], 1, 64, e3.p,
                      View{64, 0}, s));
    // decoder
    if (!infer) k_up<<<nblk_cap(N2 * 8), 256, 0, s>>>(q1, q2, q3, 64, (uint16_t*)e3.p, C2, View{128, 0});
    ELSIM_TRY(conv<0>(q1, q2, q3, C2, 128, i2, WT(L_D2), B + b_off[L_D2], 1, 64, d2.p, View{64, 0}, s,
                      infer ? Coarse{(uint16_t*)e3.p, 64, View{64, 0}} : none));
    if (!infer) k_up<<<nblk_cap(N1 * 8), 256, 0, s>>>(m1, m2, m3, 64, (uint16_t*)d2.p, C1, View{96, 0});
    ELSIM_TRY(conv<0>(m1, m2, m3, C1, 96, i1, WT(L_D1), B + b_off[L_D1], 1, 32, d1.p, View{32, 0}, s,
                      infer ? Coarse{(uint16_t*)d2.p, 64, View{64, 0}} : none));
    if (infer && !target) {
      // d0, its ReLU and the output layer in one pass, the upsampled half
      // convolved at d1's resolution with the folded weights
      const int nt = ((m1 + D0_TX - 1) / D0_TX) * ((m2 + D0_TY - 1) / D0_TY) * ((m3 + D0_TZ - 1) / D0_TZ);
      k_dec0_out<<<std::min(nt, cu_count()), 512, 0, s>>>(a1, a2, a3, C0, View{16, 0}, (uint16_t*)d1.p, View{32, 0},
                                                          WT(L_D0), (uint16_t*)wf.p, B + b_off[L_D0], B + out_off,
                                                          pred, pred_g);
      ELSIM_CHECK_LAUNCH();
      return ELSIM_OK;
    }
    if (pred_g) {
      set_error("network forward: the FP64 ghosted output is written by the inference path only");
      return ELSIM_ERR_ARG;
    }
    if (!infer) k_up<<<nblk_cap(N0 * 4), 256, 0, s>>>(a1, a2, a3, 32, (uint16_t*)d1.p, C0, View{48, 0});
    ELSIM_TRY(conv<0>(a1, a2, a3, C0, 48, i0, WT(L_D0), B + b_off[L_D0], 1, 16, d0.p, View{16, 0}, s,
                      infer ? Coarse{(uint16_t*)d1.p, 32, View{32, 0}} : none));
    const int ob = nblk(N0);
    if (target && ob > kMaxRedBlocks) {
      set_error("grid too large for the loss reduction");
      return ELSIM_ERR_ARG;
    }
    k_out_fwd<<<ob, 256, 0, s>>>(N0, (uint16_t*)d0.p, B + out_off, target, pred, (float*)dpred.p,
                                 workspace().partials);
    ELSIM_CHECK_LAUNCH();
    (void)N3;
    if (target && loss_host) {
      std::vector<double> h(ob);
      ELSIM_HIP(hipMemcpyAsync(h.data(), workspace().partials, ob * sizeof(double), hipMemcpyDeviceToHost, s));
      ELSIM_HIP(hipStreamSynchronize(s));
      double tot = 0.0;
      for (double x : h) tot += x;
      *loss_host = tot / (double)N0;
    }
    return ELSIM_OK;
  }

  int layer_grads(int l, int a1, int a2, int a3, const uint16_t* x, View xv, const float* dact, View dv,
                  const uint16_t* act, View av) {
    hipStream_t s = stream();
    const int cin = kLayers[l].cin, cout = kLayers[l].cout;
    const int64_t N = (int64_t)a1 * a2 * a3;
    uint16_t* Z = (uint16_t*)dz.p;
    k_relu_bwd<<<nblk_cap(N * cout), 256, 0, s>>>(N, cout, dact, dv, act, av, Z);
    const int ng = (int)((N + WG_VOX - 1) / WG_VOX);
    k_wgrad<<<dim3(ng, 27), 256, 0, s>>>(a1, a2, a3, x, cin, xv, Z, cout, (float*)wpart.p);
    k_wgrad_reduce<<<nblk((int64_t)27 * cin * cout), 256, 0, s>>>(ng, cin, cout, (float*)wpart.p,
                                                                   Gr() + w_off[l]);
    const int cg = (int)((N + CSUM_VOX - 1) / CSUM_VOX);
    k_colsum<<<cg, 256, 0, s>>>(N, cout, Z, (float*)cpart.p);
    k_colsum_reduce<<<1, 128, 0, s>>>(cg, cout, (float*)cpart.p, Gr() + b_off[l]);
    ELSIM_CHECK_LAUNCH();
    return ELSIM_OK;
  }

  int backward(const uint16_t* x0) {
    hipStream_t s = stream();
    const int a1 = n1, a2 = n2, a3 = n3;
    const int m1 = a1 / 2, m2 = a2 / 2, m3 = a3 / 2, q1 = a1 / 4, q2 = a2 / 4, q3 = a3 / 4, r1 = a1 / 8, r2 = a2 / 8,
              r3 = a3 / 8;
    const int64_t N0 = (int64_t)a1 * a2 * a3, N1 = N0 / 8, N2 = N1 / 8, N3 = N2 / 8;
    uint16_t *C0 = (uint16_t*)cat0.p, *C1 = (uint16_t*)cat1.p, *C2 = (uint16_t*)cat2.p;
    float *DC0 = (float*)dcat0.p, *DC1 = (float*)dcat1.p, *DC2 = (float*)dcat2.p;
    uint16_t* Z = (uint16_t*)dz.p;
    // output layer
    const int og = (int)((N0 + CSUM_VOX - 1) / CSUM_VOX);
    k_out_bwd<<<og, 256, 0, s>>>(N0, (uint16_t*)d0.p, P() + out_off, (float*)dpred.p, (float*)dd0.p,
                                 (float*)opart.p);
    k_colsum_reduce<<<1, 128, 0, s>>>(og, 17, (float*)opart.p, Gr() + out_off);
    // d0
    ELSIM_TRY(layer_grads(L_D0, a1, a2, a3, C0, View{48, 0}, (float*)dd0.p, View{16, 0}, (uint16_t*)d0.p,
                          View{16, 0}));
    ELSIM_TRY(conv<1>(a1, a2, a3, Z, 16, View{16, 0}, WB(L_D0), nullptr, 0, 48, DC0, View{48, 0}, s));
    k_up_bwd<<<nblk_cap(N1 * 32), 256, 0, s>>>(a1, a2, a3, 32, DC0, View{48, 0}, (float*)dd1.p);
    // d1
    ELSIM_TRY(layer_grads(L_D1, m1, m2, m3, C1, View{96, 0}, (float*)dd1.p, View{32, 0}, (uint16_t*)d1.p,
                          View{32, 0}));
    ELSIM_TRY(conv<1>(m1, m2, m3, Z, 32, View{32, 0}, WB(L_D1), nullptr, 0, 96, DC1, View{96, 0}, s));
    k_up_bwd<<<nblk_cap(N2 * 64), 256, 0, s>>>(m1, m2, m3, 64, DC1, View{96, 0}, (float*)dd2.p);
    // d2
    ELSIM_TRY(layer_grads(L_D2, q1, q2, q3, C2, View{128, 0}, (float*)dd2.p, View{64, 0}, (uint16_t*)d2.p,
                          View{64, 0}));
    ELSIM_TRY(conv<1>(q1, q2, q3, Z, 64, View{64, 0}, WB(L_D2), nullptr, 0, 128, DC2, View{128, 0}, s));
    k_up_bwd<<<nblk_cap(N3 * 64), 256, 0, s>>>(q1, q2, q3, 64, DC2, View{128, 0}, (float*)de3.p);
    // e3
    ELSIM_TRY(layer_grads(L_E3, r1, r2, r3, (uint16_t*)p3.p, View{64, 0}, (float*)de3.p, View{64, 0},
                          (uint16_t*)e3.p, View{64, 0}));
    ELSIM_TRY(conv<1>(r1, r2, r3, Z, 64, View{64, 0}, WB(L_E3), nullptr, 0, 64, dp3.p, View{64, 0}, s));
    k_pool_bwd<<<nblk_cap(N2 * 64), 256, 0, s>>>(q1, q2, q3, 64, (float*)dp3.p, DC2, View{128, 64});
    // e2
    ELSIM_TRY(layer_grads(L_E2, q1, q2, q3, (uint16_t*)p2.p, View{32, 0}, DC2, View{128, 64}, C2, View{128, 64}));
    ELSIM_TRY(conv<1>(q1, q2, q3, Z, 64, View{64, 0}, WB(L_E2), nullptr, 0, 32, dp2.p, View{32, 0}, s));
    k_pool_bwd<<<nblk_cap(N1 * 32), 256, 0, s>>>(m1, m2, m3, 32, (float*)dp2.p, DC1, View{96, 64});
    // e1
    ELSIM_TRY(layer_grads(L_E1, m1, m2, m3, (uint16_t*)p1.p, View{16, 0}, DC1, View{96, 64}, C1, View{96, 64}));
    ELSIM_TRY(conv<1>(m1, m2, m3, Z, 32, View{32, 0}, WB(L_E1), nullptr, 0, 16, dp1.p, View{16, 0}, s));
    k_pool_bwd<<<nblk_cap(N0 * 16), 256, 0, s>>>(a1, a2, a3, 16, (float*)dp1.p, DC0, View{48, 32});
    // e0
    ELSIM_TRY(layer_grads(L_E0, a1, a2, a3, x0, View{16, 0}, DC0, View{48, 32}, C0, View{48, 32}));
    ELSIM_CHECK_LAUNCH();
    return ELSIM_OK;
  }

  // loss = e.(-A e)/N with e = pred - target and A the homogeneous phP
  // operator of the coefficient kap (ghosted FP32, scale 1); dpred = -2 A e/N.
  int energy_loss(const float* pred, const float* target, const float* kap, double* loss_host) {
    hipStream_t s = stream();
    const Dims d{n1, n2, n3};
    const int64_t N = d.N(), G = d.G();
    ELSIM_TRY(dev_alloc(eg_dg, G * sizeof(float)));
    ELSIM_TRY(dev_alloc(eg_w, N * sizeof(float)));
    ELSIM_HIP(hipMemsetAsync(eg_dg.p, 0, G * sizeof(float), s));
    k_delta_g<<<nblk(N), 256, 0, s>>>(n1, n2, n3, pred, target, (float*)eg_dg.p);
    ELSIM_CHECK_LAUNCH();
    ELSIM_TRY(launch_matvec_f32(ELSIM_OP_PHP, d, (const float*)eg_dg.p, kap, nullptr, 1.0f, (float*)eg_w.p,
                                (double*)loss.p, true));
    k_scale_f<<<nblk(N), 256, 0, s>>>(N, (const float*)eg_w.p, -2.f / (float)N, (float*)dpred.p);
    ELSIM_CHECK_LAUNCH();
    double dot = 0.0;
    ELSIM_HIP(hipMemcpyAsync(&dot, loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
    ELSIM_HIP(hipStreamSynchronize(s));
    if (loss_host) *loss_host = -dot / (double)N;
    return ELSIM_OK;
  }

  int adam(float lr) {
    t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    const int64_t n = nn_param_count();
    k_adam<<<nblk(n), 256, 0, stream()>>>(n, P(), Gr(), (float*)m.p, (float*)v.p, lr, b1, b2, 1e-8f, c1, c2);
    ELSIM_CHECK_LAUNCH();
    packed = false;
    return ELSIM_OK;
  }

  void release() {
    for (DevBuf* b : {&params, &grads, &m, &v, &wt, &wb, &wf, &cat0, &p1, &cat1, &p2, &cat2, &p3, &e3, &d2, &d1, &d0,
                      &dpred, &dd0, &dcat0, &dd1, &dcat1, &dd2, &dcat2, &de3, &dp3, &dp2, &dp1, &dz, &wpart, &cpart,
                      &opart, &loss, &eg_dg, &eg_w})
      dev_free(*b);
  }
};

int nn_forward(NnNet* net, int n1, int n2, int n3, const uint16_t* x, float* pred, int in_used) {
  ELSIM_TRY(net->ensure(n1, n2, n3, false));
  return net->forward(x, pred, nullptr, nullptr, true, in_used);
}

int nn_forward_g(NnNet* net, int n1, int n2, int n3, const uint16_t* x, double* pred_g, int in_used) {
  ELSIM_TRY(net->ensure(n1, n2, n3, false));
  return net->forward(x, nullptr, nullptr, nullptr, true, in_used, pred_g);
}

}  // namespace elsim

using namespace elsim;

namespace {
// Networks are held in an id map, like the SBM states: a handle that was
// never issued, or whose network was destroyed, is rejected instead of
// dereferenced.
std::mutex g_nn_mu;
std::map<int64_t, std::unique_ptr<NnNet>> g_nn;
int64_t g_nn_next = 1;
NnNet* net_of(const int64_t* h) {
  if (!h || !*h) return nullptr;
  std::lock_guard<std::mutex> lk(g_nn_mu);
  auto it = g_nn.find(*h);
  if (it == g_nn.end()) {
    set_error("unknown network handle %lld", (long long)*h);
    return nullptr;
  }
  return it->second.get();
}
}  // namespace

namespace elsim {
NnNet* nn_lookup(int64_t h) { return net_of(&h); }
}  // namespace elsim

extern "C" {

int elsim_nn_param_count(void) { return (int)nn_param_count(); }

int elsim_nn_create(int64_t* handle) {
  if (!handle) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  auto n = std::make_unique<NnNet>();
  const int rc = n->init();
  if (rc < 0) {
    n->release();
    return rc;
  }
  std::lock_guard<std::mutex> lk(g_nn_mu);
  const int64_t id = g_nn_next++;
  g_nn[id] = std::move(n);
  *handle = id;
  return ELSIM_OK;
}

int elsim_nn_destroy(const int64_t* handle) {
  NnNet* n = net_of(handle);
  if (!n) return ELSIM_ERR_ARG;
  (void)hipStreamSynchronize(stream());
  // every SBM state that still holds this network falls back to the
  // extrapolated start (no state keeps a pointer to freed memory)
  sbm_forget_net(n);
  n->release();
  std::lock_guard<std::mutex> lk(g_nn_mu);
  g_nn.erase(*handle);
  return ELSIM_OK;
}

int elsim_nn_set_params(const int64_t* handle, const float* host) {
  NnNet* n = net_of(handle);
  if (!n || !host) return ELSIM_ERR_ARG;
  ELSIM_HIP(hipMemcpyAsync(n->params.p, host, nn_param_count() * sizeof(float), hipMemcpyHostToDevice, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  n->packed = false;
  return ELSIM_OK;
}

int elsim_nn_get_params(const int64_t* handle, float* host) {
  NnNet* n = net_of(handle);
  if (!n || !host) return ELSIM_ERR_ARG;
  ELSIM_HIP(hipMemcpyAsync(host, n->params.p, nn_param_count() * sizeof(float), hipMemcpyDeviceToHost, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  return ELSIM_OK;
}

int elsim_nn_get_grads(const int64_t* handle, float* host) {
  NnNet* n = net_of(handle);
  if (!n || !host) return ELSIM_ERR_ARG;
  ELSIM_HIP(hipMemcpyAsync(host, n->grads.p, nn_param_count() * sizeof(float), hipMemcpyDeviceToHost, stream()));
  ELSIM_HIP(hipStreamSynchronize(stream()));
  return ELSIM_OK;
}

int elsim_nn_forward(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, float* pred) {
  NnNet* n = net_of(handle);
  if (!n || !x || !pred) return ELSIM_ERR_ARG;
  return nn_forward(n, n1, n2, n3, x, pred, NN_IN_CH);
}

int elsim_nn_forward_used(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, int in_used,
                          float* pred) {
  NnNet* n = net_of(handle);
  if (!n || !x || !pred) return ELSIM_ERR_ARG;
  if (in_used != 8 && in_used != NN_IN_CH) {
    set_error("elsim_nn_forward_used: in_used %d (8 or 16)", in_used);
    return ELSIM_ERR_ARG;
  }
  return nn_forward(n, n1, n2, n3, x, pred, in_used);
}

int elsim_nn_loss_grad(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, const float* target,
                       float* pred, double* loss) {
  NnNet* n = net_of(handle);
  if (!n || !x || !target || !pred) return ELSIM_ERR_ARG;
  ELSIM_TRY(n->ensure(n1, n2, n3, true));
  ELSIM_TRY(n->forward(x, pred, target, loss));
  return n->backward(x);
}

int elsim_nn_train_forward(const int64_t* handle, int n1, int n2, int n3, const uint16_t* x, float* pred) {
  NnNet* n = net_of(handle);
  if (!n || !x || !pred) return ELSIM_ERR_ARG;
  ELSIM_TRY(n->ensure(n1, n2, n3, true));
  return n->forward(x, pred, nullptr, nullptr);
}

int elsim_nn_energy_loss(const int64_t* handle, const float* pred, const float* target, const float* kap,
                         double* loss) {
  NnNet* n = net_of(handle);
  if (!n || !pred || !target || !kap || !n->dpred.p) return ELSIM_ERR_ARG;
  return n->energy_loss(pred, target, kap, loss);
}

int elsim_nn_backward(const int64_t* handle, const uint16_t* x) {
  NnNet* n = net_of(handle);
  if (!n || !x || !n->dpred.p) return ELSIM_ERR_ARG;
  return n->backward(x);
}

int elsim_nn_adam(const int64_t* handle, float lr) {
  NnNet* n = net_of(handle);
  if (!n || !(lr > 0.f)) return ELSIM_ERR_ARG;
  return n->adam(lr);
}

int elsim_nn_conv3(int n1, int n2, int n3, const uint16_t* x, int cin, const uint16_t* wt, const float* bias,
                   int cout, int relu, uint16_t* out) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !x || !wt || !bias || !out || cout > 64) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  return conv<0>(n1, n2, n3, x, cin, View{cin, 0}, wt, bias, relu, cout, out, View{cout, 0}, stream());
}

int elsim_nn_conv3_bwd_data(int n1, int n2, int n3, const uint16_t* dz, int cout, const uint16_t* wb, int cin,
                            float* dx) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !dz || !wb || !dx) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  return conv<1>(n1, n2, n3, dz, cout, View{cout, 0}, wb, nullptr, 0, cin, dx, View{cin, 0}, stream());
}

int elsim_nn_conv3_bwd_weight(int n1, int n2, int n3, const uint16_t* x, int cin, const uint16_t* dz, int cout,
                              float* dw) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !x || !dz || !dw || cin % 16 || cout % 16 ||
      (cin / 16) * (cout / 16) > 4 * WG_MAXP)
    return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const int64_t N = (int64_t)n1 * n2 * n3;
  const int ng = (int)((N + WG_VOX - 1) / WG_VOX);
  float* part = (float*)scratch(0, (size_t)ng * 27 * cin * cout * sizeof(float));
  if (!part) return ELSIM_ERR_ALLOC;
  k_wgrad<<<dim3(ng, 27), 256, 0, stream()>>>(n1, n2, n3, x, cin, View{cin, 0}, dz, cout, part);
  k_wgrad_reduce<<<nblk((int64_t)27 * cin * cout), 256, 0, stream()>>>(ng, cin, cout, part, dw);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

}  // extern "C"
