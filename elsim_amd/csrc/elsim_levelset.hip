// elsim_levelset.hip -- level-set reinitialisation on the GPU:
// NoGrdBc3D (LevelSet_3D.f90:164-190), AbsGrd3DUW_XYZ (LevelSet_3D.f90:265-322)
// and ReDist3D_XYZ (LevelSet_3D.f90:194-261; RM_UCNCT_3D.f90:209-276).
//
// One sweep is one kernel: every interior cell computes the Godunov upwind
// |grad f| (the reference evaluates the gradient on the interior and then
// mirrors it into the ghosts with NoGrdBc3D, so a ghost uses its mirror's
// gradient) and applies the update with the sign-flip guard to itself and to
// the ghost cells that mirror it, each with its own f and sign, exactly as
// the reference loop over 0..p+1 does.  f ping-pongs between two buffers; the
// reference's extra Gf and nf arrays and the nf->f copy disappear.
// FP64 (bit-identical to the reference) and FP32 (f in FP32, sign as int8:
// 9 bytes per voxel per sweep) variants.  Also the geometry front-end that
// prepares the sign function (LevelSet_3D.f90:84-144).
#include "elsim_common.h"

namespace elsim {
namespace {

template <typename T>
__device__ __forceinline__ T godunov(T c, T ag_m, T bg_p, T cg_m, T dg_p, T eg_m, T fg_p, bool pos) {
  // ag = c - f(i-1), bg = f(i+1) - c, ...
  const T z = T(0);
  const T ag = c - ag_m, bg = bg_p - c, cg = c - cg_m, dg = dg_p - c, eg = c - eg_m, fg = fg_p - c;
  T a1, a2, b1, b2, c1, c2;
  if (pos) {
    a1 = fmax(ag, z); a2 = fmax(-bg, z);
    b1 = fmax(cg, z); b2 = fmax(-dg, z);
    c1 = fmax(eg, z); c2 = fmax(-fg, z);
  } else {
    a1 = fmax(-ag, z); a2 = fmax(bg, z);
    b1 = fmax(-cg, z); b2 = fmax(dg, z);
    c1 = fmax(-eg, z); c2 = fmax(fg, z);
  }
  return sqrt(fmax(a1 * a1, a2 * a2) + fmax(b1 * b1, b2 * b2) + fmax(c1 * c1, c2 * c2));
}

__device__ __forceinline__ double sgn_val(const double* sg, int64_t q) { return sg[q]; }
__device__ __forceinline__ float sgn_val(const int8_t* sg, int64_t q) { return (float)sg[q]; }

// |grad f| at interior cell c (ghosted index).
template <typename T, typename S>
__device__ __forceinline__ T grad_at(const Dims& d, const T* __restrict__ f, const S* __restrict__ sg, int64_t c) {
  const T s = sgn_val(sg, c);
  if (s == T(0)) return T(0);
  const int64_t sx = d.g1(), sz = d.g12();
  return godunov<T>(f[c], f[c - 1], f[c + 1], f[c - sx], f[c + sx], f[c - sz], f[c + sz], s > T(0));
}

// The index i itself (first) and the ghost indices that mirror it in a
// dimension of n interior cells (0 if i == 1, n + 1 if i == n).
__device__ __forceinline__ int mirror_set(int i, int n, int (&out)[3]) {
  int c = 0;
  out[c++] = i;
  if (i == 1) out[c++] = 0;
  if (i == n) out[c++] = n + 1;
  return c;
}

// One sweep on the register-blocked tile of the CG operator (256 d1 voxels x
// SJ rows x SKC planes per block): each thread keeps its rows of f for the
// planes k-1, k, k+1 in registers, takes its d1 neighbours from the adjacent
// lanes (DPP) and loads only the rows above and below the tile and the
// wave-edge columns.  It updates its interior cell and the ghost cells across
// the d2 and d3 faces whose clamped interior cell it is (k_sweep_ghosts does
// the d1 faces, where the ghost work would split the lanes of a wave), each
// ghost with its own f and sign and the gradient of that interior cell -- the
// reference's rule for the ghost layer (LS:236-250 over 0..p+1).  Same
// arithmetic as grad_at.
template <typename T>
__device__ __forceinline__ T sweep_update(T fo, T s, T g, T dt) {
  T v = fo + s * (T(1) - g) * dt;
  if (!(v * fo > T(0))) v = fo + dt * T(1.0e-7) * s;
  return v;
}
// rows per thread of the sweep tile (512^3, us per sweep: FP64 948 with 2
// rows, 931 with 3, 941 with 4; FP32 423 with 4, 421 with 6, 445 with 8)
template <typename T>
struct SweepRows;
template <>
struct SweepRows<float> {
  static constexpr int SJ = 4;
};
template <>
struct SweepRows<double> {
  static constexpr int SJ = 3;
};
template <typename T>
inline int sweep_ntiles(const Dims& d) {
  const TileGrid t = stencil_tiles(d, SweepRows<T>::SJ);
  return t.gx * t.gy * t.gz;
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_sweep_rb(Dims d, const S* __restrict__ sg, const T* __restrict__ f,
                                                  T* __restrict__ nf, T dt) {
  constexpr int SJ = SweepRows<T>::SJ;
  const StIdx st = stencil_idx<SJ>(d, blockIdx.x);
  const int64_t sz = d.g12();
  const int i = st.i;
  const bool col_ok = i <= d.n1;
  int64_t rc[SJ];
#pragma unroll
  for (int t = 0; t < SJ; t++) rc[t] = d.gi(st.ic, min(st.j0 + t, d.n2 + 1), st.k0);
  const int64_t clo = d.gi(st.ic, min(st.j0 - 1, d.n2 + 1), st.k0),
                chi = d.gi(st.ic, min(st.j0 + SJ, d.n2 + 1), st.k0);
  const bool ld_left = st.lane == 0 && col_ok, ld_right = st.lane == 63 && col_ok;
  T fc[SJ], fm[SJ];
#pragma unroll
  for (int t = 0; t < SJ; t++) {
    fc[t] = f[rc[t]];
    fm[t] = f[rc[t] - sz];
  }
  int64_t off = 0;
  for (int k = st.k0; k <= st.k1; k++, off += sz) {
    T fp[SJ], sv[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      fp[t] = f[rc[t] + off + sz];
      sv[t] = sgn_val(sg, rc[t] + off);
    }
    const T fjlo = f[clo + off], fjhi = f[chi + off];
    T fl[SJ], fr[SJ];
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      fl[t] = lane_shift<true>(fc[t]);
      fr[t] = lane_shift<false>(fc[t]);
    }
    if (ld_left) {
#pragma unroll
      for (int t = 0; t < SJ; t++) fl[t] = f[rc[t] + off - 1];
    }
    if (ld_right) {
#pragma unroll
      for (int t = 0; t < SJ; t++) fr[t] = f[rc[t] + off + 1];
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      const int j = st.j0 + t;
      if (!col_ok || j > d.n2) continue;
      const T x = fc[t], s = sv[t];
      const T fjm = t == 0 ? fjlo : fc[t - 1];
      const T fjp = t == SJ - 1 ? fjhi : fc[t + 1];
      const T g = s == T(0) ? T(0) : godunov<T>(x, fl[t], fr[t], fjm, fjp, fm[t], fp[t], s > T(0));
      const int64_t c = rc[t] + off;
      nf[c] = sweep_update<T>(x, s, g, dt);
      // ghosts across d2 / d3 faces (uniform per row and plane); those across
      // the d1 faces are k_sweep_ghosts'
      if (j == 1 || j == d.n2 || k == 1 || k == d.n3) {
        int J[3], K[3];
        const int nj = mirror_set(j, d.n2, J), nk = mirror_set(k, d.n3, K);
        for (int cc = 0; cc < nk; cc++)
          for (int b = 0; b < nj; b++) {
            if ((b | cc) == 0) continue;
            const int64_t q = d.gi(i, J[b], K[cc]);
            nf[q] = sweep_update<T>(f[q], sgn_val(sg, q), g, dt);
          }
      }
    }
#pragma unroll
    for (int t = 0; t < SJ; t++) {
      fm[t] = fc[t];
      fc[t] = fp[t];
    }
  }
}

// The ghost cells across the two d1 faces of one sweep: one thread per
// interior cell with i = 1 or i = n1 recomputes that cell's gradient (same
// arithmetic as the tile kernel) and updates the ghost cells (0 or n1+1, J,
// K) that mirror it, each with its own f and sign; k_sweep_rb updates the
// rest of the ghost layer.  Reads the same input f; writes ghosts only.
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_sweep_ghosts(Dims d, const S* __restrict__ sg, const T* __restrict__ f,
                                                      T* __restrict__ nf, T dt) {
  const int64_t nf1 = d.n1 > 1 ? 2 : 1, face = (int64_t)d.n2 * d.n3;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf1 * face) return;
  const int i = q < face ? 1 : d.n1;
  q %= face;
  const int j = (int)(q % d.n2) + 1, k = (int)(q / d.n2) + 1;
  const T g = grad_at<T, S>(d, f, sg, d.gi(i, j, k));
  int I[3], J[3], K[3];
  const int ni = mirror_set(i, d.n1, I), nj = mirror_set(j, d.n2, J), nk = mirror_set(k, d.n3, K);
  for (int cc = 0; cc < nk; cc++)
    for (int b = 0; b < nj; b++)
      for (int a = 1; a < ni; a++) {
        const int64_t qq = d.gi(I[a], J[b], K[cc]);
        nf[qq] = sweep_update<T>(f[qq], sgn_val(sg, qq), g, dt);
      }
}
inline int64_t d1_face_cells(const Dims& d) { return (d.n1 > 1 ? 2 : 1) * (int64_t)d.n2 * d.n3; }

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_absgrd(Dims d, const S* __restrict__ sg, const T* __restrict__ f,
                                                T* __restrict__ gr) {
  const int i = blockIdx.x * 64 + threadIdx.x + 1, j = blockIdx.y * 4 + threadIdx.y + 1;
  if (i > d.n1 || j > d.n2) return;
  const int k0 = blockIdx.z * 8 + 1, k1 = min(k0 + 7, d.n3);
  for (int k = k0; k <= k1; k++) gr[d.ni(i, j, k)] = grad_at<T, S>(d, f, sg, d.gi(i, j, k));
}

// every ghost = value of its clamped interior cell (NoGrdBc3D result)
__global__ void k_mirror(Dims d, double* __restrict__ f) {
  const int64_t G = d.G();
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < G; q += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % d.g1()), j = (int)((q / d.g1()) % (d.n2 + 2)), k = (int)(q / d.g12());
    const int ic = min(max(i, 1), d.n1), jc = min(max(j, 1), d.n2), kc = min(max(k, 1), d.n3);
    if (ic != i || jc != j || kc != k) f[q] = f[d.gi(ic, jc, kc)];
  }
}

template <typename T, typename S>
int redist(int lit, const Dims& d, const S* sg, T* f, T dt) {
  ELSIM_TRY(ensure_init());
  const size_t G = (size_t)d.G();
  T* tmp = (T*)scratch(1, G * sizeof(T));
  if (!tmp) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  T* a = f;
  T* b = tmp;
  const int nt = sweep_ntiles<T>(d);
  for (int m = 0; m < lit; m++) {
    k_sweep_rb<T, S><<<nt, 256, 0, s>>>(d, sg, a, b, dt);
    k_sweep_ghosts<T, S><<<(int)((d1_face_cells(d) + 255) / 256), 256, 0, s>>>(d, sg, a, b, dt);
    T* t = a;
    a = b;
    b = t;
  }
  ELSIM_CHECK_LAUNCH();
  if (a != f) ELSIM_HIP(hipMemcpyAsync(f, a, G * sizeof(T), hipMemcpyDeviceToDevice, s));
  return ELSIM_OK;
}

// ------------------------------------------------- geometry front-end --
// LevelSet_3D's preparation of the sign function (LS:84-144), one frame:
// reorient the voxel image TmpA(yo,xo,zo) (LS:84-87), zero two cylindrical
// tunnels (LS:89-108), stack the thickness copies behind a separator of sep
// voxels (LS:76, 110-121), NoGrdBc3D, n_ac Allen-Cahn iterations
// (LS:123-135), sign function (LS:138-143).

struct FrontGeom {
  int yo, xo, zo, sep;
  double tr;
};

// Writes v into the interior cell (i,j,k) and into every ghost cell whose
// clamped interior cell it is: the state NoGrdBc3D leaves (LS:164-190).
__device__ __forceinline__ void put_mirrored(const Dims& d, double* __restrict__ f, int i, int j, int k, double v) {
  int I[3], J[3], K[3];
  const int ni = mirror_set(i, d.n1, I), nj = mirror_set(j, d.n2, J), nk = mirror_set(k, d.n3, K);
  for (int c = 0; c < nk; c++)
    for (int b = 0; b < nj; b++)
      for (int a = 0; a < ni; a++) f[d.gi(I[a], J[b], K[c])] = v;
}

// Inside one of the two tunnels of the loop's cell (i, k) (LS:94-105), with
// the reference's axes at (1, pz) and (py, 1).
__device__ __forceinline__ bool in_tunnel(const Dims& d, int i, int k, double tr) {
  const double a = (double)i - 1.0, b = (double)k - (double)d.n3;
  if (sqrt(a * a + b * b) < tr) return true;
  const double c = (double)i - (double)d.n1, e = (double)k - 1.0;
  return sqrt(c * c + e * e) < tr;
}

// dsF(i,j,k) with its mirrored ghosts.  The tunnel loop (LS:91-108) runs
// j over 1..px although TmpB's second extent is zo, so it zeroes the element
// at the column-major offset (i-1) + yo*(j-1) + yo*zo*(k-1); the TmpB element
// (i, jb, kb) is therefore zero if any loop cell (j0, k0) with
// (j0-1) + zo*(k0-1) = (jb-1) + zo*(kb-1) lies in a tunnel.
__global__ __launch_bounds__(256) void k_front_build(Dims d, FrontGeom g, const double* __restrict__ tmpa,
                                                     double* __restrict__ dsf) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.N()) return;
  const int i = (int)(q % d.n1) + 1, j = (int)((q / d.n1) % d.n2) + 1, k = (int)(q / ((int64_t)d.n1 * d.n2)) + 1;
  double v = 0.0;
  if (j > g.sep) {
    const int r = (j - g.sep - 1) % g.zo, blk = (j - g.sep - 1) / g.zo;
    const int jb = (blk & 1) ? g.zo - r : r + 1;
    // TmpB(i, jb, k) = TmpA(i, k, zo - jb + 1) (LS:86)
    v = tmpa[(int64_t)(i - 1) + (int64_t)g.yo * ((int64_t)(k - 1) + (int64_t)g.xo * (g.zo - jb))];
    const int64_t m = (int64_t)(jb - 1) + (int64_t)g.zo * (k - 1);
    const int64_t lo = m - (d.n2 - 1) > 0 ? (m - (d.n2 - 1) + g.zo - 1) / g.zo : 0;
    const int64_t hi = min((int64_t)(d.n3 - 1), m / g.zo);
    for (int64_t k0 = lo; k0 <= hi; k0++)
      if (in_tunnel(d, i, (int)k0 + 1, g.tr)) {
        v = 0.0;
        break;
      }
  }
  put_mirrored(d, dsf, i, j, k, v);
}

// One Allen-Cahn iteration (LS:127-133) in the reference's evaluation order
// (REAL(4) literals 0.64 and 0.05), new value and its mirrored ghosts to dn.
__global__ __launch_bounds__(256) void k_allen_cahn(Dims d, const double* __restrict__ f, double* __restrict__ dn) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.N()) return;
  const int i = (int)(q % d.n1) + 1, j = (int)((q / d.n1) % d.n2) + 1, k = (int)(q / ((int64_t)d.n1 * d.n2)) + 1;
  const int64_t c = d.gi(i, j, k), sx = d.g1(), sz = d.g12();
  const double x = f[c];
  const double lap = f[c + 1] + f[c - 1] + f[c + sx] + f[c - sx] + f[c + sz] + f[c - sz] - 6 * x;
  const double sg = 2 * x * (1.0 - x) * (1.0 - 2 * x) - (double)0.64f * lap;
  put_mirrored(d, dn, i, j, k, x - (double)0.05f * sg);
}

// sgF = +1 where dsF > 0.5, else -1, over the ghosted box (LS:138-143).
__global__ __launch_bounds__(256) void k_front_sign(int64_t G, const double* __restrict__ f, double* __restrict__ sg) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < G) sg[q] = f[q] > 0.5 ? 1.0 : -1.0;
}

// --------------------------------------------------- 2x refinement --
// The step between RM_UCNCT_3D and the SBM program that the reference does
// not contain (its SBM reads FP32 distance functions on a grid twice as fine,
// SBM:138-152, 181-190): fine cell I of 2n lies at coarse coordinate
// (I + 0.5)/2, i.e. the two fine cells of coarse cell c at c -+ 0.25.  The
// value is the trilinear interpolation of the eight surrounding coarse cells
// (weights 1/4 and 3/4, exact in binary; coarse indices clamped to 1..n,
// the Neumann rule of the ghost layer), times 2 (distances in fine voxels),
// rounded to FP32.  Interpolation order: d1, then d2, then d3.
__device__ __forceinline__ void refine_axis(int I, int n, int& c0, int& c1, double& w0, double& w1) {
  // fine I (1-based) -> coarse cells c0 <= c1 and weights
  const int c = (I + 1) / 2;                  // coarse cell containing fine I
  const bool lo = (I & 1) != 0;               // first of the two fine cells: at c - 0.25
  c0 = lo ? c - 1 : c;
  c1 = c0 + 1;
  w0 = lo ? 0.25 : 0.75;
  w1 = lo ? 0.75 : 0.25;
  c0 = min(max(c0, 1), n);
  c1 = min(max(c1, 1), n);
}
__global__ __launch_bounds__(256) void k_refine2(Dims d, const double* __restrict__ sdf, float* __restrict__ out) {
  const int64_t m1 = 2 * (int64_t)d.n1, m2 = 2 * (int64_t)d.n2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m1 * m2 * 2 * (int64_t)d.n3) return;
  const int I = (int)(q % m1) + 1, J = (int)((q / m1) % m2) + 1, K = (int)(q / (m1 * m2)) + 1;
  int i0, i1, j0, j1, k0, k1;
  double a0, a1, b0, b1, c0, c1;
  refine_axis(I, d.n1, i0, i1, a0, a1);
  refine_axis(J, d.n2, j0, j1, b0, b1);
  refine_axis(K, d.n3, k0, k1, c0, c1);
  auto at = [&](int i, int j, int k) { return sdf[d.ni(i, j, k)]; };
  const double x00 = a0 * at(i0, j0, k0) + a1 * at(i1, j0, k0);
  const double x10 = a0 * at(i0, j1, k0) + a1 * at(i1, j1, k0);
  const double x01 = a0 * at(i0, j0, k1) + a1 * at(i1, j0, k1);
  const double x11 = a0 * at(i0, j1, k1) + a1 * at(i1, j1, k1);
  const double y0 = b0 * x00 + b1 * x10, y1 = b0 * x01 + b1 * x11;
  out[q] = (float)(2.0 * (c0 * y0 + c1 * y1));
}

bool front_ok(int yo, int xo, int zo, int py, int px, int pz, int sep) {
  if (yo <= 0 || xo <= 0 || zo <= 0 || !dims_ok(py, px, pz) || py > yo || pz > xo || sep < 0 || px <= sep ||
      (px - sep) % zo != 0)
    return false;
  // the tunnel loop's last offset must stay inside TmpB (yo*zo*xo elements)
  const int64_t last = (int64_t)(py - 1) + (int64_t)yo * ((int64_t)(px - 1) + (int64_t)zo * (pz - 1));
  return last < (int64_t)yo * zo * xo;
}

// Front-end on device arrays: tmpa (yo*xo*zo), sgf and dsf ghosted; the
// smoothing ping-pongs between dsf and sgf, the sign is written last.
int front_end(const Dims& d, const FrontGeom& g, int n_ac, const double* tmpa, double* sgf, double* dsf) {
  hipStream_t s = stream();
  const int nb = (int)((d.N() + 255) / 256);
  k_front_build<<<nb, 256, 0, s>>>(d, g, tmpa, dsf);
  double* a = dsf;
  double* b = sgf;
  for (int it = 0; it < n_ac; it++) {
    k_allen_cahn<<<nb, 256, 0, s>>>(d, a, b);
    double* t = a;
    a = b;
    b = t;
  }
  if (a != dsf) ELSIM_HIP(hipMemcpyAsync(dsf, a, (size_t)d.G() * 8, hipMemcpyDeviceToDevice, s));
  k_front_sign<<<(int)((d.G() + 255) / 256), 256, 0, s>>>(d.G(), dsf, sgf);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

}  // namespace
}  // namespace elsim

using namespace elsim;

extern "C" {

int elsim_dev_redist3d_f64(int lit, int py, int px, int pz, const double* sg, double* f, double dt) {
  if (lit < 0 || !dims_ok(py, px, pz) || !sg || !f) return ELSIM_ERR_ARG;
  return redist<double, double>(lit, Dims{py, px, pz}, sg, f, dt);
}

int elsim_dev_redist3d_f32(int lit, int py, int px, int pz, const int8_t* sg, float* f, float dt) {
  if (lit < 0 || !dims_ok(py, px, pz) || !sg || !f) return ELSIM_ERR_ARG;
  return redist<float, int8_t>(lit, Dims{py, px, pz}, sg, f, dt);
}

int elsim_redist3d_xyz(const int* lit, const int* py, const int* px, const int* pz, const double* sg, double* f,
                       const double* dt) {
  if (!lit || !py || !px || !pz || !sg || !f || !dt || *lit < 0 || !dims_ok(*py, *px, *pz)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t G = (size_t)d.G() * 8;
  double* dsg = (double*)scratch(2, G);
  double* df = (double*)scratch(3, G);
  if (!dsg || !df) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(dsg, sg, G, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(df, f, G, hipMemcpyHostToDevice, s));
  ELSIM_TRY((redist<double, double>(*lit, d, dsg, df, *dt)));
  ELSIM_HIP(hipMemcpyAsync(f, df, G, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_absgrd3duw_xyz(const int* py, const int* px, const int* pz, const double* sg, const double* f, double* gr) {
  if (!py || !px || !pz || !sg || !f || !gr || !dims_ok(*py, *px, *pz)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t G = (size_t)d.G() * 8, N = (size_t)d.N() * 8;
  double* dsg = (double*)scratch(2, G);
  double* df = (double*)scratch(3, G);
  double* dg = (double*)scratch(4, N);
  if (!dsg || !df || !dg) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(dsg, sg, G, hipMemcpyHostToDevice, s));
  ELSIM_HIP(hipMemcpyAsync(df, f, G, hipMemcpyHostToDevice, s));
  k_absgrd<double, double><<<dim3((d.n1 + 63) / 64, (d.n2 + 3) / 4, (d.n3 + 7) / 8), dim3(64, 4, 1), 0, s>>>(d, dsg, df,
                                                                                                              dg);
  ELSIM_CHECK_LAUNCH();
  ELSIM_HIP(hipMemcpyAsync(gr, dg, N, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_nogrdbc3d(const int* py, const int* px, const int* pz, double* f) {
  if (!py || !px || !pz || !f || !dims_ok(*py, *px, *pz)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t G = (size_t)d.G() * 8;
  double* df = (double*)scratch(3, G);
  if (!df) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(df, f, G, hipMemcpyHostToDevice, s));
  k_mirror<<<flat_blocks(d.G()), FLAT_T, 0, s>>>(d, df);
  ELSIM_CHECK_LAUNCH();
  ELSIM_HIP(hipMemcpyAsync(f, df, G, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_dev_refine2_sdf(int py, int px, int pz, const double* sdf, float* out) {
  if (!dims_ok(py, px, pz) || !dims_ok(2 * py, 2 * px, 2 * pz) || !sdf || !out) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{py, px, pz};
  const int64_t M = 8 * d.N();
  k_refine2<<<(int)((M + 255) / 256), 256, 0, stream()>>>(d, sdf, out);
  ELSIM_CHECK_LAUNCH();
  return ELSIM_OK;
}

int elsim_refine2_sdf(const int* py, const int* px, const int* pz, const double* sdf, float* out) {
  if (!py || !px || !pz || !sdf || !out || !dims_ok(*py, *px, *pz) || !dims_ok(2 * *py, 2 * *px, 2 * *pz))
    return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t N = (size_t)d.N();
  double* a = (double*)scratch(3, N * 8);
  float* b = (float*)scratch(4, 8 * N * 4);
  if (!a || !b) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(a, sdf, N * 8, hipMemcpyHostToDevice, s));
  ELSIM_TRY(elsim_dev_refine2_sdf(*py, *px, *pz, a, b));
  ELSIM_HIP(hipMemcpyAsync(out, b, 8 * N * 4, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

int elsim_dev_ls_frontend(int yo, int xo, int zo, const double* tmpa, int py, int px, int pz, int sep, double tr,
                          int n_ac, double* sgf, double* dsf) {
  if (!tmpa || !sgf || !dsf || n_ac < 0 || !front_ok(yo, xo, zo, py, px, pz, sep)) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  return front_end(Dims{py, px, pz}, FrontGeom{yo, xo, zo, sep, tr}, n_ac, tmpa, sgf, dsf);
}

int elsim_ls_frontend(const int* yo, const int* xo, const int* zo, const double* tmpa, const int* py, const int* px,
                      const int* pz, const int* sep, const double* tr, const int* n_ac, double* sgf, double* dsf) {
  if (!yo || !xo || !zo || !tmpa || !py || !px || !pz || !sep || !tr || !n_ac || !sgf || !dsf || *n_ac < 0 ||
      !front_ok(*yo, *xo, *zo, *py, *px, *pz, *sep))
    return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  const Dims d{*py, *px, *pz};
  const size_t G = (size_t)d.G() * 8, T = (size_t)*yo * *xo * *zo * 8;
  double* dt = (double*)scratch(2, T);
  double* dsg = (double*)scratch(3, G);
  double* dds = (double*)scratch(4, G);
  if (!dt || !dsg || !dds) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  ELSIM_HIP(hipMemcpyAsync(dt, tmpa, T, hipMemcpyHostToDevice, s));
  ELSIM_TRY(front_end(d, FrontGeom{*yo, *xo, *zo, *sep, *tr}, *n_ac, dt, dsg, dds));
  ELSIM_HIP(hipMemcpyAsync(sgf, dsg, G, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipMemcpyAsync(dsf, dds, G, hipMemcpyDeviceToHost, s));
  ELSIM_HIP(hipStreamSynchronize(s));
  return ELSIM_OK;
}

}  // extern "C"
