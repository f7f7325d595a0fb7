// elsim_levelset.hip -- level-set reinitialisation on the GPU:
// NoGrdBc3D (LevelSet_3D.f90:164-190), AbsGrd3DUW_XYZ (LevelSet_3D.f90:265-322)
// and ReDist3D_XYZ (LevelSet_3D.f90:194-261; RM_UCNCT_3D.f90:209-276).
//
// One sweep is one kernel over the ghosted box: a cell computes the Godunov
// upwind |grad f| of its clamped interior cell (the reference evaluates the
// gradient on the interior and then mirrors it into the ghosts with
// NoGrdBc3D, so a ghost uses its mirror's gradient) and applies the update
// with the sign-flip guard to itself, ghosts included, exactly as the
// reference loop over 0..p+1 does.  f ping-pongs between two buffers; the
// reference's extra Gf and nf arrays and the nf->f copy disappear.
// FP64 (bit-identical to the reference) and FP32 (f in FP32, sign as int8:
// 9 bytes per voxel per sweep) variants.
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

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_sweep(Dims d, const S* __restrict__ sg, const T* __restrict__ f,
                                               T* __restrict__ nf, T dt) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i > d.n1 + 1 || j > d.n2 + 1) return;
  const int k0 = blockIdx.z * 8, k1 = min(k0 + 7, d.n3 + 1);
  const int ic = min(max(i, 1), d.n1), jc = min(max(j, 1), d.n2);
  const T tiny = T(1.0e-7);
  for (int k = k0; k <= k1; k++) {
    const int kc = min(max(k, 1), d.n3);
    const T g = grad_at<T, S>(d, f, sg, d.gi(ic, jc, kc));
    const int64_t q = d.gi(i, j, k);
    const T fo = f[q], s = sgn_val(sg, q);
    T v = fo + s * (T(1) - g) * dt;
    if (!(v * fo > T(0))) v = fo + dt * tiny * s;
    nf[q] = v;
  }
}

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

dim3 sweep_grid(const Dims& d) { return dim3((d.n1 + 2 + 63) / 64, (d.n2 + 2 + 3) / 4, (d.n3 + 2 + 7) / 8); }

template <typename T, typename S>
int redist(int lit, const Dims& d, const S* sg, T* f, T dt) {
  ELSIM_TRY(ensure_init());
  const size_t G = (size_t)d.G();
  T* tmp = (T*)scratch(1, G * sizeof(T));
  if (!tmp) return ELSIM_ERR_ALLOC;
  hipStream_t s = stream();
  T* a = f;
  T* b = tmp;
  for (int m = 0; m < lit; m++) {
    k_sweep<T, S><<<sweep_grid(d), dim3(64, 4, 1), 0, s>>>(d, sg, a, b, dt);
    T* t = a;
    a = b;
    b = t;
  }
  ELSIM_CHECK_LAUNCH();
  if (a != f) ELSIM_HIP(hipMemcpyAsync(f, a, G * sizeof(T), hipMemcpyDeviceToDevice, s));
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

}  // extern "C"
