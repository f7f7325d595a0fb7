// elsim_runtime.hip -- device selection, the library stream, error state,
// device memory and timers of libelsim_hip.so.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "elsim_solver.h"

namespace elsim {

namespace {
thread_local char g_err[1024] = "";
std::mutex g_mu;
bool g_init = false;
int g_device = 0;
hipStream_t g_stream = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
Workspace g_ws;
std::vector<DevBuf> g_scratch(8);
}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

hipStream_t stream() { return g_stream; }
bool initialized() { return g_init; }
Workspace& workspace() { return g_ws; }

int ensure_init() {
  if (g_init) return ELSIM_OK;
  return elsim_init(g_device);
}

int dev_alloc(DevBuf& b, size_t bytes) {
  ELSIM_TRY(ensure_init());
  if (b.p && b.bytes >= bytes) return ELSIM_OK;
  if (b.p) ELSIM_HIP(hipFree(b.p));
  b.p = nullptr;
  b.bytes = 0;
  if (bytes == 0) return ELSIM_OK;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed", bytes);
    b.p = nullptr;
    return ELSIM_ERR_ALLOC;
  }
  b.bytes = bytes;
  return ELSIM_OK;
}

void dev_free(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

void* scratch(int slot, size_t bytes) {
  if (slot < 0 || slot >= (int)g_scratch.size()) return nullptr;
  if (dev_alloc(g_scratch[slot], bytes) != ELSIM_OK) return nullptr;
  return g_scratch[slot].p;
}

}  // namespace elsim

using namespace elsim;

extern "C" {

int elsim_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) return ELSIM_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_error("no HIP device available");
    return ELSIM_ERR_HIP;
  }
  if (device < 0 || device >= n) {
    set_error("device %d out of range (%d devices)", device, n);
    return ELSIM_ERR_ARG;
  }
  ELSIM_HIP(hipSetDevice(device));
  g_device = device;
  ELSIM_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  ELSIM_HIP(hipEventCreate(&g_ev0));
  ELSIM_HIP(hipEventCreate(&g_ev1));
  ELSIM_HIP(hipMalloc(&g_ws.partials, sizeof(double) * kMaxRedBlocks * kMaxRedVals));
  ELSIM_HIP(hipMalloc(&g_ws.counter, sizeof(unsigned) * 4));
  ELSIM_HIP(hipMemset(g_ws.counter, 0, sizeof(unsigned) * 4));
  ELSIM_TRY(noalloc_init(device));
  ELSIM_HIP(hipDeviceSynchronize());
  g_init = true;
  return ELSIM_OK;
}

const char* elsim_last_error(void) { return g_err; }

int elsim_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int elsim_synchronize(void) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipStreamSynchronize(g_stream));
  return ELSIM_OK;
}

int elsim_malloc(void** ptr, size_t bytes) {
  if (!ptr) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  *ptr = nullptr;
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed", bytes);
    *ptr = nullptr;
    return ELSIM_ERR_ALLOC;
  }
  return ELSIM_OK;
}

int elsim_free(void* ptr) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipStreamSynchronize(g_stream));
  ELSIM_HIP(hipFree(ptr));
  return ELSIM_OK;
}

int elsim_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  ELSIM_HIP(hipStreamSynchronize(g_stream));
  return ELSIM_OK;
}

int elsim_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  ELSIM_HIP(hipStreamSynchronize(g_stream));
  return ELSIM_OK;
}

int elsim_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
  return ELSIM_OK;
}

int elsim_memset(void* dst, int value, size_t bytes) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipMemsetAsync(dst, value, bytes, g_stream));
  return ELSIM_OK;
}

int elsim_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return ELSIM_ERR_ARG;
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return ELSIM_OK;
}

int elsim_timer_start(void) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipEventRecord(g_ev0, g_stream));
  return ELSIM_OK;
}

int elsim_timer_stop(float* ms) {
  ELSIM_TRY(ensure_init());
  ELSIM_HIP(hipEventRecord(g_ev1, g_stream));
  ELSIM_HIP(hipEventSynchronize(g_ev1));
  float t = 0.f;
  ELSIM_HIP(hipEventElapsedTime(&t, g_ev0, g_ev1));
  if (ms) *ms = t;
  return ELSIM_OK;
}

// Reference literal values with REAL(4) kinds resolved
// (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:88-136, 247, 258, 287, 324, 341, 430, 474, 617, 621).
void elsim_sbm_default_params(elsim_sbm_params* p) {
  std::memset(p, 0, sizeof(*p));
  const double cst1 = 1.6021766e-19 / (1.3806488e-23 * (double)300.0f);  // SBM:101
  const double de = 2.0 * (4.0e-6 * 1.25e-6) / (4.0e-6 + 1.25e-6);       // SBM:118
  const double tm = 4.0e-6 / (4.0e-6 + 1.25e-6);                          // SBM:121
  p->dx = 3.25e-5 * (double)0.5f;                                         // SBM:90
  p->dt = (double)(1.05625e-02f * 0.125f * 0.75f);                        // SBM:91
  p->zeta = (double)(1.0f * 0.5f);                                        // SBM:94
  p->psp_off = 1.0e-7;
  p->pse_off = 1.0e-6;
  p->avp_min = 1.0e-2;
  p->eps = 6.76e-10;                                                      // SBM:136
  p->rho = (double)0.0312f;                                               // SBM:100
  p->ch_psi_min = 1.0e-5;
  p->frd = (double)96485.3365f;                                           // SBM:102
  p->cst1 = cst1;
  p->t_minus = tm;
  p->d0 = (de / 2.598606299407882e-06) * (double)0.006667f;               // SBM:124
  p->tc1 = (2 * tm - (double)1.0f) / (2 * tm * ((double)1.0f - tm));      // SBM:125
  p->tc2 = (double)1.0f / (2 * tm * ((double)1.0f - tm)) * cst1;          // SBM:126
  p->dab_c0 = (double)(-7.02f);
  p->dab_c1 = 0.83e3;
  p->dab_c2 = (double)0.05f;
  p->dab_c3 = 1.0e3;
  p->kap_scale = (double)3.3f;
  p->i0_scale = 1.0e-3;
  p->kfw_cref = (double)0.001f;
  p->bvp = (double)(-0.1f);
  p->bve = (double)(-0.4495f);
  p->vsr = 1.0e-2 * 1.0e1 / (double)(3.25f * 3.25f) * (double)2.5f;       // SBM:112
  p->trgi = 0.0;
  p->cne_target = 1.0e-3;
  p->rtol = 1.0e-8;
  p->crt_e = 1.0e-7;
  p->crt_p = 1.0e-7;
  p->crt_l = 1.0e-6;
  p->outer_tol = 1.0e-8;
  p->inner_rtol = 0.5;  // an FP64 reliable update each time the FP32 residual halves
  p->mass_corr_period = 100;
  p->skip_cne_first = 1;
  p->stop_rule = ELSIM_STOP_REFERENCE;
  p->max_it_e = 50000;
  p->max_it_p = 30000;
  p->max_it_l = 100000;
  p->max_outer = 1000;
}

// The time-loop control of the SBM program (SBM:358-383) for the next step:
// the voltage scan rate in five tiers of the current error, the time step in
// three tiers of the lithiation fraction, each tier scaling the rate again.
// Literals are default REAL in the reference: comparisons use their REAL(4)
// values and dt is the REAL(4) product, promoted.
int elsim_sbm_time_control(const double* xrf, const double* scrnt, const double* trgi, const double* vsr0,
                           double* dt, double* vsr) {
  if (!xrf || !scrnt || !trgi || !vsr0 || !dt || !vsr) return ELSIM_ERR_ARG;
  const double dc = std::fabs(*trgi - *scrnt), at = std::fabs(*trgi);
  double v;
  if (dc < (double)0.07f * at)
    v = *vsr0 * 0.5;
  else if (dc < (double)0.20f * at)
    v = *vsr0;
  else if (dc < (double)0.40f * at)
    v = *vsr0 * 3.0;
  else if (dc < (double)0.75f * at)
    v = *vsr0 * 5.0;
  else
    v = *vsr0 * 8.0;
  const float base = 1.05625e-02f * 0.125f;
  if (*xrf < (double)0.12f) {
    *dt = (double)(base * 0.75f);
    v = v * 2;
  } else if (*xrf < (double)0.20f) {
    *dt = (double)(base * 0.5f);
    v = v * 10;
  } else {
    *dt = (double)(base * 0.5f);
    v = v * 15;
  }
  *vsr = v;
  return ELSIM_OK;
}

}  // extern "C"
