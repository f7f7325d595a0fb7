// elsim_noalloc.hip -- the allocation-free variant of the standalone
// potential solve, for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its CG work space -- seven FP64 and seven FP32 vectors of the ghosted grid,
// the reduction partials and the device scalars -- lives in `scratch`, a
// file-scope device array of fixed size: grids up to ELSIM_NOALLOC_MAX_G
// ghosted elements (64^3 interior: 66^3 = 287,496), 28,347,136 bytes that
// every GPU the library runs on holds until the process ends.  Every call on a GPU
// shares that GPU's array: the call clears the vectors it needs zeroed (with
// a kernel, no memset), and two calls that overlap on one GPU race.
#include <map>
#include <mutex>
#include <string>

#include "elsim_solver.h"

namespace {

constexpr int64_t kMaxG = ELSIM_NOALLOC_MAX_G;
// CgWork::bound_bytes of the largest grid: 7 FP64 + 7 FP32 vectors of the even
// length rounded to 256 bytes, two partial slots, the scalars
constexpr size_t kVec64 = (((kMaxG + 1) & ~1ll) * 8 + 255) & ~(size_t)255;
constexpr size_t kVec32 = (((kMaxG + 1) & ~1ll) * 4 + 255) & ~(size_t)255;
constexpr size_t kBytes = 7 * kVec64 + 7 * kVec32 + 2 * (size_t)elsim::kMaxRedBlocks * 8 + 256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kBytes];

namespace {
std::mutex g_mu;
std::map<int, unsigned char*> g_ptr;  // device -> address of that device's `scratch`
std::string g_setup_error;            // why the set-up failed (the rest of the library is unaffected)
elsim::CgDev g_host;                  // host mirror of the scalars (pageable)
elsim::CgWork g_work;
}  // namespace

namespace elsim {

// Once per device, with that device current: the device address of
// `scratch` (in host code the array's name is not a device address).
int noalloc_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr.count(device)) return ELSIM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));  // the array, not elsim::scratch()
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_setup_error = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    return ELSIM_OK;  // only the allocation-free entry point is unavailable; it reports this
  }
  g_ptr[device] = (unsigned char*)p;
  return ELSIM_OK;
}

}  // namespace elsim

extern "C" int elsim_dev_solve_pot_noalloc(int n1, int n2, int n3, double dx, const double* kap, const double* rhs,
                                           double* phi, int bc_lo, int bc_hi, int mode, double rtol, int maxit,
                                           int* iters, double* relres) {
  using namespace elsim;
  ELSIM_TRY(pot_args_check(n1, n2, n3, dx, kap, rhs, phi, bc_lo, bc_hi, mode, rtol, maxit));
  if (mode == ELSIM_MODE_MIXED_MG) {
    set_error("allocation-free solve: the multigrid mode builds its hierarchy in allocated memory (modes 0-2 only)");
    return ELSIM_ERR_ARG;
  }
  const Dims d{n1, n2, n3};
  if (d.G() > kMaxG) {
    set_error("allocation-free solve: %lld ghosted elements exceed the maximum of %lld", (long long)d.G(),
              (long long)kMaxG);
    return ELSIM_ERR_ARG;
  }
  if (!initialized()) {
    set_error("allocation-free solve: call elsim_init first (initialising the library allocates)");
    return ELSIM_ERR_ARG;
  }
  int dev = 0;
  ELSIM_HIP(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ptr.find(dev);
    if (it != g_ptr.end()) base = it->second;
  }
  if (!base) {
    std::lock_guard<std::mutex> lk(g_mu);
    set_error("allocation-free solve: device %d was not set up by elsim_init (%s)", dev, g_setup_error.c_str());
    return ELSIM_ERR_HIP;
  }
  ELSIM_TRY(g_work.bind(d, base, kBytes, &g_host));
  return solve_pot_with(g_work, n1, n2, n3, dx, kap, rhs, phi, bc_lo, bc_hi, mode, rtol, maxit, iters, relres);
}
