// Host-side checks of the C-ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer (the library built with both on its host code):
// every entry point that runs without a device -- the geometry generators,
// the property tables, the time-control law, parameter defaults -- and the
// argument and handle checks that return before any device call.  Prints
// one line per check; exits 1 on the first unexpected return code.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "elsim.h"

static int g_fail = 0;
static void expect(const char* what, int got, int want) {
  std::printf("%-44s rc %d%s\n", what, got, got == want ? "" : "  <-- unexpected");
  if (got != want) g_fail = 1;
}

int main() {
  elsim_sbm_params prm;
  elsim_sbm_default_params(&prm);
  std::printf("default params: dx %g dt %g rtol %g\n", prm.dx, prm.dt, prm.rtol);

  // generators and tables
  const int n1 = 24, n2 = 20, n3 = 18;
  std::vector<float> sdf((size_t)n1 * n2 * n3);
  int ns = 0;
  double frac = 0.0;
  expect("gen_sphere_pack", elsim_gen_sphere_pack(n1, n2, n3, 7, 3.0, 6.0, 0.5, 0.85, 1, 12.0, sdf.data(), &ns, &frac),
         ELSIM_OK);
  std::printf("  %d spheres, solid %.3f\n", ns, frac);
  expect("gen_ellipsoid_pack",
         elsim_gen_ellipsoid_pack(n1, n2, n3, 3, 3.0, 6.0, 15.0, 0.5, 0.3, 12.0, sdf.data(), &ns, &frac), ELSIM_OK);
  expect("gen_tunnel", elsim_gen_tunnel(n1, n2, n3, 3, 2.0, sdf.data()), ELSIM_OK);
  std::vector<double> tab(8 * 101);
  expect("gen_tables", elsim_gen_tables(101, tab.data()), ELSIM_OK);
  expect("gen_sphere_pack, zero size", elsim_gen_sphere_pack(0, n2, n3, 7, 3.0, 6.0, 0.5, 0.85, 1, 12.0, sdf.data(),
                                                             &ns, &frac), ELSIM_ERR_ARG);

  // the time-control law over its branches
  const double trgi = -4.0, vsr0 = 0.1;
  for (double xrf : {0.01, 0.5, 0.97})
    for (double scrnt : {-8.0, -4.0, -1.0, 0.5}) {
      double dt = 0.0, vsr = 0.0;
      const int rc = elsim_sbm_time_control(&xrf, &scrnt, &trgi, &vsr0, &dt, &vsr);
      if (rc != ELSIM_OK || !std::isfinite(dt) || !std::isfinite(vsr)) {
        std::printf("time_control xrf %g scrnt %g: rc %d dt %g vsr %g  <-- unexpected\n", xrf, scrnt, rc, dt, vsr);
        g_fail = 1;
      }
    }
  std::printf("time_control: 12 cases\n");

  // argument errors of a potential solve, before any device call
  const int m1 = 8, m2 = 8, m3 = 8;
  std::vector<double> kap((m1 + 2) * (m2 + 2) * (m3 + 2), 1.0), phi(kap.size(), 0.0), rhs(m1 * m2 * m3, 0.0);
  double dx = 1.0, rtol = 1e-8, rr = 0.0, bad_dx = -1.0, bad_rtol = 0.0;
  int lo = 0, hi = 1, mode = ELSIM_MODE_FP64, maxit = 10, it = 0, zero = 0, seven = 7, two = 2;
  expect("solve_pot n1 = 0", elsim_solve_pot(&zero, &m2, &m3, &dx, kap.data(), rhs.data(), phi.data(), &lo, &hi,
                                             &mode, &rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot both faces Neumann", elsim_solve_pot(&m1, &m2, &m3, &dx, kap.data(), rhs.data(), phi.data(), &lo,
                                                         &lo, &mode, &rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot face condition 2", elsim_solve_pot(&m1, &m2, &m3, &dx, kap.data(), rhs.data(), phi.data(), &two,
                                                       &hi, &mode, &rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot mode 7", elsim_solve_pot(&m1, &m2, &m3, &dx, kap.data(), rhs.data(), phi.data(), &lo, &hi, &seven,
                                             &rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot rtol 0", elsim_solve_pot(&m1, &m2, &m3, &dx, kap.data(), rhs.data(), phi.data(), &lo, &hi, &mode,
                                             &bad_rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot maxit 0", elsim_solve_pot(&m1, &m2, &m3, &dx, kap.data(), rhs.data(), phi.data(), &lo, &hi, &mode,
                                              &rtol, &zero, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot dx < 0", elsim_solve_pot(&m1, &m2, &m3, &bad_dx, kap.data(), rhs.data(), phi.data(), &lo, &hi,
                                             &mode, &rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  expect("solve_pot null kap", elsim_solve_pot(&m1, &m2, &m3, &dx, nullptr, rhs.data(), phi.data(), &lo, &hi, &mode,
                                               &rtol, &maxit, &it, &rr), ELSIM_ERR_ARG);
  std::printf("  last error: %s\n", elsim_last_error());

  // unknown handles are looked up, never dereferenced
  const int64_t bogus = 987654321, as_ptr = (int64_t)(intptr_t)&prm;
  elsim_step_stats st;
  int mm = ELSIM_MODE_MIXED;
  expect("sbm_step, unknown handle", elsim_sbm_step(&bogus, &mm, &st), ELSIM_ERR_ARG);
  expect("sbm_destroy, unknown handle", elsim_sbm_destroy(&bogus), ELSIM_ERR_ARG);
  expect("sbm_set_warm_start, unknown handle", elsim_sbm_set_warm_start(&bogus, nullptr), ELSIM_ERR_ARG);
  expect("nn_destroy, unknown handle", elsim_nn_destroy(&bogus), ELSIM_ERR_ARG);
  std::vector<float> params(16);
  expect("nn_get_params, pointer-like handle", elsim_nn_get_params(&as_ptr, params.data()), ELSIM_ERR_ARG);
  const char* msg = elsim_last_error();
  std::printf("  last error: %s\n", msg);
  if (!std::strstr(msg, "unknown network handle")) {
    std::printf("  <-- the handle error was not reported\n");
    g_fail = 1;
  }
  std::printf(g_fail ? "FAILED\n" : "ALL OK\n");
  return g_fail;
}
