// elsim_gen.cpp -- deterministic synthetic microstructures and property tables
// (host code).  The reference ships none of its inputs (the voxel image, the
// refined SDFs and the 101-point tables are absent), so every run here starts
// from these generators.  Same seed -> same bytes on every machine: the PRNG is
// SplitMix64-seeded xoshiro256** and all geometry is computed in double.
//
// Voxel (i,j,k) (1-based, Fortran order) has its centre at (i-0.5, j-0.5,
// k-0.5) in a box [0,n1) x [0,n2) x [0,n3).  Output SDFs are interior FP32
// arrays, column-major with d1 fastest, positive inside the solid.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/elsim.h"

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

int worker_count() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  return (int)std::min(h, 16u);
}

// Run f(k_lo, k_hi) over d3 slabs on worker threads.
template <class F>
void parallel_slabs(int n3, F f) {
  int w = std::min(worker_count(), std::max(1, n3));
  std::vector<std::thread> th;
  for (int t = 0; t < w; t++) {
    int lo = (int)((int64_t)n3 * t / w), hi = (int)((int64_t)n3 * (t + 1) / w);
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

struct Sphere {
  double x, y, z, r;
};

}  // namespace

// 6-connected flood fill over an interior voxel mask from the voxels of the d2
// plane `jface` (1-based); returns reached flags.
static std::vector<uint8_t> flood_from_face(int n1, int n2, int n3, const std::vector<uint8_t>& in, int jface) {
  const size_t N = (size_t)n1 * n2 * n3;
  std::vector<uint8_t> seen(N, 0);
  std::vector<int64_t> q;
  for (int k = 0; k < n3; k++)
    for (int i = 0; i < n1; i++) {
      size_t id = (size_t)i + (size_t)n1 * ((jface - 1) + (size_t)n2 * k);
      if (in[id]) {
        seen[id] = 1;
        q.push_back((int64_t)id);
      }
    }
  for (size_t h = 0; h < q.size(); h++) {
    int64_t id = q[h];
    int i = (int)(id % n1), j = (int)((id / n1) % n2), k = (int)(id / ((int64_t)n1 * n2));
    const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    for (const auto& v : nb) {
      if (v[0] < 0 || v[0] >= n1 || v[1] < 0 || v[1] >= n2 || v[2] < 0 || v[2] >= n3) continue;
      size_t r = (size_t)v[0] + (size_t)n1 * (v[1] + (size_t)n2 * v[2]);
      if (in[r] && !seen[r]) {
        seen[r] = 1;
        q.push_back((int64_t)r);
      }
    }
  }
  return seen;
}

extern "C" int elsim_gen_sphere_pack(int n1, int n2, int n3, uint64_t seed, double rmin, double rmax,
                                     double target, double hardcore, int connect, double dcut,
                                     float* sdf, int* nspheres, double* frac) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !sdf || rmin <= 0 || rmax < rmin) return ELSIM_ERR_ARG;
  Rng rng(seed);
  const double cell = 2.0 * rmax;
  const int c1 = std::max(1, (int)std::ceil(n1 / cell)), c2 = std::max(1, (int)std::ceil(n2 / cell)),
            c3 = std::max(1, (int)std::ceil(n3 / cell));
  std::vector<std::vector<int>> grid((size_t)c1 * c2 * c3);
  std::vector<Sphere> sp;
  const double N = (double)n1 * n2 * n3;
  std::vector<uint8_t> occ((size_t)n1 * n2 * n3, 0);
  int64_t covered = 0;
  int64_t attempts = 0, fails = 0;
  const int64_t max_attempts = 4000000, max_fails = 400000;
  while ((double)covered < target * N && attempts < max_attempts && fails < max_fails) {
    attempts++;
    Sphere s;
    s.r = rng.uniform(rmin, rmax);
    s.x = rng.uniform() * n1;
    s.y = rng.uniform() * n2;
    s.z = rng.uniform() * n3;
    int g1 = std::min(c1 - 1, (int)(s.x / cell)), g2 = std::min(c2 - 1, (int)(s.y / cell)),
        g3 = std::min(c3 - 1, (int)(s.z / cell));
    bool ok = true;
    for (int a = g1 - 1; a <= g1 + 1 && ok; a++)
      for (int b = g2 - 1; b <= g2 + 1 && ok; b++)
        for (int c = g3 - 1; c <= g3 + 1 && ok; c++) {
          if (a < 0 || b < 0 || c < 0 || a >= c1 || b >= c2 || c >= c3) continue;
          for (int id : grid[(size_t)a + (size_t)c1 * (b + (size_t)c2 * c)]) {
            const Sphere& o = sp[id];
            double dx = s.x - o.x, dy = s.y - o.y, dz = s.z - o.z;
            double lim = hardcore * (s.r + o.r);
            if (dx * dx + dy * dy + dz * dz < lim * lim) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) {
      fails++;
      continue;
    }
    fails = 0;
    grid[(size_t)g1 + (size_t)c1 * (g2 + (size_t)c2 * g3)].push_back((int)sp.size());
    sp.push_back(s);
    // voxel centres strictly inside the ball, counted once
    int i0 = std::max(1, (int)std::floor(s.x - s.r)), i1 = std::min(n1, (int)std::ceil(s.x + s.r) + 1);
    int j0 = std::max(1, (int)std::floor(s.y - s.r)), j1 = std::min(n2, (int)std::ceil(s.y + s.r) + 1);
    int k0 = std::max(1, (int)std::floor(s.z - s.r)), k1 = std::min(n3, (int)std::ceil(s.z + s.r) + 1);
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
          double dx = i - 0.5 - s.x, dy = j - 0.5 - s.y, dz = k - 0.5 - s.z;
          size_t id = (size_t)(i - 1) + (size_t)n1 * ((j - 1) + (size_t)n2 * (k - 1));
          if (dx * dx + dy * dy + dz * dz < s.r * s.r && !occ[id]) {
            occ[id] = 1;
            covered++;
          }
        }
  }
  // SDF = max_s (r_s - |x - c_s|), clamped below at -dcut.
  const float lo = (float)(-dcut);
  auto build_sdf = [&]() {
    parallel_slabs(n3, [&](int klo, int khi) {
      for (int k = klo; k < khi; k++)
        for (int j = 0; j < n2; j++)
          for (int i = 0; i < n1; i++) sdf[(size_t)i + (size_t)n1 * (j + (size_t)n2 * k)] = lo;
      for (const Sphere& s : sp) {
        double reach = s.r + dcut;
        int ka = std::max(klo, (int)std::floor(s.z - reach)), kb = std::min(khi - 1, (int)std::ceil(s.z + reach));
        if (ka > kb) continue;
        int ja = std::max(0, (int)std::floor(s.y - reach)), jb = std::min(n2 - 1, (int)std::ceil(s.y + reach));
        int ia = std::max(0, (int)std::floor(s.x - reach)), ib = std::min(n1 - 1, (int)std::ceil(s.x + reach));
        for (int k = ka; k <= kb; k++) {
          double dz = k + 0.5 - s.z;
          for (int j = ja; j <= jb; j++) {
            double dy = j + 0.5 - s.y;
            float* row = sdf + (size_t)n1 * (j + (size_t)n2 * k);
            for (int i = ia; i <= ib; i++) {
              double dx = i + 0.5 - s.x;
              float d = (float)(s.r - std::sqrt(dx * dx + dy * dy + dz * dz));
              if (d > row[i]) row[i] = d;
            }
          }
        }
      }
    });
  };
  build_sdf();
  if (connect) {
    // Keep only particles joined to the current collector (high d2 face) and
    // fill electrolyte pockets cut off from the separator (low d2 face): an
    // isolated particle or pocket has no conduction path, so the coupled
    // potential problem around it has no bounded solution.
    const size_t NN = (size_t)n1 * n2 * n3;
    std::vector<uint8_t> solid(NN);
    for (size_t q = 0; q < NN; q++) solid[q] = sdf[q] > 0.0f;
    std::vector<uint8_t> reach = flood_from_face(n1, n2, n3, solid, n2);
    std::vector<Sphere> kept;
    for (const Sphere& s : sp) {
      int i = std::min(n1 - 1, std::max(0, (int)std::floor(s.x))), j = std::min(n2 - 1, std::max(0, (int)std::floor(s.y))),
          k = std::min(n3 - 1, std::max(0, (int)std::floor(s.z)));
      size_t id = (size_t)i + (size_t)n1 * (j + (size_t)n2 * k);
      if (solid[id] ? reach[id] : true) kept.push_back(s);
    }
    sp.swap(kept);
    build_sdf();
    std::vector<uint8_t> pore(NN);
    for (size_t q = 0; q < NN; q++) pore[q] = !(sdf[q] > 0.0f);
    std::vector<uint8_t> open = flood_from_face(n1, n2, n3, pore, 1);
    for (size_t q = 0; q < NN; q++)
      if (pore[q] && !open[q]) sdf[q] = 0.5f;
  }
  int64_t pos = 0;
  for (size_t q = 0; q < (size_t)n1 * n2 * n3; q++) pos += sdf[q] > 0.0f;
  if (nspheres) *nspheres = (int)sp.size();
  if (frac) *frac = (double)pos / N;
  return ELSIM_OK;
}

extern "C" int elsim_gen_ellipsoid_pack(int n1, int n2, int n3, uint64_t seed, double amin, double amax,
                                        double cone_deg, double target, double overlap, double dcut, float* sdf,
                                        int* nshapes, double* frac) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !sdf || amin <= 0 || amax < amin || !(overlap >= 0.0 && overlap <= 1.0))
    return ELSIM_ERR_ARG;
  struct Ell {
    double c[3], ax[3], R[9];  // rows of R: body axes in world coords
  };
  Rng rng(seed);
  const size_t N = (size_t)n1 * n2 * n3;
  std::vector<uint8_t> occ(N, 0);
  std::vector<Ell> es;
  int64_t covered = 0, attempts = 0, fails = 0;
  const int64_t max_attempts = 2000000, max_fails = 100000;
  const double cone = cone_deg * M_PI / 180.0;
  std::vector<size_t> cand;
  while ((double)covered < target * (double)N && attempts < max_attempts && fails < max_fails) {
    attempts++;
    Ell e;
    double a = rng.uniform(amin, amax), b = rng.uniform(amin, amax), c = rng.uniform(amin, amax);
    double major = std::max(a, std::max(b, c)), m1 = std::min(a, std::min(b, c)), m2 = a + b + c - major - m1;
    e.ax[0] = major;  // along the tilted d2 axis
    e.ax[1] = m2;
    e.ax[2] = m1;
    // major axis: d2 tilted by theta <= cone about a random azimuth; random roll
    double th = std::acos(1.0 - rng.uniform() * (1.0 - std::cos(cone)));
    double ph = rng.uniform() * 2.0 * M_PI, roll = rng.uniform() * 2.0 * M_PI;
    double u[3] = {std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
    // any vector orthogonal to u
    double t[3] = {1.0, 0.0, 0.0};
    if (std::fabs(u[0]) > 0.9) { t[0] = 0.0; t[2] = 1.0; }
    double d = t[0] * u[0] + t[1] * u[1] + t[2] * u[2];
    double v[3] = {t[0] - d * u[0], t[1] - d * u[1], t[2] - d * u[2]};
    double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (double& x : v) x /= vn;
    double w[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    double cr = std::cos(roll), sr = std::sin(roll);
    double v2[3], w2[3];
    for (int q = 0; q < 3; q++) {
      v2[q] = cr * v[q] + sr * w[q];
      w2[q] = -sr * v[q] + cr * w[q];
    }
    for (int q = 0; q < 3; q++) {
      e.R[q] = u[q];
      e.R[3 + q] = v2[q];
      e.R[6 + q] = w2[q];
    }
    e.c[0] = rng.uniform() * n1;
    e.c[1] = rng.uniform() * n2;
    e.c[2] = rng.uniform() * n3;
    // rasterise; reject when more than `overlap` of the shape's voxels are
    // already solid (0: voxel-level non-overlap, 1: free overlap)
    cand.clear();
    bool ok = true;
    int64_t hit = 0;
    int lo[3], hi[3], n[3] = {n1, n2, n3};
    for (int q = 0; q < 3; q++) {
      lo[q] = std::max(1, (int)std::floor(e.c[q] - major));
      hi[q] = std::min(n[q], (int)std::ceil(e.c[q] + major) + 1);
    }
    for (int k = lo[2]; k <= hi[2] && ok; k++)
      for (int j = lo[1]; j <= hi[1] && ok; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          double p[3] = {i - 0.5 - e.c[0], j - 0.5 - e.c[1], k - 0.5 - e.c[2]};
          double r2 = 0.0;
          for (int q = 0; q < 3; q++) {
            double b0 = (e.R[3 * q] * p[0] + e.R[3 * q + 1] * p[1] + e.R[3 * q + 2] * p[2]) / e.ax[q];
            r2 += b0 * b0;
          }
          if (r2 < 1.0) {
            size_t id = (size_t)(i - 1) + (size_t)n1 * ((j - 1) + (size_t)n2 * (k - 1));
            if (occ[id]) {
              hit++;
              if (overlap == 0.0) {
                ok = false;
                break;
              }
            }
            cand.push_back(id);
          }
        }
    if (ok && overlap < 1.0 && (double)hit > overlap * (double)cand.size()) ok = false;
    if (!ok) {
      fails++;
      continue;
    }
    fails = 0;
    for (size_t id : cand) {
      covered += occ[id] == 0;
      occ[id] = 1;
    }
    es.push_back(e);
  }
  // approximate SDF: for each ellipsoid, (1 - rho) * a_min with rho the
  // normalised radius; exact for spheres; max over shapes; clamp at -dcut.
  const float lof = (float)(-dcut);
  parallel_slabs(n3, [&](int klo, int khi) {
    for (int k = klo; k < khi; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) sdf[(size_t)i + (size_t)n1 * (j + (size_t)n2 * k)] = lof;
    for (const Ell& e : es) {
      double reach = e.ax[0] + dcut;
      int ka = std::max(klo, (int)std::floor(e.c[2] - reach)), kb = std::min(khi - 1, (int)std::ceil(e.c[2] + reach));
      if (ka > kb) continue;
      int ja = std::max(0, (int)std::floor(e.c[1] - reach)), jb = std::min(n2 - 1, (int)std::ceil(e.c[1] + reach));
      int ia = std::max(0, (int)std::floor(e.c[0] - reach)), ib = std::min(n1 - 1, (int)std::ceil(e.c[0] + reach));
      double amin_e = e.ax[2];
      for (int k = ka; k <= kb; k++)
        for (int j = ja; j <= jb; j++) {
          float* row = sdf + (size_t)n1 * (j + (size_t)n2 * k);
          for (int i = ia; i <= ib; i++) {
            double p[3] = {i + 0.5 - e.c[0], j + 0.5 - e.c[1], k + 0.5 - e.c[2]};
            double r2 = 0.0;
            for (int q = 0; q < 3; q++) {
              double b0 = (e.R[3 * q] * p[0] + e.R[3 * q + 1] * p[1] + e.R[3 * q + 2] * p[2]) / e.ax[q];
              r2 += b0 * b0;
            }
            float d = (float)((1.0 - std::sqrt(r2)) * amin_e);
            if (d > row[i]) row[i] = d;
          }
        }
    }
  });
  int64_t pos = 0;
  for (size_t q = 0; q < N; q++) pos += sdf[q] > 0.0f;
  if (nshapes) *nshapes = (int)es.size();
  if (frac) *frac = (double)pos / (double)N;
  return ELSIM_OK;
}

extern "C" int elsim_gen_tunnel(int n1, int n2, int n3, uint64_t seed, double radius, float* sdf) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !sdf || radius <= 0) return ELSIM_ERR_ARG;
  Rng rng(seed);
  // biased random walk in d2 from the low to the high face, unit steps
  std::vector<double> px, py, pz;
  double x = rng.uniform(radius, n1 - radius), y = 0.0, z = rng.uniform(radius, n3 - radius);
  while (y < n2) {
    px.push_back(x);
    py.push_back(y);
    pz.push_back(z);
    double a = rng.uniform() * 2.0 * M_PI, c = rng.uniform(-1.0, 1.0);
    double s = std::sqrt(1.0 - c * c);
    double dx = s * std::cos(a), dy = 0.6 + 0.4 * std::fabs(c), dz = s * std::sin(a);
    double nrm = std::sqrt(dx * dx + dy * dy + dz * dz);
    x = std::min(std::max(x + dx / nrm, radius), n1 - radius);
    y += dy / nrm;
    z = std::min(std::max(z + dz / nrm, radius), n3 - radius);
  }
  px.push_back(x);
  py.push_back(y);
  pz.push_back(z);
  const double reach = radius + 8.0;
  parallel_slabs(n3, [&](int klo, int khi) {
    for (size_t q = 0; q < px.size(); q++) {
      int ka = std::max(klo, (int)std::floor(pz[q] - reach)), kb = std::min(khi - 1, (int)std::ceil(pz[q] + reach));
      if (ka > kb) continue;
      int ja = std::max(0, (int)std::floor(py[q] - reach)), jb = std::min(n2 - 1, (int)std::ceil(py[q] + reach));
      int ia = std::max(0, (int)std::floor(px[q] - reach)), ib = std::min(n1 - 1, (int)std::ceil(px[q] + reach));
      for (int k = ka; k <= kb; k++)
        for (int j = ja; j <= jb; j++) {
          float* row = sdf + (size_t)n1 * (j + (size_t)n2 * k);
          for (int i = ia; i <= ib; i++) {
            double dx = i + 0.5 - px[q], dy = j + 0.5 - py[q], dz = k + 0.5 - pz[q];
            float d = (float)(std::sqrt(dx * dx + dy * dy + dz * dz) - radius);
            if (d < row[i]) row[i] = d;
          }
        }
    }
  });
  return ELSIM_OK;
}

// Synthetic nondimensional property tables on the uniform grid X(k)=(k-1)/(sz-1):
//   mu(c)  = Omega(1-2c) + ln((c+d)/(1-c+d)), Omega = 2.5, d = 1e-3 (regular solution)
//   i0(c)  = 1 (scaled by the 1e-3 of SBM:617)
//   OCV(c) = 0.1 - 0.2 c
//   M(c)   = c(1-c) + 0.05
extern "C" int elsim_gen_tables(int sz, double* t) {
  if (sz < 3 || !t) return ELSIM_ERR_ARG;
  for (int tb = 0; tb < 4; tb++)
    for (int k = 0; k < sz; k++) t[(2 * tb) * sz + k] = (double)k / (double)(sz - 1);
  const double om = 2.5, d = 1.0e-3;
  for (int k = 0; k < sz; k++) {
    double c = t[k];
    t[1 * sz + k] = om * (1.0 - 2.0 * c) + std::log((c + d) / (1.0 - c + d));
    t[3 * sz + k] = 1.0;
    t[5 * sz + k] = 0.1 - 0.2 * c;
    t[7 * sz + k] = c * (1.0 - c) + 0.05;
  }
  return ELSIM_OK;
}
