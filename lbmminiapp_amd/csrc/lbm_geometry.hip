// lbm_geometry.hip -- host-side geometry and field generators behind the C ABI:
// the solid masks every configuration of the engine runs on, and the smooth
// seeded per-node fields (Brinkman gamma, tau).  Replaces Geometry::setupGrid
// (src/geometries/Geometry.cpp:100-151), whose only case is the cylinder.
//
// Mask convention: solid[x + nx * (y + ny * z)] != 0 marks a solid cell.
// Random numbers: U = top53(splitmix64(seed * 0x9E3779B97F4A7C15 + counter)) / 2^53,
// the counter-based generator the rest of the engine uses.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "lbm_internal.hpp"

namespace {

uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t seed, uint64_t counter) {
  return (double)(splitmix64(seed * 0x9E3779B97F4A7C15ull + counter) >> 11) *
         (1.0 / 9007199254740992.0);
}

inline int64_t cell_of(int x, int y, int z, int nx, int ny) {
  return x + (int64_t)nx * (y + (int64_t)ny * z);
}

}  // namespace

extern "C" {

// Reference cylinder (src/geometries/Geometry.cpp:60-70, :113-131): radius
// int(8 s), length int(84 s), box 84s x (2r+3) x (2r+3), axis at (r + 1.5, r + 1.5)
// in (y, z); a cell is fluid when int(sqrt(dy^2 + dz^2)) <= r.  With solid = NULL
// only the dimensions are returned.
lbm_status lbm_geometry_cylinder(double simulation_size, int32_t* dims, uint8_t* solid) {
  LBM_REQUIRE(dims && simulation_size > 0.0, LBM_ERR_ARG,
              "lbm_geometry_cylinder: need dims and simulation_size > 0");
  const int radius_i = (int)(8.0 * simulation_size);
  const int length = (int)(84.0 * simulation_size);
  LBM_REQUIRE(radius_i >= 1 && length >= 1, LBM_ERR_ARG, "lbm_geometry_cylinder: size too small");
  dims[0] = length;
  dims[1] = 2 * radius_i + 3;
  dims[2] = 2 * radius_i + 3;
  if (!solid) return LBM_OK;
  const double radius = (double)((dims[1] - 3) / 2);
  const double center = radius + 1.5;
  for (int z = 0; z < dims[2]; z++)
    for (int y = 0; y < dims[1]; y++) {
      const int dist = (int)std::sqrt(std::pow(y - center, 2.0) + std::pow(z - center, 2.0));
      const uint8_t v = dist <= radius ? 0 : 1;
      for (int x = 0; x < dims[0]; x++) solid[cell_of(x, y, z, dims[0], dims[1])] = v;
    }
  return LBM_OK;
}

// Channel: solid layers at y = 0 and y = ny - 1, everything else fluid.
lbm_status lbm_geometry_channel(int32_t nx, int32_t ny, int32_t nz, uint8_t* solid) {
  LBM_REQUIRE(solid && nx > 0 && ny > 2 && nz > 0, LBM_ERR_ARG, "lbm_geometry_channel: bad args");
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        solid[cell_of(x, y, z, nx, ny)] = (y == 0 || y == ny - 1) ? 1 : 0;
  return LBM_OK;
}

// Seeded random spheres added to an existing mask (walls are kept) until the
// solid fraction of the box reaches `fraction`.  Sphere k has centre
// (U(4k) nx, U(4k+1) ny, U(4k+2) nz) and radius rmin + (rmax - rmin) U(4k+3);
// spheres wrap around the faces flagged in `periodic`.  Overlaps are allowed
// (random sequential addition), so a high fraction gives a packed porous medium.
// Returns the number of spheres placed in *count (may be NULL).
lbm_status lbm_geometry_add_spheres(int32_t nx, int32_t ny, int32_t nz, const int32_t* periodic,
                                    double rmin, double rmax, double fraction, uint64_t seed,
                                    uint8_t* solid, int64_t* count) {
  LBM_REQUIRE(solid && periodic && nx > 0 && ny > 0 && nz > 0, LBM_ERR_ARG,
              "lbm_geometry_add_spheres: bad args");
  LBM_REQUIRE(rmin > 0.0 && rmax >= rmin && fraction > 0.0 && fraction < 0.95, LBM_ERR_ARG,
              "lbm_geometry_add_spheres: need 0 < rmin <= rmax and 0 < fraction < 0.95");
  const int64_t ncell = (int64_t)nx * ny * nz;
  int64_t nsolid = 0;
  for (int64_t c = 0; c < ncell; c++) nsolid += solid[c] != 0;
  const int64_t target = (int64_t)std::ceil(fraction * (double)ncell);
  const int dims[3] = {nx, ny, nz};
  int64_t k = 0;
  const int64_t max_spheres = 100000000;
  while (nsolid < target && k < max_spheres) {
    const double c[3] = {uniform01(seed, 4 * k) * nx, uniform01(seed, 4 * k + 1) * ny,
                         uniform01(seed, 4 * k + 2) * nz};
    const double r = rmin + (rmax - rmin) * uniform01(seed, 4 * k + 3);
    k++;
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = (int)std::floor(c[d] - r);
      hi[d] = (int)std::ceil(c[d] + r);
      if (!periodic[d]) {
        lo[d] = std::max(lo[d], 0);
        hi[d] = std::min(hi[d], dims[d] - 1);
      }
    }
    for (int z = lo[2]; z <= hi[2]; z++)
      for (int y = lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++) {
          const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
          if (dx * dx + dy * dy + dz * dz > r * r) continue;
          const int p[3] = {((x % nx) + nx) % nx, ((y % ny) + ny) % ny, ((z % nz) + nz) % nz};
          uint8_t& v = solid[cell_of(p[0], p[1], p[2], nx, ny)];
          if (!v) {
            v = 1;
            nsolid++;
          }
        }
  }
  if (count) *count = k;
  LBM_REQUIRE(nsolid >= target, LBM_ERR_ARG, "lbm_geometry_add_spheres: fraction not reached");
  return LBM_OK;
}

// Smooth seeded field in [0, 1]: white noise U(seed, cell), periodic Gaussian
// filter with standard deviation sigma (separable, kernel truncated at
// 4 sigma and renormalised), then min-max rescaled to [0, 1].
lbm_status lbm_field_smooth_random(int32_t nx, int32_t ny, int32_t nz, double sigma,
                                   uint64_t seed, double* out) {
  LBM_REQUIRE(out && nx > 0 && ny > 0 && nz > 0 && sigma > 0.0, LBM_ERR_ARG,
              "lbm_field_smooth_random: bad args");
  const int64_t ncell = (int64_t)nx * ny * nz;
  std::vector<double> a(ncell), b(ncell);
  for (int64_t c = 0; c < ncell; c++) a[c] = uniform01(seed, (uint64_t)c);
  const int radius = (int)(4.0 * sigma + 0.5);
  std::vector<double> wts(2 * radius + 1);
  double wsum = 0.0;
  for (int i = -radius; i <= radius; i++) {
    wts[i + radius] = std::exp(-0.5 * (double)i * i / (sigma * sigma));
    wsum += wts[i + radius];
  }
  for (double& w : wts) w /= wsum;
  // One pass per dimension.  Each output x-row is accumulated tap by tap
  // (i = -radius .. radius, acc = 0 + w_i a_i + ...), so every cell sees the
  // same sum in the same order whichever dimension is filtered; the y and z
  // passes read whole source rows (contiguous), and the planes are split over
  // host threads (the result does not depend on the split).
  auto pass = [&](int d, int z0, int z1) {
    std::vector<double> acc(nx);
    for (int z = z0; z < z1; z++)
      for (int y = 0; y < ny; y++) {
        double* dst = &b[cell_of(0, y, z, nx, ny)];
        if (d == 0) {
          const double* row = &a[cell_of(0, y, z, nx, ny)];
          for (int x = 0; x < nx; x++) {
            double s = 0.0;
            for (int i = -radius; i <= radius; i++)
              s += wts[i + radius] * row[((x + i) % nx + nx) % nx];
            dst[x] = s;
          }
          continue;
        }
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int i = -radius; i <= radius; i++) {
          const int yy = d == 1 ? ((y + i) % ny + ny) % ny : y;
          const int zz = d == 2 ? ((z + i) % nz + nz) % nz : z;
          const double* row = &a[cell_of(0, yy, zz, nx, ny)];
          const double w = wts[i + radius];
          for (int x = 0; x < nx; x++) acc[x] += w * row[x];
        }
        std::copy(acc.begin(), acc.end(), dst);
      }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const int nthreads = std::max(1, std::min<int>(nz, std::min<unsigned>(hw ? hw : 1, 16)));
  for (int d = 0; d < 3; d++) {
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
      pool.emplace_back(pass, d, (int)((int64_t)nz * t / nthreads),
                        (int)((int64_t)nz * (t + 1) / nthreads));
    for (auto& th : pool) th.join();
    a.swap(b);
  }
  const auto mm = std::minmax_element(a.begin(), a.end());
  const double lo = *mm.first, hi = *mm.second;
  const double span = hi > lo ? hi - lo : 1.0;
  for (int64_t c = 0; c < ncell; c++) out[c] = (a[c] - lo) / span;
  return LBM_OK;
}

}  // extern "C"
