// lbm_lattice.hpp -- compile-time D3Q19 / D3Q27 velocity sets for host and device.
//
// Direction order, opposite map and weights follow the reference exactly
// (src/lattices/LatticeD3Q19.cpp:12-86, src/lattices/LatticeD3Q27.cpp:12-118):
// 0-5 axis (+x,-x,+y,-y,+z,-z), 6-17 edges (xy, xz, yz pairs), D3Q27 18-25
// corners, rest last.  The fluid-index storage order of populations (slot s)
// uses this order, so a state written by the reference kernels is readable here.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbm {

template <int Q>
struct Lattice;

// The 27 directions in reference order; D3Q19 uses entries 0-17 and the rest (26).
__host__ __device__ constexpr int dir27(int s, int d) {
  constexpr int t[27][3] = {
      {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},    {0, 0, -1}, {1, 1, 0},
      {1, -1, 0},  {1, 0, 1},   {1, 0, -1},  {-1, 1, 0},  {-1, -1, 0},  {-1, 0, 1}, {-1, 0, -1},
      {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},    {1, 1, -1}, {1, -1, 1},
      {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}, {0, 0, 0}};
  return t[s][d];
}

template <>
struct Lattice<19> {
  static constexpr int Q = 19;
  __host__ __device__ static constexpr int c(int s, int d) { return s == 18 ? 0 : dir27(s, d); }
  __host__ __device__ static constexpr int opp(int s) {
    constexpr int o[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
    return o[s];
  }
  // weights 1/18 (axis), 1/36 (edge), 1/3 (rest), as double expressions so the
  // host and device constants are the correctly rounded quotients.
  __host__ __device__ static constexpr double w(int s) {
    return s < 6 ? 1.0 / 18.0 : (s < 18 ? 1.0 / 36.0 : 1.0 / 3.0);
  }
};

template <>
struct Lattice<27> {
  static constexpr int Q = 27;
  __host__ __device__ static constexpr int c(int s, int d) { return dir27(s, d); }
  __host__ __device__ static constexpr int opp(int s) {
    constexpr int o[27] = {1,  0,  3,  2,  5,  4,  11, 10, 13, 12, 7,  6,  9, 8,
                           17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18, 26};
    return o[s];
  }
  __host__ __device__ static constexpr double w(int s) {
    return s < 6 ? 2.0 / 27.0 : (s < 18 ? 1.0 / 54.0 : (s < 26 ? 1.0 / 216.0 : 8.0 / 27.0));
  }
};

}  // namespace lbm
