// lbm_dense_kernels.hpp -- structured-grid collide-stream kernels (the fast path).
//
// Domain: one GPU's nx*ny*nz box, cell = x + nx*(y + ny*z).  SOA slot (s, cell)
// lives at s*ncell + cell, so the 64 lanes of a wavefront, which take 64
// consecutive x of one row, touch 64 consecutive elements of each direction's
// array: every load and store is a coalesced row segment.  AOS puts the Q
// values of a cell together (cell*Q + s).
//
// Streaming is index arithmetic with periodic wrap in every dimension (the
// single-GPU form of the reference's adjacency x-wrap, src/geometries/GeometryAAAB.cpp:56-66,
// and of its MPI halo exchange, src/Communication.cpp:379-481).  Solid cells
// and non-periodic faces use halfway bounce-back through a 32-bit link mask
// per cell: bit s set <=> the neighbour at x + c_s is solid or outside a
// non-periodic face; bit 31 marks a solid cell (skipped).  This is the same
// rule as the reference adjacency (src/geometries/GeometryAAAB.cpp:67-85),
// evaluated in-kernel instead of read from an N*Q int32 table.
//
// Kernel kinds and the slots they touch (halfway bounce-back in brackets):
//   AA_EVEN  read slot opp(s) at x, write slot s at x
//            (src/kernels/KernelAA_SOA_Base.cpp:173-254)
//   AA_ODD   read slot s at x - c_s [slot opp(s) at x],
//            write slot opp(s) at x + c_s [slot s at x]
//            (src/kernels/KernelAA_SOA_Base.cpp:257-342)
//   AB_PUSH  read slot s at x, write slot s at x + c_s [slot opp(s) at x]
//            (src/kernels/KernelAB_SOA_Base.cpp:111-186)
//   AB_PULL  read slot s at x - c_s [slot opp(s) at x], write slot s at x
//            (src/kernels/KernelAB_SOA_Pull.cpp:113-194)
// AA touches the same Q slots for reading and writing, so it runs in place
// with no inter-workgroup synchronisation.
#pragma once
#include "lbm_collide.hpp"
#include "lbm_internal.hpp"

namespace lbm {

constexpr uint32_t SOLID_BIT = 1u << 31;

enum DenseKind { K_AA_EVEN = 0, K_AA_ODD = 1, K_AB_PUSH = 2, K_AB_PULL = 3 };

struct DenseArgs {
  int nx, ny, nz;
  int64_t ncell;
  const uint32_t* links;       // nullptr when every dimension is periodic and nothing is solid
  const double* omega_field;   // per-cell omega = 1/tau, or nullptr
  const double* gamma_field;   // per-cell Brinkman gamma, or nullptr
  double omega;
  double F0, F1, F2;
  float scale, inv_scale;      // FP16 deviation scaling
};

template <int STORE>
struct Storage;
template <>
struct Storage<LBM_STORE_F64> {
  using T = double;
  using C = double;
};
template <>
struct Storage<LBM_STORE_F32S> {
  using T = float;
  using C = float;
};
template <>
struct Storage<LBM_STORE_F16S> {
  using T = _Float16;
  using C = float;
};

template <int STORE>
__device__ __forceinline__ typename Storage<STORE>::C ld(const typename Storage<STORE>::T* p,
                                                         int64_t idx, float inv_scale) {
  if constexpr (STORE == LBM_STORE_F16S) return (float)p[idx] * inv_scale;
  else return p[idx];
}

template <int STORE>
__device__ __forceinline__ void st(typename Storage<STORE>::T* p, int64_t idx,
                                   typename Storage<STORE>::C v, float scale) {
  if constexpr (STORE == LBM_STORE_F16S) p[idx] = (_Float16)(v * scale);
  else p[idx] = v;
}

template <int Q, int LAYOUT>
__device__ __forceinline__ int64_t slot(int64_t ncell, int s, int64_t cell) {
  return LAYOUT == LBM_LAYOUT_SOA ? ncell * s + cell : cell * Q + s;
}

// Neighbour offsets with periodic wrap: o[0] = step to coordinate - 1, o[2] = + 1.
struct Offsets {
  int64_t ox[3], oy[3], oz[3];
  __device__ __forceinline__ Offsets(int x, int y, int z, int nx, int ny, int nz) {
    const int64_t plane = (int64_t)nx * ny;
    ox[0] = x == 0 ? (int64_t)(nx - 1) : -1;
    ox[1] = 0;
    ox[2] = x == nx - 1 ? -(int64_t)(nx - 1) : 1;
    oy[0] = (y == 0 ? (int64_t)(ny - 1) : -1) * nx;
    oy[1] = 0;
    oy[2] = (y == ny - 1 ? -(int64_t)(ny - 1) : 1) * (int64_t)nx;
    oz[0] = (z == 0 ? (int64_t)(nz - 1) : -1) * plane;
    oz[1] = 0;
    oz[2] = (z == nz - 1 ? -(int64_t)(nz - 1) : 1) * plane;
  }
  // offset of the neighbour at x + (cx, cy, cz), each in {-1, 0, 1}
  __device__ __forceinline__ int64_t at(int cx, int cy, int cz) const {
    return ox[1 + cx] + oy[1 + cy] + oz[1 + cz];
  }
};

template <int Q, int STORE, int LAYOUT, int KIND, bool MASK>
__global__ void __launch_bounds__(256)
    k_dense(const typename Storage<STORE>::T* f0, typename Storage<STORE>::T* f1, DenseArgs a) {
  using C = typename Storage<STORE>::C;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.nx) return;
  const int y = blockIdx.y, z = blockIdx.z;
  const int64_t cell = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
  uint32_t lk = 0;
  if constexpr (MASK) {
    lk = a.links[cell];
    if (lk & SOLID_BIT) return;
  }
  const Offsets off(x, y, z, a.nx, a.ny, a.nz);
  const int64_t n = a.ncell;

  C d[Q];
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int o = Lattice<Q>::opp(s);
    const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
    int64_t idx;
    if constexpr (KIND == K_AA_EVEN) {
      idx = slot<Q, LAYOUT>(n, o, cell);
    } else if constexpr (KIND == K_AB_PUSH) {
      idx = slot<Q, LAYOUT>(n, s, cell);
    } else {
      if (MASK && ((lk >> o) & 1u)) idx = slot<Q, LAYOUT>(n, o, cell);
      else idx = slot<Q, LAYOUT>(n, s, cell + off.at(-cx, -cy, -cz));
    }
    d[s] = ld<STORE>(f0, idx, a.inv_scale);
  }

  const double om = a.omega_field ? a.omega_field[cell] : a.omega;
  const double F[3] = {a.F0, a.F1, a.F2};
  if constexpr (STORE == LBM_STORE_F64) {
    if (a.gamma_field) collide_ref<Q, double, true>(d, d, om, F, a.gamma_field[cell]);
    else collide_ref<Q, double, false>(d, d, om, F, 0.0);
  } else {
    if (a.gamma_field) collide_shifted<Q, true>(d, om, F, a.gamma_field[cell]);
    else collide_shifted<Q, false>(d, om, F, 0.0);
  }

#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int o = Lattice<Q>::opp(s);
    const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
    int64_t idx;
    if constexpr (KIND == K_AA_EVEN || KIND == K_AB_PULL) {
      idx = slot<Q, LAYOUT>(n, s, cell);
    } else {
      const bool bb = MASK && ((lk >> s) & 1u);
      if constexpr (KIND == K_AA_ODD)
        idx = bb ? slot<Q, LAYOUT>(n, s, cell) : slot<Q, LAYOUT>(n, o, cell + off.at(cx, cy, cz));
      else
        idx = bb ? slot<Q, LAYOUT>(n, o, cell) : slot<Q, LAYOUT>(n, s, cell + off.at(cx, cy, cz));
    }
    st<STORE>(f1, idx, d[s], a.scale);
  }
}

using DenseLaunchFn = void (*)(dim3 grid, dim3 block, hipStream_t stream, void* f0, void* f1,
                               const DenseArgs& a);

template <int Q, int STORE, int LAYOUT, int KIND, bool MASK>
void launch_dense(dim3 grid, dim3 block, hipStream_t stream, void* f0, void* f1,
                  const DenseArgs& a) {
  using T = typename Storage<STORE>::T;
  hipLaunchKernelGGL((k_dense<Q, STORE, LAYOUT, KIND, MASK>), grid, block, 0, stream,
                     (const T*)f0, (T*)f1, a);
}

template <int Q, int STORE, int LAYOUT>
DenseLaunchFn pick_kind(int kind, bool mask) {
  switch (kind) {
    case K_AA_EVEN:
      return mask ? launch_dense<Q, STORE, LAYOUT, K_AA_EVEN, true>
                  : launch_dense<Q, STORE, LAYOUT, K_AA_EVEN, false>;
    case K_AA_ODD:
      return mask ? launch_dense<Q, STORE, LAYOUT, K_AA_ODD, true>
                  : launch_dense<Q, STORE, LAYOUT, K_AA_ODD, false>;
    case K_AB_PUSH:
      return mask ? launch_dense<Q, STORE, LAYOUT, K_AB_PUSH, true>
                  : launch_dense<Q, STORE, LAYOUT, K_AB_PUSH, false>;
    case K_AB_PULL:
      return mask ? launch_dense<Q, STORE, LAYOUT, K_AB_PULL, true>
                  : launch_dense<Q, STORE, LAYOUT, K_AB_PULL, false>;
  }
  return nullptr;
}

template <int Q>
DenseLaunchFn pick_dense_q(int storage, int layout, int kind, bool mask) {
  if (layout == LBM_LAYOUT_SOA) {
    switch (storage) {
      case LBM_STORE_F64: return pick_kind<Q, LBM_STORE_F64, LBM_LAYOUT_SOA>(kind, mask);
      case LBM_STORE_F32S: return pick_kind<Q, LBM_STORE_F32S, LBM_LAYOUT_SOA>(kind, mask);
      case LBM_STORE_F16S: return pick_kind<Q, LBM_STORE_F16S, LBM_LAYOUT_SOA>(kind, mask);
    }
  } else {
    switch (storage) {
      case LBM_STORE_F64: return pick_kind<Q, LBM_STORE_F64, LBM_LAYOUT_AOS>(kind, mask);
      case LBM_STORE_F32S: return pick_kind<Q, LBM_STORE_F32S, LBM_LAYOUT_AOS>(kind, mask);
      case LBM_STORE_F16S: return pick_kind<Q, LBM_STORE_F16S, LBM_LAYOUT_AOS>(kind, mask);
    }
  }
  return nullptr;
}

// Defined in lbm_dense_q19.hip and lbm_dense_q27.hip (one translation unit per Q).
DenseLaunchFn pick_dense19(int storage, int layout, int kind, bool mask);
DenseLaunchFn pick_dense27(int storage, int layout, int kind, bool mask);

}  // namespace lbm
