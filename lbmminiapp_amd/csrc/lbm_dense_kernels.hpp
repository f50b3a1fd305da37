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
// non-periodic face; bit 31 marks a solid cell (no collision; its slots are
// never read and, with DenseArgs::fill_solid, receive zeros so that every row
// segment is stored whole).  This is the same rule as the reference adjacency
// (src/geometries/GeometryAAAB.cpp:67-85), evaluated in-kernel instead of read
// from an N*Q int32 table.
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
//
// Interleaved rows (SOA boxes with nx a multiple of 64, DenseArgs::xk = K =
// nx / 64): element x of every row is stored at position
// pos(x) = (x mod K) * 64 + x / K, so the 64 positions 64w .. 64w + 63 of block
// w hold x = w, w + K, ..., w + 63K.  Shifting x by +-1 maps block w onto block
// w +- 1 lane for lane, and the two end blocks onto each other with the lanes
// rotated by one (the periodic wrap).  A wavefront that covers one block
// therefore reads and writes one whole aligned block of each row in the AA odd
// step too (K_AA_ODD_IR on unmasked boxes), exactly as in the even step, which
// is position-local and unchanged.  The other kinds take interleaved rows at
// run time (a.xk): their x-neighbour offsets come from ir_shift, and the tests
// that need the coordinate (walls, pressure-drop crossings) use x itself.  The
// link mask and the per-cell fields are stored in the same order.  Readout and
// initialisation map x to pos(x).
#pragma once
#include "lbm_collide.hpp"
#include "lbm_internal.hpp"

namespace lbm {

constexpr uint32_t SOLID_BIT = 1u << 31;

// K_AA_ODD_IR: the AA odd step on interleaved rows
enum DenseKind { K_AA_EVEN = 0, K_AA_ODD = 1, K_AB_PUSH = 2, K_AB_PULL = 3, K_AA_ODD_IR = 4 };

// Interleaved rows: position of x, x at position p, and the position of
// x + d (d in {-1, 0, 1}, periodic) from the position p of x.
__host__ __device__ __forceinline__ int ir_pos(int x, int K) { return (x % K) * 64 + x / K; }
__host__ __device__ __forceinline__ int ir_x(int p, int K) { return (p & 63) * K + (p >> 6); }
__host__ __device__ __forceinline__ int ir_shift(int p, int d, int K) {
  const int w = p >> 6, l = p & 63;
  if (d > 0) return w + 1 < K ? p + 64 : ((l + 1) & 63);
  if (d < 0) return w > 0 ? p - 64 : (K - 1) * 64 + ((l - 1) & 63);
  return p;
}

struct DenseArgs {
  int nx, ny, nz;
  int64_t ncell;
  uint32_t periodic_bits;      // bit k: dimension k is periodic
  int fill_solid;              // write zeros into the slots of solid cells (whole-line stores)
  const uint32_t* links;       // nullptr when every dimension is periodic and nothing is solid
  const double* omega_field;   // per-cell omega = 1/tau, or nullptr
  const double* gamma_field;   // per-cell Brinkman gamma, or nullptr
  double omega;
  double F0, F1, F2;
  float scale, inv_scale;      // FP16 deviation scaling
  double djump;                // density jump across the periodic x boundary (0: none)
  int z0;                      // first z-plane of the launch (z = blockIdx.z + z0): 0 except in
                               // the z-slab schedule of the AA steps
  int xk;                      // interleaved rows: nx / 64; 0: element x at position x
};

// Pressure-drop boundary: the value a population carries across the periodic
// x boundary changes by sign * w_s * djump (sign +1 forward, from x = nx - 1 to
// 0; -1 backward).  Gathers: node x reads from x - c_s; scatters: node x writes
// to x + c_s.
__device__ __forceinline__ int gather_crossing(int cx, int x, int nx) {
  return (cx > 0 && x == 0) ? 1 : ((cx < 0 && x == nx - 1) ? -1 : 0);
}
__device__ __forceinline__ int scatter_crossing(int cx, int x, int nx) {
  return (cx > 0 && x == nx - 1) ? 1 : ((cx < 0 && x == 0) ? -1 : 0);
}
template <int Q, typename C>
__device__ __forceinline__ C with_jump(C v, int sign, int s, double djump) {
  if (sign == 0 || djump == 0.0) return v;
  const C wj = (C)(Lattice<Q>::w(s) * djump);
  return sign > 0 ? v + wj : v - wj;
}

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

// Streaming (non-temporal) variants for the time-step kernels: every population
// is touched once per step, so there is nothing to keep in the caches.
template <int STORE>
__device__ __forceinline__ typename Storage<STORE>::C ld_nt(const typename Storage<STORE>::T* p,
                                                            int64_t idx, float inv_scale) {
  const typename Storage<STORE>::T v = __builtin_nontemporal_load(p + idx);
  if constexpr (STORE == LBM_STORE_F16S) return (float)v * inv_scale;
  else return v;
}

template <int STORE>
__device__ __forceinline__ void st_nt(typename Storage<STORE>::T* p, int64_t idx,
                                      typename Storage<STORE>::C v, float scale) {
  if constexpr (STORE == LBM_STORE_F16S) __builtin_nontemporal_store((_Float16)(v * scale), p + idx);
  else __builtin_nontemporal_store(v, p + idx);
}

// Streaming accesses of the time-step kernels: non-temporal when the grid
// exceeds the memory-side cache (NT), cached when it fits there.
template <int STORE, bool NT>
__device__ __forceinline__ typename Storage<STORE>::C ld_s(const typename Storage<STORE>::T* p,
                                                           int64_t idx, float inv_scale) {
  if constexpr (NT) return ld_nt<STORE>(p, idx, inv_scale);
  else return ld<STORE>(p, idx, inv_scale);
}
template <int STORE, bool NT>
__device__ __forceinline__ void st_s(typename Storage<STORE>::T* p, int64_t idx,
                                     typename Storage<STORE>::C v, float scale) {
  if constexpr (NT) st_nt<STORE>(p, idx, v, scale);
  else st<STORE>(p, idx, v, scale);
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

// SOA element (s, x', y', z') addressed as the row (s, y', z') plus the lane's
// x'.  Within a workgroup every lane has the same (y, z), so the row address
// is uniform.  row_elem moves it into scalar registers (readfirstlane: the
// value is the same in every lane) as a global-memory pointer and adds the
// lane's 32-bit byte offset, so that each access is one load or store with a
// scalar base and a 32-bit lane offset, instead of a 64-bit per-lane address
// sum per access (the compiler otherwise folds the 19 row bases into one
// per-lane address chain).
template <typename T>
__device__ __forceinline__ T* row_of(T* f, int64_t n, int s, const DenseArgs& a, int yy, int zz) {
  return f + (n * s + (int64_t)a.nx * (yy + (int64_t)a.ny * zz));
}
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ global_ptr<T> row_elem(T* row, uint32_t xx) {
  const uint64_t v = (uint64_t)row;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  const auto base = (global_ptr<const char>)(((uint64_t)hi << 32) | lo);
  return (global_ptr<T>)(base + xx * (uint32_t)sizeof(T));
}
// Loads and stores of one row element (non-temporal when NT), converting FP16
// deviations as ld / st do.
template <int STORE, bool NT>
__device__ __forceinline__ typename Storage<STORE>::C ld_row(const typename Storage<STORE>::T* row,
                                                             uint32_t xx, float inv_scale) {
  const auto p = row_elem(row, xx);
  typename Storage<STORE>::T v;
  if constexpr (NT) v = __builtin_nontemporal_load(p);
  else v = *p;
  if constexpr (STORE == LBM_STORE_F16S) return (float)v * inv_scale;
  else return v;
}
template <int STORE, bool NT>
__device__ __forceinline__ void st_row(typename Storage<STORE>::T* row, uint32_t xx,
                                       typename Storage<STORE>::C v, float scale) {
  using T = typename Storage<STORE>::T;
  const auto p = row_elem(row, xx);
  T w;
  if constexpr (STORE == LBM_STORE_F16S) w = (T)(v * scale);
  else w = v;
  if constexpr (NT) __builtin_nontemporal_store(w, p);
  else *p = w;
}
// periodic wrap of v + (-1, 0 or 1) into [0, nv)
__device__ __forceinline__ int wrap1(int v, int nv) {
  return v < 0 ? v + nv : (v >= nv ? v - nv : v);
}

// true when x + c leaves the box through a non-periodic face
__device__ __forceinline__ bool leaves_box(const DenseArgs& a, int x, int y, int z, int cx, int cy,
                                           int cz) {
  const uint32_t p = a.periodic_bits;
  return (!(p & 1u) && ((cx < 0 && x == 0) || (cx > 0 && x == a.nx - 1))) ||
         (!(p & 2u) && ((cy < 0 && y == 0) || (cy > 0 && y == a.ny - 1))) ||
         (!(p & 4u) && ((cz < 0 && z == 0) || (cz > 0 && z == a.nz - 1)));
}

// One node's collision, with the Brinkman term when a gamma field is given.
// ONE: a single collision body with gamma = 0 where there is no field (the
// factor 1 - 0 = 1 leaves every velocity bit unchanged, so the results are the
// same); otherwise separate bodies with and without the Brinkman term.  The
// single body needs fewer VGPRs but schedules differently; which is faster
// depends on the kernel (one_collision_body).
template <int Q, int STORE, bool FORCE, bool ONE>
__device__ __forceinline__ void collide_cell(typename Storage<STORE>::C* d, double om,
                                             const double* F, const double* gamma_field,
                                             int64_t cell) {
  if constexpr (ONE) {
    const double gm = gamma_field ? gamma_field[cell] : 0.0;
    if constexpr (STORE == LBM_STORE_F64) collide_ref<Q, double, true, FORCE>(d, d, om, F, gm);
    else collide_shifted<Q, true, FORCE>(d, om, F, gm);
  } else if constexpr (STORE == LBM_STORE_F64) {
    if (gamma_field) collide_ref<Q, double, true, FORCE>(d, d, om, F, gamma_field[cell]);
    else collide_ref<Q, double, false, FORCE>(d, d, om, F, 0.0);
  } else {
    if (gamma_field) collide_shifted<Q, true, FORCE>(d, om, F, gamma_field[cell]);
    else collide_shifted<Q, false, FORCE>(d, om, F, 0.0);
  }
}

// Single or separate collision bodies, from 512^3 timings of every kind with
// alternating launches on one array (ratio one / two bodies):
//   FP16 deviations: one everywhere (0.77-0.98; even steps 0.97-1.01)
//   FP32 deviations: one for AA odd steps with forcing (0.81 D3Q19, 0.83-0.93
//     D3Q27) and D3Q27 AA odd steps without (0.90); two elsewhere (D3Q19 odd
//     without forcing 1.05, D3Q19 push 1.04-1.05, even and pull 1.00-1.03)
//   FP64: two everywhere (D3Q19 odd with forcing 1.07, D3Q27 push 1.05,
//     the rest 0.98-1.02)
template <int Q, int STORE, int KIND, bool FORCE>
constexpr bool one_collision_body() {
  if (STORE == LBM_STORE_F16S) return true;
  if (STORE == LBM_STORE_F32S) return (KIND == K_AA_ODD || KIND == K_AA_ODD_IR) && (FORCE || Q == 27);
  return false;
}

// Blocks are one x-row segment of up to kDenseMaxBlock cells (a full row when
// nx <= kDenseMaxBlock, so that every cache line shared by x-shifted accesses is
// used by one workgroup).
constexpr int kDenseMaxBlock = 512;

// Solid cells (MASK): their slots are never read.  With a.fill_solid every
// (slot, cell) of the box is still stored exactly once per step, zeros going
// to solid cells, so that no cache line is left partly written (a partly
// written line costs the memory a read-modify-write):
//   local-write kinds: a solid cell stores zeros into its own slots;
//   scatter kinds, direction s, target T = x + c_s:
//     T fluid          fluid x streams; solid x stores nothing
//     T solid (inside) fluid x bounces back, and either x stores a zero into
//                      the slot at T that streaming would have written
//     T outside        fluid x bounces back; solid x stores a zero into that
//                      bounce-back slot
template <int Q, int STORE, int LAYOUT, int KIND, bool MASK, bool FORCE, bool NT = true,
          bool ONE = one_collision_body<Q, STORE, KIND, FORCE>()>
__global__ void __launch_bounds__(kDenseMaxBlock)
    k_dense(const typename Storage<STORE>::T* f0, typename Storage<STORE>::T* f1, DenseArgs a) {
  using C = typename Storage<STORE>::C;
  constexpr bool IR = KIND == K_AA_ODD_IR;
  static_assert(!IR || (!MASK && LAYOUT == LBM_LAYOUT_SOA), "interleaved rows: unmasked SOA only");
  // x: the thread's row position (on interleaved rows, a.xk != 0, element xr
  // of the row).  Interleaved rows make every x-shifted access one aligned
  // block as well (see the file comment).
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.nx) return;
  const int xr = a.xk ? ir_x(x, a.xk) : x;
  const bool irows = IR || a.xk != 0;
  const int y = blockIdx.y, z = blockIdx.z + a.z0;
  const int64_t cell = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
  uint32_t lk = 0;
  bool solid = false;
  if constexpr (MASK) {
    lk = a.links[cell];
    solid = (lk & SOLID_BIT) != 0;
    // a wavefront of solid cells only completes no partly written line: its
    // zeros would be pure extra traffic (whole solid regions of a box).  The
    // vote is taken by every lane of the wavefront, before any lane leaves.
    const bool all_solid = __all(solid);
    if (solid && (!a.fill_solid || all_solid)) return;
  }
  Offsets off(xr, y, z, a.nx, a.ny, a.nz);
  if (a.xk) {
    off.ox[0] = ir_shift(x, -1, a.xk) - x;
    off.ox[2] = ir_shift(x, 1, a.xk) - x;
  }
  const int64_t n = a.ncell;

  // Fully periodic SOA boxes (no mask): every access is a row element, see row_of.
  constexpr bool ROWS = !MASK && LAYOUT == LBM_LAYOUT_SOA;
  C d[Q];
  if (!solid) {
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      // Row-aligned accesses stream past the caches; x-shifted ones (c_x != 0 in
      // the gathering kinds) share their boundary cache line with the
      // neighbouring wavefront and must stay cacheable.
      bool aligned = true;
      int cross = 0;  // pressure-drop boundary crossed by this gather
      if constexpr (ROWS) {
        int ss = s, xx = x, yy = y, zz = z;
        if constexpr (KIND == K_AA_EVEN) {
          ss = o;
        } else if constexpr (IR) {
          // the element at x - c_s: a whole aligned block per wavefront
          xx = ir_shift(x, -cx, a.xk);
          yy = wrap1(y - cy, a.ny);
          zz = wrap1(z - cz, a.nz);
          cross = gather_crossing(cx, xr, a.nx);
        } else if constexpr (KIND == K_AA_ODD || KIND == K_AB_PULL) {
          aligned = cx == 0 || irows;
          xx = irows ? ir_shift(x, -cx, a.xk) : wrap1(x - cx, a.nx);
          yy = wrap1(y - cy, a.ny);
          zz = wrap1(z - cz, a.nz);
          cross = gather_crossing(cx, xr, a.nx);
        }
        const auto* rp = row_of(f0, n, ss, a, yy, zz);
        if (aligned) d[s] = ld_row<STORE, NT>(rp, (uint32_t)xx, a.inv_scale);
        else d[s] = ld_row<STORE, false>(rp, (uint32_t)xx, a.inv_scale);
        d[s] = with_jump<Q>(d[s], cross, s, a.djump);
        continue;
      }
      int64_t idx;
      if constexpr (KIND == K_AA_EVEN) {
        idx = slot<Q, LAYOUT>(n, o, cell);
      } else if constexpr (KIND == K_AB_PUSH) {
        idx = slot<Q, LAYOUT>(n, s, cell);
      } else {
        aligned = cx == 0 || irows;
        if (MASK && ((lk >> o) & 1u)) {
          idx = slot<Q, LAYOUT>(n, o, cell);
        } else {
          idx = slot<Q, LAYOUT>(n, s, cell + off.at(-cx, -cy, -cz));
          cross = gather_crossing(cx, xr, a.nx);
        }
      }
      if (aligned && LAYOUT == LBM_LAYOUT_SOA) d[s] = ld_s<STORE, NT>(f0, idx, a.inv_scale);
      else d[s] = ld<STORE>(f0, idx, a.inv_scale);
      d[s] = with_jump<Q>(d[s], cross, s, a.djump);
    }

    const double om = a.omega_field ? a.omega_field[cell] : a.omega;
    const double F[3] = {a.F0, a.F1, a.F2};
    collide_cell<Q, STORE, FORCE, ONE>(d, om, F, a.gamma_field, cell);
  }

  if constexpr (ROWS) {
    auto put_row = [&](typename Storage<STORE>::T* rp, int xx, C v, bool aligned) {
      if (aligned) st_row<STORE, NT>(rp, (uint32_t)xx, v, a.scale);
      else st_row<STORE, false>(rp, (uint32_t)xx, v, a.scale);
    };
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      if constexpr (KIND == K_AA_EVEN || KIND == K_AB_PULL) {
        put_row(row_of(f1, n, s, a, y, z), x, d[s], true);
      } else if constexpr (IR) {
        put_row(row_of(f1, n, o, a, wrap1(y + cy, a.ny), wrap1(z + cz, a.nz)), ir_shift(x, cx, a.xk),
                with_jump<Q>(d[s], scatter_crossing(cx, xr, a.nx), s, a.djump), true);
      } else {
        // streaming target: AA odd writes slot opp(s) at x + c_s, AB push slot s
        put_row(row_of(f1, n, KIND == K_AA_ODD ? o : s, a, wrap1(y + cy, a.ny), wrap1(z + cz, a.nz)),
                irows ? ir_shift(x, cx, a.xk) : wrap1(x + cx, a.nx),
                with_jump<Q>(d[s], scatter_crossing(cx, xr, a.nx), s, a.djump), cx == 0 || irows);
      }
    }
    return;
  }
  auto put = [&](int64_t idx, C v, bool aligned) {
    if (aligned && LAYOUT == LBM_LAYOUT_SOA) st_s<STORE, NT>(f1, idx, v, a.scale);
    else st<STORE>(f1, idx, v, a.scale);
  };
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int o = Lattice<Q>::opp(s);
    const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
    if constexpr (KIND == K_AA_EVEN || KIND == K_AB_PULL) {
      put(slot<Q, LAYOUT>(n, s, cell), solid ? C(0) : d[s], true);
    } else {
      const bool aligned = cx == 0 || irows;
      // streaming target: AA odd writes slot opp(s) at x + c_s, AB push slot s
      const int64_t sidx = slot<Q, LAYOUT>(n, KIND == K_AA_ODD ? o : s, cell + off.at(cx, cy, cz));
      const bool bb = MASK && ((lk >> s) & 1u);
      if (!bb) {
        if (!solid) put(sidx, with_jump<Q>(d[s], scatter_crossing(cx, xr, a.nx), s, a.djump), aligned);
      } else {
        // halfway bounce-back slot: AA odd slot s at x, AB push slot opp(s) at x
        const int64_t bidx = slot<Q, LAYOUT>(n, KIND == K_AA_ODD ? s : o, cell);
        const bool out = leaves_box(a, xr, y, z, cx, cy, cz);
        if (!solid) put(bidx, d[s], aligned);
        else if (out) put(bidx, C(0), aligned);
        if (!out && a.fill_solid) put(sidx, C(0), aligned);
      }
    }
  }
}

// number of directions with c_x != 0
template <int Q>
__host__ __device__ constexpr int n_xshift() {
  int k = 0;
  for (int s = 0; s < Q; s++) k += Lattice<Q>::c(s, 0) != 0;
  return k;
}

// LDS-staged variant of the gathering / scattering kinds (AA_ODD, AB_PUSH,
// AB_PULL) for fully periodic SOA boxes: one workgroup is one whole x-row
// (blockDim.x == nx, nx a multiple of 64).  Every global access is the aligned
// element x of some row, so each wavefront moves whole cache lines exactly
// once (streamed past the caches).  The x-shift of the directions with
// c_x != 0, including the periodic wrap at the row ends, is applied in LDS:
//   gather : thread x stores element x of row (y-cy, z-cz) in LDS, node x then
//            reads LDS[(x - cx) mod nx]
//   scatter: node x writes its result to LDS[(x + cx) mod nx], thread x then
//            stores LDS[x] as element x of row (y+cy, z+cz)
// Same reads, writes and arithmetic as k_dense, so the results are identical.
// MASK (AB_PULL only): every thread still loads its aligned row element (the
// periodic offsets keep rows in bounds across non-periodic faces too); a node
// whose neighbour x - c_s is solid or outside then takes slot opp(s) at x
// instead, solid nodes skip the collision and store zeros (a.fill_solid), and
// all threads stay for the barrier.
// Waves per SIMD the staged kernel is compiled for.  A whole-row workgroup of
// 512 threads is 8 wavefronts, 2 per SIMD, so a CU holds two of them only at
// 4 waves per SIMD, i.e. at most 128 VGPRs.  D3Q19 FP64 needs 130 (144 with
// the Guo terms) unbounded; bounded to 128 (7 spilled with the Guo terms), two
// rows per CU overlap each other's barriers.  512^3 AA odd step, alternating
// launches on one array: with the Guo terms 7.47-7.51 ms bounded vs 8.17-8.24
// unbounded, without 7.83-7.95 vs 7.99-8.02.  Bounds of 5-8 waves spill more
// and ran slower (8.8-18 ms).  Q27 spills heavily at 128 and ran 10-45%
// slower, FP32/FP16 deviations gained nothing, so they stay unbounded
// (1 = no constraint).
template <int Q, int STORE>
constexpr int staged_min_waves() {
  return (Q == 19 && STORE == LBM_STORE_F64) ? 4 : 1;
}

template <int Q, int STORE, int KIND, bool FORCE, bool MASK = false, bool NT = true,
          int WAVES = staged_min_waves<Q, STORE>(),
          bool ONE = one_collision_body<Q, STORE, KIND, FORCE>()>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kDenseMaxBlock),
                               amdgpu_waves_per_eu(WAVES)))
    k_dense_staged(const typename Storage<STORE>::T* f0, typename Storage<STORE>::T* f1,
                   DenseArgs a) {
  static_assert(!MASK || KIND == K_AB_PULL, "masked staging covers the pull gather");
  using C = typename Storage<STORE>::C;
  constexpr int NSH = n_xshift<Q>();
  extern __shared__ unsigned char lbm_smem[];
  C* lds = reinterpret_cast<C*>(lbm_smem);  // [NSH][nx]
  const int nx = a.nx;
  const int x = threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z + a.z0;
  const int64_t cell = x + (int64_t)nx * (y + (int64_t)a.ny * z);
  const int64_t n = a.ncell;
  constexpr bool GATHER = KIND == K_AA_ODD || KIND == K_AB_PULL;
  constexpr bool SCATTER = KIND == K_AA_ODD || KIND == K_AB_PUSH;

  C d[Q];
  {
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      if constexpr (GATHER) {
        // element x of slot s in row (y - cy, z - cz)
        const C v = ld_row<STORE, NT>(
            row_of(f0, n, s, a, wrap1(y - cy, a.ny), wrap1(z - cz, a.nz)), (uint32_t)x, a.inv_scale);
        if (cx == 0) d[s] = v;
        else lds[(k++) * nx + x] = v;
      } else {
        d[s] = ld_row<STORE, NT>(row_of(f0, n, s, a, y, z), (uint32_t)x, a.inv_scale);
      }
    }
  }
  if constexpr (GATHER) {
    __syncthreads();
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int cx = Lattice<Q>::c(s, 0);
      if (cx != 0) {
        int xs = x - cx;
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        d[s] = with_jump<Q>(lds[(k++) * nx + xs], gather_crossing(cx, x, nx), s, a.djump);
      }
    }
  }

  uint32_t lk = 0;
  bool solid = false;
  if constexpr (MASK) {
    lk = a.links[cell];
    solid = (lk & SOLID_BIT) != 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      if (!solid && ((lk >> o) & 1u))
        d[s] = ld_row<STORE, NT>(row_of(f0, n, o, a, y, z), (uint32_t)x, a.inv_scale);
    }
  }

  if (!solid) {
    const double om = a.omega_field ? a.omega_field[cell] : a.omega;
    const double F[3] = {a.F0, a.F1, a.F2};
    collide_cell<Q, STORE, FORCE, ONE>(d, om, F, a.gamma_field, cell);
  }

  if constexpr (SCATTER) {
    if constexpr (GATHER) __syncthreads();  // every LDS read of the gather is done
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      const int tgt = KIND == K_AA_ODD ? o : s;
      if (cx == 0) {
        st_row<STORE, NT>(row_of(f1, n, tgt, a, wrap1(y + cy, a.ny), wrap1(z + cz, a.nz)),
                          (uint32_t)x, d[s], a.scale);
      } else {
        int xt = x + cx;
        xt = xt < 0 ? xt + nx : (xt >= nx ? xt - nx : xt);
        lds[(k++) * nx + xt] = with_jump<Q>(d[s], scatter_crossing(cx, x, nx), s, a.djump);
      }
    }
    __syncthreads();
    k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      const int tgt = KIND == K_AA_ODD ? o : s;
      if (cx != 0)
        st_row<STORE, NT>(row_of(f1, n, tgt, a, wrap1(y + cy, a.ny), wrap1(z + cz, a.nz)),
                          (uint32_t)x, lds[(k++) * nx + x], a.scale);
    }
  } else {
    if (solid && !a.fill_solid) return;
#pragma unroll
    for (int s = 0; s < Q; s++)
      st_row<STORE, NT>(row_of(f1, n, s, a, y, z), (uint32_t)x, solid ? C(0) : d[s], a.scale);
  }
}

// Segmented LDS-staged AA odd step for fully periodic SOA boxes: a workgroup
// is a segment of W = blockDim.x cells of one x-row (W a multiple of 64 that
// divides nx), so the odd step can run the small, capped workgroups that made
// the even step fast, instead of whole rows.  As in k_dense_staged, every
// access of a segment is its aligned element x of some row, and the x-shift of
// the c_x != 0 directions is applied in LDS; what a segment needs beyond its
// own W elements is one element per shifted direction:
//   gather : node x0 (c_x = +1) reads element x0 - 1 and node x0 + W - 1
//            (c_x = -1) element x0 + W of the source row; lane k < NSH loads
//            direction k's element into the LDS halo;
//   scatter: node x0 + W - 1 (c_x = +1) writes element x0 + W and node x0
//            (c_x = -1) element x0 - 1 of the target row; lane k stores it
//            from the LDS halo, and the owner segment skips that element.
// Each such element is read and written only by the one node of this segment
// whose AA read set and write set it belongs to (a node reads and writes the
// same Q slots), so neighbouring segments never touch it: the update is race
// free in place, and its values are the whole-row kernel's bit for bit.  The
// halo element shares a cache line with the neighbouring segment's aligned
// accesses; the segments of one row are placed on one XCD (seg_coords) so
// that the line meets in that XCD's L2, and the halo accesses stay cacheable.
//
// Linear block b -> (row, segment): the hardware hands blocks to the 8 XCDs in
// turn (b mod 8), so with t = b / 8 block b takes segment t mod nseg of row
// (t / nseg) * 8 + b mod 8: the nseg segments of a row are consecutive blocks
// of one XCD.  Needs the number of rows to be a multiple of 8 (else plain
// order: row b / nseg, segment b mod nseg).
__device__ __forceinline__ void seg_coords(int nseg, int nrows, int& row, int& seg) {
  // rows are numbered y + ny * (z - z0) over the planes of the launch
  const int b = blockIdx.x;
  if (nrows % 8 == 0) {
    const int t = b / 8;
    seg = t % nseg;
    row = (t / nseg) * 8 + b % 8;
  } else {
    seg = b % nseg;
    row = b / nseg;
  }
}

template <int Q, int STORE, bool FORCE, bool NT = true,
          bool ONE = one_collision_body<Q, STORE, K_AA_ODD, FORCE>()>
__global__ void __launch_bounds__(256) k_dense_seg_odd(typename Storage<STORE>::T* f, DenseArgs a) {
  using T = typename Storage<STORE>::T;
  using C = typename Storage<STORE>::C;
  constexpr int NSH = n_xshift<Q>();
  extern __shared__ unsigned char lbm_smem[];
  const int W = blockDim.x;
  C* lds = reinterpret_cast<C*>(lbm_smem);  // [NSH][W + 2]: halo, W elements, halo
  const int nx = a.nx;
  const int nseg = nx / W;
  int row, seg;
  seg_coords(nseg, (int)gridDim.x / nseg, row, seg);  // one block per segment of every row
  const int y = row % a.ny, z = row / a.ny + a.z0;
  const int tid = threadIdx.x;
  const int x0 = seg * W, x = x0 + tid;
  const int xl = x0 == 0 ? nx - 1 : x0 - 1, xr = x0 + W == nx ? 0 : x0 + W;  // halo positions
  const int64_t cell = x + (int64_t)nx * (y + (int64_t)a.ny * z);
  const int64_t n = a.ncell;
  const int L = W + 2;

  // Gather.  All loads are issued before any LDS store: lane k < NSH also
  // loads direction k's halo element, through an address selected per lane (a
  // load inside a lane-divergent branch would make the wave wait for every
  // load issued before it).
  const T* hp = f + cell;  // the halo element of lane k < NSH (lanes >= NSH: own element)
  int hpos = 1 + tid;      // its LDS position
  {
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      if (cx != 0) {
        if (tid == k) {
          hp = row_of(f, n, s, a, wrap1(y - cy, a.ny), wrap1(z - cz, a.nz)) + (cx > 0 ? xl : xr);
          hpos = k * L + (cx > 0 ? 0 : W + 1);
        }
        k++;
      }
    }
  }
  const C hv = ld<STORE>(hp, 0, a.inv_scale);
  // keep the halo load ahead of the aligned ones: scheduled after them, its
  // latency would be waited out on its own before the first barrier
  __builtin_amdgcn_sched_barrier(0);
  C d[Q];
  C v[NSH];
  {
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      // element x of slot s in row (y - cy, z - cz)
      const T* rp = row_of(f, n, s, a, wrap1(y - cy, a.ny), wrap1(z - cz, a.nz));
      const C val = ld_row<STORE, NT>(rp, (uint32_t)x, a.inv_scale);
      if (cx == 0) d[s] = val;
      else v[k++] = val;
    }
  }
#pragma unroll
  for (int k = 0; k < NSH; k++) lds[k * L + 1 + tid] = v[k];
  if (tid < NSH) lds[hpos] = hv;
  __syncthreads();
  {
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int cx = Lattice<Q>::c(s, 0);
      if (cx != 0) {
        d[s] = with_jump<Q>(lds[k * L + 1 + tid - cx], gather_crossing(cx, x, nx), s, a.djump);
        k++;
      }
    }
  }
  {
    const double om = a.omega_field ? a.omega_field[cell] : a.omega;
    const double F[3] = {a.F0, a.F1, a.F2};
    collide_cell<Q, STORE, FORCE, ONE>(d, om, F, a.gamma_field, cell);
  }
  __syncthreads();  // every LDS read of the gather is done
  {
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      if (cx == 0) {
        st_row<STORE, NT>(row_of(f, n, o, a, wrap1(y + cy, a.ny), wrap1(z + cz, a.nz)), (uint32_t)x,
                          d[s], a.scale);
      } else {
        lds[k * L + 1 + tid + cx] = with_jump<Q>(d[s], scatter_crossing(cx, x, nx), s, a.djump);
        k++;
      }
    }
  }
  __syncthreads();
  // Scatter: aligned elements, then lane k < NSH stores direction k's halo
  // element (through a per-lane address, as in the gather).
  {
    T* sp = f + cell;
    int spos = 0;
    int k = 0;
#pragma unroll
    for (int s = 0; s < Q; s++) {
      const int o = Lattice<Q>::opp(s);
      const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
      if (cx != 0) {
        T* rp = row_of(f, n, o, a, wrap1(y + cy, a.ny), wrap1(z + cz, a.nz));
        // the element the neighbouring segment's node writes through its halo
        const bool foreign = (cx > 0 && tid == 0) || (cx < 0 && tid == W - 1);
        if (!foreign) st_row<STORE, NT>(rp, (uint32_t)x, lds[k * L + 1 + tid], a.scale);
        if (tid == k) {
          sp = rp + (cx > 0 ? xr : xl);
          spos = k * L + (cx > 0 ? W + 1 : 0);
        }
        k++;
      }
    }
    if (tid < NSH) st<STORE>(sp, 0, lds[spos], a.scale);
  }
}

template <int Q, int STORE>
constexpr size_t seg_lds_bytes(int w) {
  return (size_t)n_xshift<Q>() * (w + 2) * sizeof(typename Storage<STORE>::C);
}

template <int Q, int STORE>
constexpr size_t staged_lds_bytes(int nx) {
  return (size_t)n_xshift<Q>() * nx * sizeof(typename Storage<STORE>::C);
}

// lds_min: dynamic LDS to request per workgroup at least (0 = what the kernel
// needs).  A larger request caps the resident workgroups per CU, which bounds
// the number of concurrently open population streams per CU.
using DenseLaunchFn = void (*)(dim3 grid, dim3 block, hipStream_t stream, void* f0, void* f1,
                               const DenseArgs& a, size_t lds_min);

template <int Q, int STORE, int LAYOUT, int KIND, bool MASK, bool FORCE, bool NT>
void launch_dense(dim3 grid, dim3 block, hipStream_t stream, void* f0, void* f1,
                  const DenseArgs& a, size_t lds_min) {
  using T = typename Storage<STORE>::T;
  hipLaunchKernelGGL((k_dense<Q, STORE, LAYOUT, KIND, MASK, FORCE, NT>), grid, block, lds_min, stream,
                     (const T*)f0, (T*)f1, a);
}

template <int Q, int STORE, bool FORCE, bool NT>
void launch_seg_odd(dim3 grid, dim3 block, hipStream_t stream, void* f0, void* f1,
                    const DenseArgs& a, size_t lds_min) {
  using T = typename Storage<STORE>::T;
  (void)f0;  // in place: f0 == f1
  size_t lds = seg_lds_bytes<Q, STORE>((int)block.x);
  if (lds < lds_min) lds = lds_min;
  hipLaunchKernelGGL((k_dense_seg_odd<Q, STORE, FORCE, NT>), grid, block, lds, stream, (T*)f1, a);
}

template <int Q, bool NT>
DenseLaunchFn pick_seg_odd_qn(int storage, bool force) {
  switch (storage) {
    case LBM_STORE_F64:
      return force ? launch_seg_odd<Q, LBM_STORE_F64, true, NT>
                   : launch_seg_odd<Q, LBM_STORE_F64, false, NT>;
    case LBM_STORE_F32S:
      return force ? launch_seg_odd<Q, LBM_STORE_F32S, true, NT>
                   : launch_seg_odd<Q, LBM_STORE_F32S, false, NT>;
    case LBM_STORE_F16S:
      return force ? launch_seg_odd<Q, LBM_STORE_F16S, true, NT>
                   : launch_seg_odd<Q, LBM_STORE_F16S, false, NT>;
  }
  return nullptr;
}

template <int Q>
DenseLaunchFn pick_seg_odd_q(int storage, bool force, bool nt) {
  return nt ? pick_seg_odd_qn<Q, true>(storage, force) : pick_seg_odd_qn<Q, false>(storage, force);
}

template <int Q, int STORE, int KIND, bool FORCE, bool MASK, bool NT>
void launch_staged(dim3 grid, dim3 block, hipStream_t stream, void* f0, void* f1,
                   const DenseArgs& a, size_t lds_min) {
  using T = typename Storage<STORE>::T;
  size_t lds = staged_lds_bytes<Q, STORE>(a.nx);
  if (lds < lds_min) lds = lds_min;
  hipLaunchKernelGGL((k_dense_staged<Q, STORE, KIND, FORCE, MASK, NT>), grid, block, lds, stream,
                     (const T*)f0, (T*)f1, a);
}

template <int Q, int STORE, int KIND, bool MASK, bool NT>
DenseLaunchFn pick_staged_force(bool force) {
  return force ? launch_staged<Q, STORE, KIND, true, MASK, NT>
               : launch_staged<Q, STORE, KIND, false, MASK, NT>;
}

template <int Q, int STORE, bool NT>
DenseLaunchFn pick_staged_kind(int kind, bool force, bool mask) {
  if (mask)
    return kind == K_AB_PULL ? pick_staged_force<Q, STORE, K_AB_PULL, true, NT>(force) : nullptr;
  switch (kind) {
    case K_AA_ODD: return pick_staged_force<Q, STORE, K_AA_ODD, false, NT>(force);
    case K_AB_PUSH: return pick_staged_force<Q, STORE, K_AB_PUSH, false, NT>(force);
    case K_AB_PULL: return pick_staged_force<Q, STORE, K_AB_PULL, false, NT>(force);
  }
  return nullptr;
}

// LDS-staged kernel for (storage, kind, mask), or nullptr when the staged form
// does not apply (AA_EVEN needs no staging; with a mask only the pull gather
// is staged) or its LDS would exceed lds_max, the device's limit per workgroup
// (160 KiB on gfx950: a D3Q27 FP64 row of 512 needs 72 KiB).
template <int Q, bool NT>
DenseLaunchFn pick_staged_qn(int storage, int kind, bool force, int nx, bool mask, size_t lds_max) {
  if (kind == K_AA_EVEN) return nullptr;
  switch (storage) {
    case LBM_STORE_F64:
      if (staged_lds_bytes<Q, LBM_STORE_F64>(nx) > lds_max) return nullptr;
      return pick_staged_kind<Q, LBM_STORE_F64, NT>(kind, force, mask);
    case LBM_STORE_F32S:
      if (staged_lds_bytes<Q, LBM_STORE_F32S>(nx) > lds_max) return nullptr;
      return pick_staged_kind<Q, LBM_STORE_F32S, NT>(kind, force, mask);
    case LBM_STORE_F16S:
      if (staged_lds_bytes<Q, LBM_STORE_F16S>(nx) > lds_max) return nullptr;
      return pick_staged_kind<Q, LBM_STORE_F16S, NT>(kind, force, mask);
  }
  return nullptr;
}

template <int Q>
DenseLaunchFn pick_staged_q(int storage, int kind, bool force, int nx, bool mask, size_t lds_max,
                            bool nt) {
  return nt ? pick_staged_qn<Q, true>(storage, kind, force, nx, mask, lds_max)
            : pick_staged_qn<Q, false>(storage, kind, force, nx, mask, lds_max);
}

template <int Q, int STORE, int LAYOUT, int KIND, bool NT>
DenseLaunchFn pick_flags(bool mask, bool force) {
  if (mask) return force ? launch_dense<Q, STORE, LAYOUT, KIND, true, true, NT>
                         : launch_dense<Q, STORE, LAYOUT, KIND, true, false, NT>;
  return force ? launch_dense<Q, STORE, LAYOUT, KIND, false, true, NT>
               : launch_dense<Q, STORE, LAYOUT, KIND, false, false, NT>;
}

template <int Q, int STORE, int LAYOUT, bool NT>
DenseLaunchFn pick_kind(int kind, bool mask, bool force) {
  switch (kind) {
    case K_AA_EVEN: return pick_flags<Q, STORE, LAYOUT, K_AA_EVEN, NT>(mask, force);
    case K_AA_ODD: return pick_flags<Q, STORE, LAYOUT, K_AA_ODD, NT>(mask, force);
    case K_AB_PUSH: return pick_flags<Q, STORE, LAYOUT, K_AB_PUSH, NT>(mask, force);
    case K_AB_PULL: return pick_flags<Q, STORE, LAYOUT, K_AB_PULL, NT>(mask, force);
    case K_AA_ODD_IR:
      if constexpr (LAYOUT == LBM_LAYOUT_SOA) {
        if (mask) return nullptr;
        return force ? launch_dense<Q, STORE, LAYOUT, K_AA_ODD_IR, false, true, NT>
                     : launch_dense<Q, STORE, LAYOUT, K_AA_ODD_IR, false, false, NT>;
      }
      return nullptr;
  }
  return nullptr;
}

template <int Q, bool NT>
DenseLaunchFn pick_dense_qn(int storage, int layout, int kind, bool mask, bool force) {
  if (layout == LBM_LAYOUT_SOA) {
    switch (storage) {
      case LBM_STORE_F64: return pick_kind<Q, LBM_STORE_F64, LBM_LAYOUT_SOA, NT>(kind, mask, force);
      case LBM_STORE_F32S: return pick_kind<Q, LBM_STORE_F32S, LBM_LAYOUT_SOA, NT>(kind, mask, force);
      case LBM_STORE_F16S: return pick_kind<Q, LBM_STORE_F16S, LBM_LAYOUT_SOA, NT>(kind, mask, force);
    }
  } else {
    switch (storage) {
      case LBM_STORE_F64: return pick_kind<Q, LBM_STORE_F64, LBM_LAYOUT_AOS, NT>(kind, mask, force);
      case LBM_STORE_F32S: return pick_kind<Q, LBM_STORE_F32S, LBM_LAYOUT_AOS, NT>(kind, mask, force);
      case LBM_STORE_F16S: return pick_kind<Q, LBM_STORE_F16S, LBM_LAYOUT_AOS, NT>(kind, mask, force);
    }
  }
  return nullptr;
}

template <int Q>
DenseLaunchFn pick_dense_q(int storage, int layout, int kind, bool mask, bool force, bool nt) {
  return nt ? pick_dense_qn<Q, true>(storage, layout, kind, mask, force)
            : pick_dense_qn<Q, false>(storage, layout, kind, mask, force);
}

// Defined in lbm_dense_q19.hip and lbm_dense_q27.hip (one translation unit per Q).
// force = false selects the zero-body-force kernels (only valid for F = 0).
DenseLaunchFn pick_dense19(int storage, int layout, int kind, bool mask, bool force, bool nt);
DenseLaunchFn pick_dense27(int storage, int layout, int kind, bool mask, bool force, bool nt);
DenseLaunchFn pick_staged19(int storage, int kind, bool force, int nx, bool mask, size_t lds_max,
                            bool nt);
DenseLaunchFn pick_staged27(int storage, int kind, bool force, int nx, bool mask, size_t lds_max,
                            bool nt);
// segmented LDS-staged AA odd step (k_dense_seg_odd): grid = one block per row segment
DenseLaunchFn pick_seg_odd19(int storage, bool force, bool nt);
DenseLaunchFn pick_seg_odd27(int storage, bool force, bool nt);

}  // namespace lbm
