// lbm_dense.hip -- host side of the structured-grid engine: handle, link-mask
// builder, state initialisation and readout, moments, FP64 mass reduction,
// and the time loop (the replacement of PropagationAA::run / PropagationAB::run,
// src/propagation/PropagationAA.cpp:126-215, src/propagation/PropagationAB.cpp:155-219).
//
// State conventions (pre-collision populations f_s(x) at time t):
//   AA, t even      f_s(x) in slot opp(s) at x (src/propagation/PropagationAA.cpp:111)
//   AA, t odd       f_s(x) in slot s at x - c_s, or slot opp(s) at x when x - c_s is
//                   solid (src/propagation/PropagationAA.cpp:243-257)
//   AB push         f_s(x) in slot s at x of the current array
//   AB pull         the current array holds post-collision values; f_s(x) is gathered
//                   like the AA odd case (src/kernels/KernelAB_SOA_Pull.cpp:139-140)
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "lbm_dense_internal.hpp"

namespace lbm {

enum CanonMode { C_LOCAL = 0, C_OPP = 1, C_GATHER = 2 };

// storage cell of cell = (x, y, z): on interleaved rows element x of the row
// sits at position ir_pos(x)
__device__ __forceinline__ int64_t store_cell(const DenseArgs& a, int64_t cell, int x) {
  return a.xk ? cell - x + ir_pos(x, a.xk) : cell;
}

// storage index of f_s(x) for a given state convention
template <int Q, int LAYOUT>
__device__ __forceinline__ int64_t canon_index(int mode, int s, int64_t cell, int x, uint32_t lk,
                                               const Offsets& off, const DenseArgs& a) {
  const int64_t n = a.ncell;
  const int o = Lattice<Q>::opp(s);
  if (mode == C_LOCAL) return slot<Q, LAYOUT>(n, s, store_cell(a, cell, x));
  if (mode == C_OPP || ((lk >> o) & 1u)) return slot<Q, LAYOUT>(n, o, store_cell(a, cell, x));
  const int cx = Lattice<Q>::c(s, 0);
  return slot<Q, LAYOUT>(
      n, s,
      store_cell(a, cell + off.at(-cx, -Lattice<Q>::c(s, 1), -Lattice<Q>::c(s, 2)),
                 wrap1(x - cx, a.nx)));
}

__device__ __forceinline__ bool cell_coords(const DenseArgs& a, int& x, int& y, int& z,
                                            int64_t& cell) {
  x = blockIdx.x * blockDim.x + threadIdx.x;
  y = blockIdx.y;
  z = blockIdx.z;
  if (x >= a.nx) return false;
  cell = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
  return true;
}

// ---------------------------------------------------------------- link mask
template <int Q>
__global__ void __launch_bounds__(256)
    k_build_links(const uint8_t* __restrict__ solid, uint32_t* __restrict__ links, int nx, int ny,
                  int nz, int px, int py, int pz, int xk) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const int64_t cell = x + (int64_t)nx * (y + (int64_t)ny * z);
  // bit s: x + c_s is solid or outside; solid cells carry the same bits (they
  // place the zeros of the whole-line stores) plus SOLID_BIT
  uint32_t m = solid[cell] ? SOLID_BIT : 0u;
  for (int s = 0; s < Q; s++) {
    int p[3] = {x + Lattice<Q>::c(s, 0), y + Lattice<Q>::c(s, 1), z + Lattice<Q>::c(s, 2)};
    const int dims[3] = {nx, ny, nz};
    const int per[3] = {px, py, pz};
    bool wall = false;
    for (int k = 0; k < 3; k++) {
      if (p[k] < 0 || p[k] >= dims[k]) {
        if (!per[k]) wall = true;
        else p[k] = p[k] < 0 ? p[k] + dims[k] : p[k] - dims[k];
      }
    }
    if (!wall) wall = solid[p[0] + (int64_t)nx * (p[1] + (int64_t)ny * p[2])] != 0;
    if (wall) m |= 1u << s;
  }
  // stored like the populations: at the cell's position on interleaved rows
  links[xk ? cell - x + ir_pos(x, xk) : cell] = m;
}

// ---------------------------------------------------------------- state in / out
// Sign of the pressure-drop jump between the stored value at canon_index and
// the canonical f_s(x): non-zero only where a gathering convention reads
// across the periodic x boundary (canonical = stored + sign * w_s * djump).
template <int Q>
__device__ __forceinline__ int canon_jump_sign(int mode, int s, int x, uint32_t lk, int nx) {
  if (mode != C_GATHER || ((lk >> Lattice<Q>::opp(s)) & 1u)) return 0;
  return gather_crossing(Lattice<Q>::c(s, 0), x, nx);
}

// f_s(x) -> storage, for every fluid cell; src is canonical [Q][ncell] (device).
template <int Q, int STORE, int LAYOUT>
__global__ void __launch_bounds__(256)
    k_set_canonical(const double* __restrict__ src, void* fv, DenseArgs a, int mode) {
  auto* f = static_cast<typename Storage<STORE>::T*>(fv);
  int x, y, z;
  int64_t cell;
  if (!cell_coords(a, x, y, z, cell)) return;
  const uint32_t lk = a.links ? a.links[store_cell(a, cell, x)] : 0u;
  if (lk & SOLID_BIT) return;
  const Offsets off(x, y, z, a.nx, a.ny, a.nz);
  for (int s = 0; s < Q; s++) {
    double v = with_jump<Q>(src[a.ncell * s + cell], -canon_jump_sign<Q>(mode, s, x, lk, a.nx), s,
                            a.djump);
    if constexpr (STORE != LBM_STORE_F64) v -= Lattice<Q>::w(s);
    st<STORE>(f, canon_index<Q, LAYOUT>(mode, s, cell, x, lk, off, a),
              (typename Storage<STORE>::C)v, a.scale);
  }
}

template <int STORE>
__device__ __forceinline__ double to_f(typename Storage<STORE>::C v, double w) {
  if constexpr (STORE == LBM_STORE_F64) return v;
  else return w + (double)v;
}

template <int Q, int STORE, int LAYOUT>
__global__ void __launch_bounds__(256)
    k_get_canonical(const void* fv, double* __restrict__ dst, DenseArgs a, int mode) {
  const auto* f = static_cast<const typename Storage<STORE>::T*>(fv);
  int x, y, z;
  int64_t cell;
  if (!cell_coords(a, x, y, z, cell)) return;
  const uint32_t lk = a.links ? a.links[store_cell(a, cell, x)] : 0u;
  const Offsets off(x, y, z, a.nx, a.ny, a.nz);
  for (int s = 0; s < Q; s++) {
    double v = 0.0;
    if (!(lk & SOLID_BIT))
      v = with_jump<Q>(
          to_f<STORE>(ld<STORE>(f, canon_index<Q, LAYOUT>(mode, s, cell, x, lk, off, a),
                                a.inv_scale),
                      Lattice<Q>::w(s)),
          canon_jump_sign<Q>(mode, s, x, lk, a.nx), s, a.djump);
    dst[a.ncell * s + cell] = v;
  }
}

// rho, u = m / rho of the pre-collision state (reference readout: no F/2 shift,
// src/propagation/PropagationAA.cpp:260-327).
template <int Q, int STORE, int LAYOUT>
__global__ void __launch_bounds__(256)
    k_moments(const void* fv, double* __restrict__ rho_out, double* __restrict__ u_out,
              DenseArgs a, int mode) {
  const auto* f = static_cast<const typename Storage<STORE>::T*>(fv);
  int x, y, z;
  int64_t cell;
  if (!cell_coords(a, x, y, z, cell)) return;
  const uint32_t lk = a.links ? a.links[store_cell(a, cell, x)] : 0u;
  double rho = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
  if (!(lk & SOLID_BIT)) {
    const Offsets off(x, y, z, a.nx, a.ny, a.nz);
    for (int s = 0; s < Q; s++) {
      const double v = with_jump<Q>(
          to_f<STORE>(ld<STORE>(f, canon_index<Q, LAYOUT>(mode, s, cell, x, lk, off, a),
                                a.inv_scale),
                      Lattice<Q>::w(s)),
          canon_jump_sign<Q>(mode, s, x, lk, a.nx), s, a.djump);
      rho += v;
      m0 += v * Lattice<Q>::c(s, 0);
      m1 += v * Lattice<Q>::c(s, 1);
      m2 += v * Lattice<Q>::c(s, 2);
    }
    const double inv = 1.0 / rho;
    m0 *= inv;
    m1 *= inv;
    m2 *= inv;
  }
  rho_out[cell] = rho;
  u_out[3 * cell + 0] = m0;
  u_out[3 * cell + 1] = m1;
  u_out[3 * cell + 2] = m2;
}

// ---------------------------------------------------------------- initial states
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t counter) {
  const uint64_t key = seed * 0x9E3779B97F4A7C15ull + counter;
  return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

// src_kind 0: rest; 1: rho/u device arrays; 2: Taylor-Green vortex + seeded noise.
template <int Q, int STORE, int LAYOUT>
__global__ void __launch_bounds__(256)
    k_init(void* fv, DenseArgs a, int mode, int src_kind,
           const double* __restrict__ rho_in, const double* __restrict__ u_in, double u0,
           double noise, uint64_t seed) {
  auto* f = static_cast<typename Storage<STORE>::T*>(fv);
  int x, y, z;
  int64_t cell;
  if (!cell_coords(a, x, y, z, cell)) return;
  const uint32_t lk = a.links ? a.links[store_cell(a, cell, x)] : 0u;
  if (lk & SOLID_BIT) return;
  double rho = 1.0, u[3] = {0.0, 0.0, 0.0};
  if (src_kind == 1) {
    if (rho_in) rho = rho_in[cell];
    for (int k = 0; k < 3; k++) u[k] = u_in[3 * cell + k];
  } else if (src_kind == 2) {
    const double kx = 2.0 * M_PI / a.nx, ky = 2.0 * M_PI / a.ny, kz = 2.0 * M_PI / a.nz;
    u[0] = u0 * sin(kx * x) * cos(ky * y) * cos(kz * z);
    u[1] = -u0 * cos(kx * x) * sin(ky * y) * cos(kz * z);
    if (noise != 0.0)
      for (int k = 0; k < 3; k++)
        u[k] += noise * (2.0 * uniform01(seed, 3ull * (uint64_t)cell + k) - 1.0);
  }
  const Offsets off(x, y, z, a.nx, a.ny, a.nz);
  // f^eq with the collision's association (src/kernels/KernelAA_SOA_Base.cpp:227-238, F = 0)
  const double vdotv = 3.0 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int s = 0; s < Q; s++) {
    const double c0 = Lattice<Q>::c(s, 0), c1 = Lattice<Q>::c(s, 1), c2 = Lattice<Q>::c(s, 2);
    const double cdotv = 3.0 * (c0 * u[0] + c1 * u[1] + c2 * u[2]);
    const double w = Lattice<Q>::w(s);
    double feq = w * rho * (1.0 + cdotv + 0.5 * cdotv * cdotv - 0.5 * vdotv);
    if constexpr (STORE != LBM_STORE_F64) feq = w * ((rho - 1.0) + rho * (cdotv + 0.5 * cdotv * cdotv - 0.5 * vdotv));
    feq = with_jump<Q>(feq, -canon_jump_sign<Q>(mode, s, x, lk, a.nx), s, a.djump);
    st<STORE>(f, canon_index<Q, LAYOUT>(mode, s, cell, x, lk, off, a),
              (typename Storage<STORE>::C)feq, a.scale);
  }
}

// ---------------------------------------------------------------- mass
// Sum over every fluid cell of its Q stored values (any state convention is a
// permutation of the canonical populations over the fluid slots).
template <int Q, int STORE, int LAYOUT>
__global__ void __launch_bounds__(256)
    k_mass_partial(const void* fv, DenseArgs a, double* partial) {
  const auto* f = static_cast<const typename Storage<STORE>::T*>(fv);
  __shared__ double red[256];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cell < a.ncell;
       cell += stride) {
    if (a.links && (a.links[cell] & SOLID_BIT)) continue;
    for (int s = 0; s < Q; s++)
      acc += to_f<STORE>(ld<STORE>(f, slot<Q, LAYOUT>(a.ncell, s, cell), a.inv_scale),
                         Lattice<Q>::w(s));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* partial, int n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------- dispatch helpers
template <template <int, int, int> class K, typename... A>
void launch_util(int q, int storage, int layout, dim3 g, dim3 b, hipStream_t st, A... args) {
#define LBM_UTIL_CASE(QQ, SS, LL) \
  if (q == QQ && storage == SS && layout == LL) return K<QQ, SS, LL>::launch(g, b, st, args...);
#define LBM_UTIL_Q(QQ)                                      \
  LBM_UTIL_CASE(QQ, LBM_STORE_F64, LBM_LAYOUT_SOA)          \
  LBM_UTIL_CASE(QQ, LBM_STORE_F32S, LBM_LAYOUT_SOA)         \
  LBM_UTIL_CASE(QQ, LBM_STORE_F16S, LBM_LAYOUT_SOA)         \
  LBM_UTIL_CASE(QQ, LBM_STORE_F64, LBM_LAYOUT_AOS)          \
  LBM_UTIL_CASE(QQ, LBM_STORE_F32S, LBM_LAYOUT_AOS)         \
  LBM_UTIL_CASE(QQ, LBM_STORE_F16S, LBM_LAYOUT_AOS)
  LBM_UTIL_Q(19)
  LBM_UTIL_Q(27)
#undef LBM_UTIL_Q
#undef LBM_UTIL_CASE
}

#define LBM_UTIL_KERNEL(name)                                                             \
  template <int Q, int S, int L>                                                          \
  struct name##_u {                                                                       \
    template <typename... A>                                                              \
    static void launch(dim3 g, dim3 b, hipStream_t st, A... a) {                          \
      hipLaunchKernelGGL((name<Q, S, L>), g, b, 0, st, a...);                             \
    }                                                                                     \
  };
LBM_UTIL_KERNEL(k_set_canonical)
LBM_UTIL_KERNEL(k_get_canonical)
LBM_UTIL_KERNEL(k_moments)
LBM_UTIL_KERNEL(k_init)
LBM_UTIL_KERNEL(k_mass_partial)
#undef LBM_UTIL_KERNEL

}  // namespace lbm

using namespace lbm;


namespace lbm {

// Utility kernels: 256-wide x segments.
dim3 block_for(int nx) {
  int bx = nx >= 256 ? 256 : ((nx + 63) / 64) * 64;
  return dim3(bx);
}

dim3 grid_for(const lbm_dense* h) {
  const dim3 b = block_for(h->nx);
  return dim3((h->nx + b.x - 1) / b.x, h->ny, h->nz);
}

// Default block width of the unstaged time-step kernels: a whole x-row per
// workgroup when it fits (kDenseMaxBlock), else 256-wide segments.  FP32
// deviation storage measured faster with 256-wide segments when they tile the
// row exactly (D3Q19 AA at nx = 512: 35.6k vs 31.8k MLUPS); otherwise whole
// rows (porous push at nx = 384: 2.50 vs 2.95 ms per step).
int default_step_bx(int nx, int storage) {
  if (nx <= kDenseMaxBlock && !(storage == LBM_STORE_F32S && nx % 256 == 0))
    return ((nx + 63) / 64) * 64;
  return 256;
}

// Dynamic LDS that leaves room for at most k workgroups per CU (160 KiB LDS).
size_t lds_for_blocks_per_cu(int k) {
  return k > 0 ? (size_t)(160 * 1024 / k) & ~(size_t)1023 : (size_t)0;
}

// Launch shapes, fixed at create.  The SOA AA even step reads and writes only
// row-aligned slots, so it needs no whole rows; it measured fastest with
// small workgroups at low occupancy, like the 19-stream in-place copy pattern,
// whose DRAM efficiency drops as more streams are open per CU.  512^3 even
// step, best of a sweep over widths 64-512 and 0-8 workgroups per CU:
//   D3Q19 FP64   128 wide, <= 3 per CU: 7.09-7.12 ms per step, even and odd
//                averaged over 100-step runs alternating with <= 4 per CU
//                (7.16-7.20), <= 5 (7.15-7.20) and <= 2 (7.09-7.13); single
//                even steps 6.91 ms (512-wide rows: 7.86 ms)
//   D3Q19 FP32   256 wide, <= 4 per CU: 3.39 ms (256 wide, no cap: 3.47 ms)
//   D3Q19 FP16   512 wide, <= 4 per CU: 1.74 ms per step, even and odd
//                averaged over 100-step runs with the interleaved-row odd
//                step (256 wide, no cap: 1.76-1.78; 128 wide: 1.84-1.85)
//   D3Q27 FP64   128 wide, <= 2 per CU: 9.75-9.83 ms alternating with 256 wide,
//                <= 2 per CU (10.30-10.33; step pairs 10.67-10.94 vs 10.86-11.28
//                ms per step); 512-wide rows: 10.83 ms
//   D3Q27 FP32   256 wide, no cap: 4.92-4.95 ms per step, even and odd
//                averaged over 100-step runs on plain rows, alternating with
//                128 wide, no cap (5.01-5.07; in single launches right after
//                initialisation 128 wide had ranked first, 4.69 vs 4.86 ms)
//   D3Q27 FP16   512 wide, <= 4 per CU: 2.49-2.50 ms per step, even and odd
//                averaged over 100-step runs with the interleaved-row odd
//                step (128 wide, <= 8 per CU: 2.61-2.67; 256 / 4: 2.53)
// The first handle of a process measured faster than later ones, so shape
// comparisons alternate the candidates, none of them first.
// Environment overrides (tuning): LBM_DENSE_BLOCK_X and LBM_DENSE_BLOCKS_PER_CU
// for the unstaged kernels (the latter also for the staged ones),
// LBM_DENSE_EVEN_BLOCK_X and LBM_DENSE_EVEN_BLOCKS_PER_CU for the even step.
void set_launch_shapes(lbm_dense* h) {
  auto env_int = [](const char* name, int dflt) {
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
  };
  auto clamp_bx = [](int bx) { return bx < 64 ? 64 : (bx > kDenseMaxBlock ? kDenseMaxBlock : bx); };
  h->step_bx = clamp_bx(env_int("LBM_DENSE_BLOCK_X", default_step_bx(h->nx, h->storage)));
  h->step_lds = lds_for_blocks_per_cu(env_int("LBM_DENSE_BLOCKS_PER_CU", 0));
  int even_bx = h->step_bx, even_k = 0, odd_ir_bx = h->step_bx, odd_ir_k = 0;
  if (h->layout == LBM_LAYOUT_SOA) {
    static const int shape[2][3][2] = {
        // {width, workgroups per CU} for F64, F32S, F16S
        {{128, 3}, {256, 4}, {512, 4}},  // D3Q19
        {{128, 2}, {256, 0}, {512, 4}},  // D3Q27
    };
    const int* sh = shape[h->q == 19 ? 0 : 1][h->storage];
    even_bx = sh[0];
    even_k = sh[1];
    // The AA odd step on interleaved rows makes only whole aligned block
    // accesses, like the even step.  512^3, best of widths 64-512 and 0-8
    // workgroups per CU; ms per step over 100-step runs, interleaved / plain
    // rows (the plain-row odd step staged, segmented or unstaged as set
    // below):
    //   D3Q19 FP64   64 wide, <= 4 per CU: 7.11-7.14 / 7.20-7.24
    //   D3Q19 FP32  256 wide, <= 4 per CU: 3.42-3.44 / 3.41-3.44
    //   D3Q19 FP16  512 wide, no cap:      1.77-1.78 / 1.89-1.90
    //   D3Q27 FP64  128 wide, <= 4 per CU: 10.62-10.77 / 10.75-10.96
    //   D3Q27 FP16  512 wide, no cap:      2.68-2.69 / 2.87-2.89
    // D3Q27 FP32 ran slower on interleaved rows (5.16-5.18 / 4.96-5.02) and
    // keeps plain rows (dense_create).
    static const int odd_shape[2][3][2] = {
        {{64, 4}, {256, 4}, {512, 0}},   // D3Q19
        {{128, 4}, {128, 4}, {512, 0}},  // D3Q27
    };
    const int* so = odd_shape[h->q == 19 ? 0 : 1][h->storage];
    odd_ir_bx = so[0];
    odd_ir_k = so[1];
  }
  // Non-temporal streaming pays when the populations exceed the 256 MiB
  // memory-side cache; a grid that fits there runs cached (LBM_DENSE_NT=0|1
  // overrides).
  {
    const size_t pop_bytes =
        (h->pattern == LBM_PATTERN_AB ? 2 : 1) * (size_t)h->ncell * h->q * h->esize;
    h->nt = env_int("LBM_DENSE_NT", pop_bytes > ((size_t)256 << 20) ? 1 : 0) != 0;
  }
  int dev = 0, lds = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      lds > 0)
    h->lds_max = (size_t)lds;
  h->even_bx = clamp_bx(env_int("LBM_DENSE_EVEN_BLOCK_X", even_bx));
  h->even_lds = lds_for_blocks_per_cu(env_int("LBM_DENSE_EVEN_BLOCKS_PER_CU", even_k));
  // LBM_DENSE_ODD_IR_BLOCK_X, LBM_DENSE_ODD_IR_BLOCKS_PER_CU: tuning
  h->odd_ir_bx = clamp_bx(env_int("LBM_DENSE_ODD_IR_BLOCK_X", odd_ir_bx));
  h->odd_ir_lds = lds_for_blocks_per_cu(env_int("LBM_DENSE_ODD_IR_BLOCKS_PER_CU", odd_ir_k));
  // Segmented odd step (k_dense_seg_odd): unmasked periodic SOA boxes whose
  // rows split into segments of a multiple of 64 cells.  512^3 D3Q19 odd step,
  // ms (whole-row staged / best segmented shape): FP32 deviations 3.55 / 3.34
  // (256 wide, <= 4 per CU), so FP32 rows of 512 or more take it; FP64 7.46 /
  // 7.86 and FP16 1.99 (unstaged) / 2.74 stay as they are.  LBM_DENSE_ODD_SEG
  // = width (0 off) and LBM_DENSE_ODD_SEG_BLOCKS_PER_CU override (tuning).  The
  // mask is known only later in create, so dense_create clears it for masked
  // grids.
  {
    const bool dflt = h->storage == LBM_STORE_F32S && h->q == 19 && h->nx >= 512 &&
                      h->nx % 256 == 0;
    const int w = env_int("LBM_DENSE_ODD_SEG", dflt ? 256 : 0);
    const bool ok = h->layout == LBM_LAYOUT_SOA && h->pattern == LBM_PATTERN_AA && w >= 64 &&
                    w <= 256 && w % 64 == 0 && h->nx % w == 0;
    h->odd_seg_w = ok ? w : 0;
    h->odd_seg_lds = lds_for_blocks_per_cu(env_int("LBM_DENSE_ODD_SEG_BLOCKS_PER_CU", dflt ? 4 : 0));
  }
}

// state convention of the current array
int canon_mode(const lbm_dense* h) {
  if (h->pattern == LBM_PATTERN_AA) return (h->t % 2 == 0) ? C_OPP : C_GATHER;
  return h->pull ? C_GATHER : C_LOCAL;
}

// kernels with the Guo forcing terms: a nonzero force, or the policy set at create
bool use_force_terms(const lbm_dense* h) {
  return h->force_terms || h->args.F0 != 0.0 || h->args.F1 != 0.0 || h->args.F2 != 0.0;
}

DenseLaunchFn pick_step(const lbm_dense* h, int kind) {
  const bool mask = h->links != nullptr;
  const bool force = use_force_terms(h);
  return h->q == 19 ? pick_dense19(h->storage, h->layout, kind, mask, force, h->nt)
                    : pick_dense27(h->storage, h->layout, kind, mask, force, h->nt);
}

// The LDS-staged kernels cover SOA grids whose rows fit one workgroup (nx a
// multiple of 64, nx <= kDenseMaxBlock): every kind on fully periodic boxes,
// and the pull gather also with solid cells and walls.  By default (staged
// mode 1, set at create from LBM_DENSE_STAGED) they run where they measured
// faster: 512^3 steps, staged vs unstaged handles alternating in one process,
// ms per step (unmasked periodic boxes):
//               AA (odd staged)     AB push            AB pull
//   D3Q19 FP64  staged (7.68/8.47)  8.46 / 8.93        8.10 / 7.83 unstaged
//   D3Q19 FP32  staged (3.64/3.79)  4.23 / 4.94        3.86 / 3.56 unstaged
//   D3Q19 FP16  1.94 / 1.91 unst.   1.86 / 2.05        1.90 / 2.05
//   D3Q27 FP64  10.65 / 11.35       11.65 / 12.69      11.71 / 10.77 unstaged
//   D3Q27 FP32  4.64 / 4.90         5.96 / 6.45        5.75 / 5.29 unstaged
//   D3Q27 FP16  unstaged (kernel alone 3.08 / 2.90)  push 2.81 / 3.27  pull unstaged (3.12 / 3.02)
// Masked gathers run unstaged (35% solid porous pull at 384^3: 20.4k vs 18.7k
// MLUPS staged; a 5% solid channel: equal).  Mode 0 never stages, mode 2
// stages every case covered (tuning and cross-checks).
DenseLaunchFn pick_staged(const lbm_dense* h, int kind) {
  const int mode = h->staged_mode;
  const bool mask = h->links != nullptr;
  if (mode == 0 || h->layout != LBM_LAYOUT_SOA || h->nx % 64 != 0 || h->nx > kDenseMaxBlock ||
      h->xk)
    return nullptr;
  if (mode < 2) {
    const bool f16 = h->storage == LBM_STORE_F16S;
    const bool staged = kind == K_AB_PUSH || (kind == K_AA_ODD && !f16) ||
                        (kind == K_AB_PULL && f16 && h->q == 19);
    if (mask || !staged) return nullptr;
  }
  const bool force = use_force_terms(h);
  return h->q == 19 ? pick_staged19(h->storage, kind, force, h->nx, mask, h->lds_max, h->nt)
                    : pick_staged27(h->storage, kind, force, h->nx, mask, h->lds_max, h->nt);
}

lbm_status launch_kind(lbm_dense* h, int kind, void* f0, void* f1, int z0, int nzl) {
  if (nzl < 0) {
    z0 = 0;
    nzl = h->nz;
  }
  DenseArgs a = h->args;
  a.z0 = z0;
  if (kind == K_AA_ODD && h->xk && !h->links) {
    DenseLaunchFn fn = pick_step(h, K_AA_ODD_IR);
    LBM_REQUIRE(fn, LBM_ERR_UNSUPPORTED, "no interleaved-row odd kernel for this configuration");
    const dim3 b(h->odd_ir_bx);
    fn(dim3((h->nx + b.x - 1) / b.x, h->ny, nzl), b, h->stream, f0, f1, a, h->odd_ir_lds);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
  }
  if (kind == K_AA_ODD && h->odd_seg_w > 0) {
    const int w = h->odd_seg_w;
    const bool force = use_force_terms(h);
    DenseLaunchFn fn = h->q == 19 ? pick_seg_odd19(h->storage, force, h->nt)
                                  : pick_seg_odd27(h->storage, force, h->nt);
    LBM_REQUIRE(fn, LBM_ERR_UNSUPPORTED, "no segmented odd kernel for this configuration");
    fn(dim3((unsigned)((h->nx / w) * h->ny * nzl)), dim3(w), h->stream, f0, f1, a, h->odd_seg_lds);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
  }
  if (DenseLaunchFn st = pick_staged(h, kind)) {
    st(dim3(1, h->ny, nzl), dim3(h->nx), h->stream, f0, f1, a, h->step_lds);
  } else {
    DenseLaunchFn fn = pick_step(h, kind);
    LBM_REQUIRE(fn, LBM_ERR_UNSUPPORTED, "no kernel for this configuration");
    const bool even = kind == K_AA_EVEN;
    const dim3 b(even ? h->even_bx : h->step_bx);
    const dim3 g((h->nx + b.x - 1) / b.x, h->ny, nzl);
    fn(g, b, h->stream, f0, f1, a, even ? h->even_lds : h->step_lds);
  }
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

// One AA step pair (even, then odd) as z-slabs of P planes, so that the odd
// step of a slab runs right after the even steps it reads, while their planes
// are still in the memory-side cache.  The odd step at z reads and writes
// slots of planes z - 1 .. z + 1, all of which the even step must have
// updated first (and no odd step may precede the even step of a plane it
// writes).  Order with S slabs: even(S-1), even(0), then for k = 0 .. S-2
// even(k+1), odd(k), then odd(S-1).  Every node still gets one even and one
// odd update with the same arithmetic, so the result equals the whole-box
// launches bit for bit.
lbm_status enqueue_aa_pair_by_slabs(lbm_dense* h) {
  const int P = h->slab_planes, S = (h->nz + P - 1) / P;
  void* f = h->f[0];
  auto run = [&](int kind, int k) {
    const int z0 = k * P;
    return launch_kind(h, kind, f, f, z0, std::min(P, h->nz - z0));
  };
  if (lbm_status st = run(K_AA_EVEN, S - 1)) return st;
  if (S > 1)
    if (lbm_status st = run(K_AA_EVEN, 0)) return st;
  for (int k = 0; k + 1 < S; k++) {
    if (k + 1 < S - 1)
      if (lbm_status st = run(K_AA_EVEN, k + 1)) return st;
    if (lbm_status st = run(K_AA_ODD, k)) return st;
  }
  return run(K_AA_ODD, S - 1);
}

lbm_status enqueue_steps(lbm_dense* h, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    lbm_status st;
    if (h->pattern == LBM_PATTERN_AA && h->slab_planes > 0 && h->t % 2 == 0 && i + 1 < nsteps) {
      st = enqueue_aa_pair_by_slabs(h);
      if (st != LBM_OK) return st;
      h->t += 2;
      i++;
      continue;
    }
    if (h->pattern == LBM_PATTERN_AA) {
      st = launch_kind(h, (h->t % 2 == 0) ? K_AA_EVEN : K_AA_ODD, h->f[0], h->f[0]);
    } else {
      st = launch_kind(h, h->pull ? K_AB_PULL : K_AB_PUSH, h->f[h->cur], h->f[1 - h->cur]);
      h->cur = 1 - h->cur;
    }
    if (st != LBM_OK) return st;
    h->t++;
  }
  return LBM_OK;
}

void free_all(lbm_dense* h) {
  if (!h) return;
  for (int i = 0; i < 2; i++)
    if (h->f[i]) (void)hipFree(h->f[i]);
  if (h->links) (void)hipFree(h->links);
  if (h->omega_field) (void)hipFree(h->omega_field);
  if (h->gamma_field) (void)hipFree(h->gamma_field);
  if (h->partial) (void)hipFree(h->partial);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
}

lbm_status dmalloc(lbm_dense* h, void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("lbm_dense_create: hipMalloc of %zu bytes for %s failed: %s", bytes, what,
              hipGetErrorString(e));
    return LBM_ERR_OOM;
  }
  h->device_bytes += bytes;
  return LBM_OK;
}

lbm_status init_from(lbm_dense* h, int src_kind, const double* rho_d, const double* u_d,
                     double u0, double noise, uint64_t seed) {
  h->t = 0;
  h->cur = 0;
  const int mode = canon_mode(h);
  launch_util<k_init_u>(h->q, h->storage, h->layout, grid_for(h), block_for(h->nx), h->stream,
                        h->f[0], h->args, mode, src_kind, rho_d, u_d, u0, noise, seed);
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

lbm_status dense_create(const lbm_dense_config* cfg, lbm_dense** out, bool push_by_gather,
                        bool irows_ok) {
  LBM_REQUIRE(cfg && out, LBM_ERR_ARG, "lbm_dense_create: null pointer");
  *out = nullptr;
  LBM_REQUIRE(cfg->nx > 0 && cfg->ny > 0 && cfg->nz > 0, LBM_ERR_ARG,
              "lbm_dense_create: dimensions must be positive (got %d x %d x %d)", cfg->nx,
              cfg->ny, cfg->nz);
  LBM_REQUIRE(cfg->ny <= 65535 && cfg->nz <= 65535, LBM_ERR_ARG,
              "lbm_dense_create: ny and nz must be <= 65535");
  LBM_REQUIRE(cfg->q == 19 || cfg->q == 27, LBM_ERR_ARG, "lbm_dense_create: q must be 19 or 27");
  LBM_REQUIRE(cfg->pattern == LBM_PATTERN_AA || cfg->pattern == LBM_PATTERN_AB, LBM_ERR_ARG,
              "lbm_dense_create: bad pattern %d", cfg->pattern);
  LBM_REQUIRE(cfg->layout == LBM_LAYOUT_SOA || cfg->layout == LBM_LAYOUT_AOS, LBM_ERR_ARG,
              "lbm_dense_create: bad layout %d", cfg->layout);
  LBM_REQUIRE(cfg->storage == LBM_STORE_F64 || cfg->storage == LBM_STORE_F32S ||
                  cfg->storage == LBM_STORE_F16S,
              LBM_ERR_ARG, "lbm_dense_create: storage must be F64, F32S or F16S");
  LBM_REQUIRE(std::isfinite(cfg->omega) && cfg->omega > 0.0 && cfg->omega < 2.0, LBM_ERR_ARG,
              "lbm_dense_create: omega must be in (0, 2)");
  LBM_REQUIRE(std::isfinite(cfg->density_jump_x) && (cfg->density_jump_x == 0.0 || cfg->periodic[0]),
              LBM_ERR_ARG, "lbm_dense_create: density_jump_x needs a finite value and periodic x");
  lbm_dense* h = new lbm_dense;
  h->nx = cfg->nx;
  h->ny = cfg->ny;
  h->nz = cfg->nz;
  h->q = cfg->q;
  h->pattern = cfg->pattern;
  h->pull = cfg->pattern == LBM_PATTERN_AB ? (cfg->pull ? 1 : 0) : 0;
  h->layout = cfg->layout;
  h->storage = cfg->storage;
  for (int k = 0; k < 3; k++) h->periodic[k] = cfg->periodic[k] ? 1 : 0;
  if (const char* e = std::getenv("LBM_DENSE_STAGED")) h->staged_mode = std::atoi(e);
  if (const char* e = std::getenv("LBM_DENSE_SLAB_PLANES")) h->slab_planes = std::max(0, std::atoi(e));
  h->ncell = (int64_t)cfg->nx * cfg->ny * cfg->nz;
  h->esize = cfg->storage == LBM_STORE_F64 ? 8 : (cfg->storage == LBM_STORE_F32S ? 4 : 2);
  set_launch_shapes(h);  // after ncell and esize: the cache policy depends on the footprint
  // fluid count (host mask)
  h->nfluid = h->ncell;
  if (cfg->solid) {
    int64_t ns = 0;
    for (int64_t c = 0; c < h->ncell; c++) ns += cfg->solid[c] != 0;
    h->nfluid = h->ncell - ns;
  }
  const bool need_links =
      cfg->solid != nullptr || !h->periodic[0] || !h->periodic[1] || !h->periodic[2];
  if (need_links) h->odd_seg_w = 0;  // the segmented odd step covers unmasked periodic boxes
  // Interleaved rows (lbm_dense_kernels.hpp) for AA on unmasked periodic SOA
  // boxes whose rows are whole 64-element blocks: the odd step then makes only
  // whole aligned block accesses, like the even step.  Per-cell fields are
  // stored in the same order.  Default for grids that stream from HBM (h->nt),
  // except D3Q27 FP32 deviations, where plain rows measured faster
  // (set_launch_shapes).  Grids that fit the memory-side cache gained nothing
  // (D3Q19 100-400-step runs, MLUPS interleaved / plain rows: 64^3 FP64 18.7k
  // uncapped, 16.6k with the streaming shape / 19.1-19.5k; 128^3 FP16 65.8-67.3k
  // / 66.4-67.8k), while streaming ones did (256^3 FP16 73.0k / 68.0-68.6k;
  // 256^3 FP64 18.7k / 18.5k).  LBM_DENSE_IROWS=0|1 overrides (A/B timing).
  {
    const char* e = std::getenv("LBM_DENSE_IROWS");
    const bool dflt = h->pattern == LBM_PATTERN_AA && !need_links && h->nt &&
                      !(h->q == 27 && h->storage == LBM_STORE_F32S);
    const bool on = e ? e[0] != '0' : dflt;
    if (on && irows_ok && h->layout == LBM_LAYOUT_SOA && h->nx % 64 == 0) {
      h->xk = h->nx / 64;
      h->odd_seg_w = 0;
    }
  }
  // a per-cell field in the storage order of the populations
  auto store_order = [h](std::vector<double>& v) {
    if (!h->xk) return;
    std::vector<double> t(v.size());
    for (int64_t row = 0; row < h->ncell; row += h->nx)
      for (int x = 0; x < h->nx; x++) t[row + ir_pos(x, h->xk)] = v[row + x];
    v.swap(t);
  };
  // AB push on a masked SOA grid runs the pull (gather) kernel one streaming
  // step out of phase: with P_0 = S^-1 f_0, (C S)^T P_0 = S^-1 (S C)^T f_0, so
  // the push state f_T is the gather of the pull storage, which is how every
  // readout of a pull handle reads it.  Same per-node arithmetic, bit-identical
  // states.  The gather measured faster when the two arrays stream from HBM
  // (384^3 porous, 8.6 GB: 20.4k vs 14.7k MLUPS; reference cylinder at
  // simulationSize 8, 3.5 GB: 18.1k vs 14.9k) and slower when they nearly fit
  // the 256 MiB memory-side cache (simulationSize 4, 460 MB: 14.6k vs 17.9k),
  // so it applies above 1 GiB of populations.  LBM_DENSE_PUSH_GATHER=0|1
  // overrides (A/B timing).
  if (push_by_gather && h->pattern == LBM_PATTERN_AB && !h->pull && need_links &&
      h->layout == LBM_LAYOUT_SOA) {
    const size_t pop_bytes = 2 * (size_t)h->ncell * h->q * h->esize;
    const char* e = std::getenv("LBM_DENSE_PUSH_GATHER");
    const bool gather = e ? e[0] != '0' : pop_bytes > ((size_t)1 << 30);
    if (gather) h->pull = 1;
  }
  lbm_status st = LBM_OK;
  auto fail = [&](lbm_status s) {
    free_all(h);
    return s;
  };
  if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess) {
    set_error("lbm_dense_create: stream/event creation failed");
    return fail(LBM_ERR_HIP);
  }
  h->stream = h->own_stream;
  const size_t fbytes = (size_t)h->ncell * h->q * h->esize;
  const int narrays = h->pattern == LBM_PATTERN_AA ? 1 : 2;
  for (int i = 0; i < narrays; i++)
    if ((st = dmalloc(h, &h->f[i], fbytes, "populations"))) return fail(st);
  if ((st = dmalloc(h, (void**)&h->partial, sizeof(double) * (h->npartial + 1), "reduction")))
    return fail(st);
  if (need_links) {
    if ((st = dmalloc(h, (void**)&h->links, sizeof(uint32_t) * h->ncell, "link mask")))
      return fail(st);
    uint8_t* solid_d = nullptr;
    if (hipMalloc(&solid_d, h->ncell) != hipSuccess) {
      set_error("lbm_dense_create: hipMalloc of the solid mask failed");
      return fail(LBM_ERR_OOM);
    }
    hipError_t e = cfg->solid
                       ? hipMemcpyAsync(solid_d, cfg->solid, h->ncell, hipMemcpyHostToDevice,
                                        h->stream)
                       : hipMemsetAsync(solid_d, 0, h->ncell, h->stream);
    if (e == hipSuccess) {
      const dim3 g = grid_for(h), b = block_for(h->nx);
      if (h->q == 19)
        hipLaunchKernelGGL(k_build_links<19>, g, b, 0, h->stream, solid_d, h->links, h->nx,
                           h->ny, h->nz, h->periodic[0], h->periodic[1], h->periodic[2], h->xk);
      else
        hipLaunchKernelGGL(k_build_links<27>, g, b, 0, h->stream, solid_d, h->links, h->nx,
                           h->ny, h->nz, h->periodic[0], h->periodic[1], h->periodic[2], h->xk);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    (void)hipFree(solid_d);
    if (e != hipSuccess) {
      set_error("lbm_dense_create: link mask build failed: %s", hipGetErrorString(e));
      return fail(LBM_ERR_HIP);
    }
  }
  if (cfg->tau) {
    std::vector<double> om(h->ncell);
    for (int64_t c = 0; c < h->ncell; c++) {
      const double tau = cfg->tau[c];
      const bool is_solid = cfg->solid && cfg->solid[c];
      if (!is_solid && !(std::isfinite(tau) && tau > 0.5)) {
        set_error("lbm_dense_create: tau[%lld] = %g must be > 0.5", (long long)c, tau);
        return fail(LBM_ERR_ARG);
      }
      om[c] = is_solid ? 1.0 : 1.0 / tau;
    }
    store_order(om);
    if ((st = dmalloc(h, (void**)&h->omega_field, sizeof(double) * h->ncell, "omega field")))
      return fail(st);
    if (hipMemcpy(h->omega_field, om.data(), sizeof(double) * h->ncell, hipMemcpyHostToDevice) !=
        hipSuccess) {
      set_error("lbm_dense_create: omega field copy failed");
      return fail(LBM_ERR_HIP);
    }
  }
  if (cfg->gamma) {
    if ((st = dmalloc(h, (void**)&h->gamma_field, sizeof(double) * h->ncell, "gamma field")))
      return fail(st);
    std::vector<double> gm(cfg->gamma, cfg->gamma + h->ncell);
    store_order(gm);
    if (hipMemcpy(h->gamma_field, gm.data(), sizeof(double) * h->ncell, hipMemcpyHostToDevice) !=
        hipSuccess) {
      set_error("lbm_dense_create: gamma field copy failed");
      return fail(LBM_ERR_HIP);
    }
  }
  const double scale = cfg->f16_scale > 0.0 ? cfg->f16_scale : 4096.0;
  DenseArgs& a = h->args;
  a.nx = h->nx;
  a.ny = h->ny;
  a.nz = h->nz;
  a.ncell = h->ncell;
  a.periodic_bits = (h->periodic[0] ? 1u : 0u) | (h->periodic[1] ? 2u : 0u) |
                    (h->periodic[2] ? 4u : 0u);
  // Zeros into solid cells' slots complete the SOA row segments (+12% on a 35%
  // solid D3Q19 push step); in AOS the extra scattered stores cost more than
  // they save (-33%).  LBM_DENSE_FILL_SOLID=0 turns it off (A/B timing).
  {
    const char* e = std::getenv("LBM_DENSE_FILL_SOLID");
    a.fill_solid = h->layout == LBM_LAYOUT_SOA && !(e && e[0] == '0');
  }
  a.links = h->links;
  a.omega_field = h->omega_field;
  a.gamma_field = h->gamma_field;
  a.omega = cfg->omega;
  a.F0 = cfg->force[0];
  a.F1 = cfg->force[1];
  a.F2 = cfg->force[2];
  a.scale = (float)scale;
  a.inv_scale = (float)(1.0 / scale);
  a.djump = cfg->density_jump_x;
  a.xk = h->xk;
  // With F = 0 the Guo terms are exact zeros, so the kernels that evaluate
  // them give the same bits as the specialised ones.  For D3Q19 FP64 AA on
  // SOA boxes without solids and with plain rows of more than 256 cells
  // (whole-row staged workgroups of more than 4 wavefronts) they also
  // measured faster: two handles alternating in one process, MLUPS without /
  // with the terms, 512^3: 17,436 / 18,304; 256^3: 18,437 / 18,414; 128^3:
  // 19,083 / 18,965; 64^3: 18,373 / 16,990.  In every other case measured
  // (FP32/FP16 deviations, D3Q27, AB, the interleaved-row odd step: 7.33 /
  // 7.37 ms at 512^3) they were equal or slower.  LBM_DENSE_FORCE_TERMS=0|1
  // overrides.
  {
    const bool dflt = h->q == 19 && h->storage == LBM_STORE_F64 && h->pattern == LBM_PATTERN_AA &&
                      h->layout == LBM_LAYOUT_SOA && h->links == nullptr && h->nx > 256 && !h->xk;
    const char* e = std::getenv("LBM_DENSE_FORCE_TERMS");
    h->force_terms = e ? e[0] == '1' : dflt;
  }
  // rest state, as PropagationAA/AB::initialize (src/propagation/PropagationAA.cpp:60-123)
  if ((st = init_from(h, 0, nullptr, nullptr, 0.0, 0.0, 0))) return fail(st);
  if (h->pattern == LBM_PATTERN_AB &&
      hipMemcpyAsync(h->f[1], h->f[0], fbytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
    set_error("lbm_dense_create: initial copy failed");
    return fail(LBM_ERR_HIP);
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_error("lbm_dense_create: initialisation failed");
    return fail(LBM_ERR_HIP);
  }
  *out = h;
  return LBM_OK;
}

}  // namespace lbm

extern "C" {

lbm_status lbm_dense_create(const lbm_dense_config* cfg, lbm_dense** out) {
  return dense_create(cfg, out, true, true);
}

lbm_status lbm_dense_destroy(lbm_dense* h) {
  if (!h) return LBM_OK;
  LBM_REQUIRE(h->live_graphs == 0, LBM_ERR_STATE,
              "lbm_dense_destroy: %d graph(s) of this handle are still alive; destroy them first",
              h->live_graphs);
  (void)hipStreamSynchronize(h->stream);
  free_all(h);
  return LBM_OK;
}

lbm_status lbm_dense_get_info(lbm_dense* h, lbm_dense_info* info) {
  LBM_REQUIRE(h && info, LBM_ERR_ARG, "lbm_dense_get_info: null pointer");
  info->ncell = h->ncell;
  info->n_fluid = h->nfluid;
  info->time_step = h->t;
  info->bytes_per_lup = 2.0 * h->q * (double)h->esize;
  info->device_bytes = h->device_bytes;
  return LBM_OK;
}

lbm_status lbm_dense_set_state(lbm_dense* h, const double* f) {
  LBM_REQUIRE(h && f, LBM_ERR_ARG, "lbm_dense_set_state: null pointer");
  const size_t bytes = sizeof(double) * h->q * h->ncell;
  double* stage = nullptr;
  LBM_HIP_CHECK(hipMalloc(&stage, bytes));
  hipError_t e = hipMemcpyAsync(stage, f, bytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    h->t = 0;
    h->cur = 0;
    launch_util<k_set_canonical_u>(h->q, h->storage, h->layout, grid_for(h), block_for(h->nx),
                                   h->stream, (const double*)stage, h->f[0], h->args,
                                   canon_mode(h));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  }
  (void)hipFree(stage);
  if (e != hipSuccess) {
    set_error("lbm_dense_set_state: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  return LBM_OK;
}

lbm_status lbm_dense_get_state(lbm_dense* h, double* f) {
  LBM_REQUIRE(h && f, LBM_ERR_ARG, "lbm_dense_get_state: null pointer");
  const size_t bytes = sizeof(double) * h->q * h->ncell;
  double* stage = nullptr;
  LBM_HIP_CHECK(hipMalloc(&stage, bytes));
  launch_util<k_get_canonical_u>(h->q, h->storage, h->layout, grid_for(h), block_for(h->nx),
                                 h->stream, (const void*)h->f[h->cur], stage, h->args,
                                 canon_mode(h));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(f, stage, bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(stage);
  if (e != hipSuccess) {
    set_error("lbm_dense_get_state: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  return LBM_OK;
}

lbm_status lbm_dense_init_equilibrium(lbm_dense* h, const double* rho, const double* u) {
  LBM_REQUIRE(h && u, LBM_ERR_ARG, "lbm_dense_init_equilibrium: null pointer");
  double *rho_d = nullptr, *u_d = nullptr;
  LBM_HIP_CHECK(hipMalloc(&u_d, sizeof(double) * 3 * h->ncell));
  hipError_t e = hipMemcpyAsync(u_d, u, sizeof(double) * 3 * h->ncell, hipMemcpyHostToDevice,
                                h->stream);
  if (e == hipSuccess && rho) {
    e = hipMalloc(&rho_d, sizeof(double) * h->ncell);
    if (e == hipSuccess)
      e = hipMemcpyAsync(rho_d, rho, sizeof(double) * h->ncell, hipMemcpyHostToDevice,
                         h->stream);
  }
  lbm_status st = LBM_OK;
  if (e == hipSuccess) st = init_from(h, 1, rho_d, u_d, 0.0, 0.0, 0);
  if (e == hipSuccess && st == LBM_OK) e = hipStreamSynchronize(h->stream);
  (void)hipFree(u_d);
  if (rho_d) (void)hipFree(rho_d);
  if (st != LBM_OK) return st;
  if (e != hipSuccess) {
    set_error("lbm_dense_init_equilibrium: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  return LBM_OK;
}

lbm_status lbm_dense_init_tgv(lbm_dense* h, double u0, double noise, uint64_t seed) {
  LBM_REQUIRE(h, LBM_ERR_ARG, "lbm_dense_init_tgv: null handle");
  if (lbm_status st = init_from(h, 2, nullptr, nullptr, u0, noise, seed)) return st;
  LBM_HIP_CHECK(hipStreamSynchronize(h->stream));
  return LBM_OK;
}

lbm_status lbm_dense_step(lbm_dense* h, int32_t nsteps) {
  LBM_REQUIRE(h && nsteps >= 0, LBM_ERR_ARG, "lbm_dense_step: bad arguments");
  return enqueue_steps(h, nsteps);
}

lbm_status lbm_dense_run_timed(lbm_dense* h, int32_t nsteps, double* elapsed_ms) {
  LBM_REQUIRE(h && elapsed_ms && nsteps >= 0, LBM_ERR_ARG, "lbm_dense_run_timed: bad arguments");
  LBM_HIP_CHECK(hipEventRecord(h->e0, h->stream));
  if (lbm_status st = enqueue_steps(h, nsteps)) return st;
  LBM_HIP_CHECK(hipEventRecord(h->e1, h->stream));
  LBM_HIP_CHECK(hipEventSynchronize(h->e1));
  float ms = 0.f;
  LBM_HIP_CHECK(hipEventElapsedTime(&ms, h->e0, h->e1));
  *elapsed_ms = ms;
  return LBM_OK;
}

// Every call on the handle enqueues on h->stream.  Switching waits for the
// work already on the old stream, so the order of calls is kept.
lbm_status lbm_dense_set_stream(lbm_dense* h, void* stream) {
  LBM_REQUIRE(h, LBM_ERR_ARG, "lbm_dense_set_stream: null handle");
  LBM_HIP_CHECK(hipStreamSynchronize(h->stream));
  h->stream = stream ? (hipStream_t)stream : h->own_stream;
  return LBM_OK;
}

// A captured run of an even number of steps.  After an even number of steps
// the AA array is back in the convention it started in and the AB pair has
// swapped twice, so one replay continues the run from any state whose phase
// (t mod 2 for AA, the current array for AB) equals the phase at capture.
// The kernels take the handle's device pointers, which stay fixed for the
// handle's lifetime; capture records kernel launches only.
struct lbm_dense_graph {
  lbm_dense* h = nullptr;
  int32_t nsteps = 0;
  int phase = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool counted = false;  // included in h->live_graphs
};

static int dense_phase(const lbm_dense* h) {
  return h->pattern == LBM_PATTERN_AA ? (int)(h->t % 2) : h->cur;
}

lbm_status lbm_dense_graph_destroy(lbm_dense_graph* g) {
  if (!g) return LBM_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->counted) g->h->live_graphs--;
  delete g;
  return LBM_OK;
}

lbm_status lbm_dense_graph_create(lbm_dense* h, int32_t nsteps, lbm_dense_graph** out) {
  LBM_REQUIRE(h && out, LBM_ERR_ARG, "lbm_dense_graph_create: null pointer");
  LBM_REQUIRE(nsteps > 0 && nsteps % 2 == 0, LBM_ERR_ARG,
              "lbm_dense_graph_create: nsteps must be even and positive (got %d)", nsteps);
  *out = nullptr;
  auto* g = new lbm_dense_graph;
  g->h = h;
  g->nsteps = nsteps;
  g->phase = dense_phase(h);
  // capture on a stream of its own: the caller's stream may be the null
  // stream, which cannot be captured, and nothing captured runs now
  hipStream_t cap = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete g;
    set_error("lbm_dense_graph_create: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  const hipStream_t user = h->stream;
  const int64_t t0 = h->t;
  const int cur0 = h->cur;
  h->stream = cap;
  lbm_status st = LBM_OK;
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    st = enqueue_steps(h, nsteps);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cap, &graph);
    g->graph = graph;
    if (st == LBM_OK && e2 != hipSuccess) e = e2;
  }
  h->stream = user;
  h->t = t0;
  h->cur = cur0;
  (void)hipStreamDestroy(cap);
  if (st == LBM_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (st != LBM_OK || e != hipSuccess) {
    if (st == LBM_OK) set_error("lbm_dense_graph_create: %s", hipGetErrorString(e));
    lbm_dense_graph_destroy(g);
    return st != LBM_OK ? st : LBM_ERR_HIP;
  }
  g->counted = true;
  h->live_graphs++;
  *out = g;
  return LBM_OK;
}

lbm_status lbm_dense_graph_launch(lbm_dense_graph* g, int32_t nlaunch) {
  LBM_REQUIRE(g && nlaunch >= 0, LBM_ERR_ARG, "lbm_dense_graph_launch: bad arguments");
  lbm_dense* h = g->h;
  LBM_REQUIRE(dense_phase(h) == g->phase, LBM_ERR_ARG,
              "lbm_dense_graph_launch: the handle is an odd number of steps away from the "
              "state the graph was captured in");
  for (int32_t i = 0; i < nlaunch; i++) {
    LBM_HIP_CHECK(hipGraphLaunch(g->exec, h->stream));
    h->t += g->nsteps;
  }
  return LBM_OK;
}

lbm_status lbm_dense_moments(lbm_dense* h, double* rho, double* u) {
  LBM_REQUIRE(h && rho && u, LBM_ERR_ARG, "lbm_dense_moments: null pointer");
  double* buf = nullptr;
  LBM_HIP_CHECK(hipMalloc(&buf, sizeof(double) * 4 * h->ncell));
  launch_util<k_moments_u>(h->q, h->storage, h->layout, grid_for(h), block_for(h->nx), h->stream,
                           (const void*)h->f[h->cur], buf, buf + h->ncell, h->args,
                           canon_mode(h));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(rho, buf, sizeof(double) * h->ncell, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(u, buf + h->ncell, sizeof(double) * 3 * h->ncell, hipMemcpyDeviceToHost,
                       h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) {
    set_error("lbm_dense_moments: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  return LBM_OK;
}

lbm_status lbm_dense_mass(lbm_dense* h, double* mass) {
  LBM_REQUIRE(h && mass, LBM_ERR_ARG, "lbm_dense_mass: null pointer");
  launch_util<k_mass_partial_u>(h->q, h->storage, h->layout, dim3(h->npartial), dim3(256),
                                h->stream, (const void*)h->f[h->cur], h->args, h->partial);
  LBM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, h->stream, (const double*)h->partial,
                     h->npartial, h->partial + h->npartial);
  LBM_LAUNCH_CHECK();
  LBM_HIP_CHECK(hipMemcpyAsync(mass, h->partial + h->npartial, sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
  LBM_HIP_CHECK(hipStreamSynchronize(h->stream));
  LBM_REQUIRE(std::isfinite(*mass), LBM_ERR_NUMERIC, "lbm_dense_mass: mass is not finite (%g)",
              *mass);
  return LBM_OK;
}

lbm_status lbm_dense_synchronize(lbm_dense* h) {
  LBM_REQUIRE(h, LBM_ERR_ARG, "lbm_dense_synchronize: null handle");
  LBM_HIP_CHECK(hipStreamSynchronize(h->stream));
  return LBM_OK;
}

}  // extern "C"
