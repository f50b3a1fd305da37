// lbm_sparse.hip -- drop-in GPU replacement of the reference's kernel classes on
// its fluid-only, indirectly addressed node list.
//
// One work-item per fluid node, exactly the reference's parallel_for shape
// (src/kernels/KernelAA_SOA_Base.cpp:180, :263; KernelAB_SOA_Base.cpp:116;
// KernelAB_SOA_Pull.cpp:118), with the node's Q populations held in registers
// and the collision fully unrolled at compile time.  The arithmetic is
// collide_ref (the Base kernels' association and order, no FMA contraction),
// so FP64 results and -D_PRECISION_SINGLE_ FP32 results equal the reference's
// bit for bit.
//
// Launches go to the null stream by default, so that they are ordered with the
// caller's blocking hipMemcpy traffic, as the reference's kernel calls are
// ordered with its host code; lbm_sparse_set_stream moves them to a caller's
// stream (which can then be captured into a HIP graph: the step calls only
// launch kernels).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "lbm_collide.hpp"
#include "lbm_internal.hpp"

namespace lbm {

template <int Q, int LAYOUT>
__device__ __forceinline__ int64_t pidx(int64_t n, int s, int64_t i) {
  return LAYOUT == LBM_LAYOUT_SOA ? n * s + i : (int64_t)Q * i + s;
}

// Node-aligned accesses (a work-item's own slots and adjacency entries) are
// touched once per step.  With NT they stream past the caches, which pays when
// the arrays exceed the memory-side cache; arrays that fit in it are faster
// cached (the reference's simulationSize 4 deck, 175 MB: 15.2k vs 13.6k MLUPS
// AA; simulationSize 8, 1.35 GB: 13.7k vs 14.8k).  Accesses through the
// adjacency stay cacheable, since neighbouring work-items share their lines.
template <bool NT, typename T>
__device__ __forceinline__ T ldn(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void stn(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Node block of this workgroup.  Workgroups are dispatched to the 8 XCDs in
// turn (block b to XCD b mod 8), each with its own L2.  group > 0 (the launch
// then has a multiple of 8 * group blocks) gives each XCD runs of `group`
// consecutive node blocks, so that cache lines shared by neighbouring blocks
// (accesses through the adjacency) mostly meet in one L2: physical block p,
// round r = p / 8, XCD k = p mod 8, takes node block
// (r / group) 8 group + k group + r mod group, a bijection; blocks past the
// range exit through the count check.
__device__ __forceinline__ int64_t node_block(uint32_t group) {
  if (!group) return blockIdx.x;
  const uint32_t r = blockIdx.x / 8u, k = blockIdx.x % 8u;
  return (int64_t)(r / group) * 8 * group + (int64_t)k * group + r % group;
}

// AA even: src/kernels/KernelAA_SOA_Base.cpp:173-254 (AOS: KernelAA_AOS_Base.cpp:174-254)
template <int Q, typename T, int LAYOUT, bool NT>
__global__ void __launch_bounds__(256)
    k_sparse_aa_even(T* __restrict__ f, int64_t n, int64_t start, int64_t count, double omega,
                     T F0, T F1, T F2, uint32_t group) {
  const int64_t i = start + node_block(group) * blockDim.x + threadIdx.x;
  if (i >= start + count) return;
  T d[Q];
#pragma unroll
  for (int s = 0; s < Q; s++) d[s] = ldn<NT>(f + pidx<Q, LAYOUT>(n, Lattice<Q>::opp(s), i));
  const T F[3] = {F0, F1, F2};
  collide_ref<Q, T, false>(d, d, omega, F, 0.0);
#pragma unroll
  for (int s = 0; s < Q; s++) stn<NT>(f + pidx<Q, LAYOUT>(n, s, i), d[s]);
}

// AA odd: src/kernels/KernelAA_SOA_Base.cpp:257-342 (AOS: KernelAA_AOS_Base.cpp:258-342)
template <int Q, typename T, int LAYOUT, bool NT>
__global__ void __launch_bounds__(256)
    k_sparse_aa_odd(T* __restrict__ f, const int32_t* __restrict__ adj, int64_t n, int64_t start,
                    int64_t count, double omega, T F0, T F1, T F2, uint32_t group) {
  const int64_t i = start + node_block(group) * blockDim.x + threadIdx.x;
  if (i >= start + count) return;
  T d[Q];
  int32_t loc[Q];
#pragma unroll
  for (int s = 0; s < Q; s++) {
    loc[s] = ldn<NT>(adj + pidx<Q, LAYOUT>(n, s, i));
    d[s] = f[loc[s]];
  }
  const T F[3] = {F0, F1, F2};
  collide_ref<Q, T, false>(d, d, omega, F, 0.0);
#pragma unroll
  for (int s = 0; s < Q; s++) f[loc[Lattice<Q>::opp(s)]] = d[s];
}

// AB push: src/kernels/KernelAB_SOA_Base.cpp:111-186 (AOS: KernelAB_AOS_Base.cpp:108-182)
template <int Q, typename T, int LAYOUT, bool NT>
__global__ void __launch_bounds__(256)
    k_sparse_ab_push(const T* __restrict__ src, T* __restrict__ tgt,
                     const int32_t* __restrict__ adj, int64_t n, int64_t start, int64_t count,
                     double omega, T F0, T F1, T F2, uint32_t group) {
  const int64_t i = start + node_block(group) * blockDim.x + threadIdx.x;
  if (i >= start + count) return;
  T d[Q];
#pragma unroll
  for (int s = 0; s < Q; s++) d[s] = ldn<NT>(src + pidx<Q, LAYOUT>(n, s, i));
  const T F[3] = {F0, F1, F2};
  collide_ref<Q, T, false>(d, d, omega, F, 0.0);
#pragma unroll
  for (int s = 0; s < Q; s++) tgt[ldn<NT>(adj + pidx<Q, LAYOUT>(n, s, i))] = d[s];
}

// AB pull: src/kernels/KernelAB_SOA_Pull.cpp:113-194
template <int Q, typename T, int LAYOUT, bool NT>
__global__ void __launch_bounds__(256)
    k_sparse_ab_pull(const T* __restrict__ src, T* __restrict__ tgt,
                     const int32_t* __restrict__ adj, int64_t n, int64_t start, int64_t count,
                     double omega, T F0, T F1, T F2, uint32_t group) {
  const int64_t i = start + node_block(group) * blockDim.x + threadIdx.x;
  if (i >= start + count) return;
  T d[Q];
#pragma unroll
  for (int s = 0; s < Q; s++) d[s] = src[ldn<NT>(adj + pidx<Q, LAYOUT>(n, s, i))];
  const T F[3] = {F0, F1, F2};
  collide_ref<Q, T, false>(d, d, omega, F, 0.0);
#pragma unroll
  for (int s = 0; s < Q; s++) stn<NT>(tgt + pidx<Q, LAYOUT>(n, s, i), d[s]);
}

// PropagationAA::initialize (slot opp(s) := w_s) and PropagationAB::initialize (slot s := w_s).
template <int Q, typename T, int LAYOUT, bool NT>
__global__ void __launch_bounds__(256) k_sparse_init_rest(T* __restrict__ f, int64_t n, int aa) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int slot = aa ? Lattice<Q>::opp(s) : s;
    f[pidx<Q, LAYOUT>(n, slot, i)] = T(Lattice<Q>::w(s));
  }
}

// FP64 mass of a fluid-index range: the sum of the Q stored values of every
// node in [start, start + count) (over the whole range this is the total mass
// in any slot convention, which only permutes the populations over the
// slots).  Deterministic two-pass reduction with a fixed shape: workgroup b of
// 256 lanes sums the kMassChunk nodes start + b kMassChunk + ... (lane t takes
// nodes t, t + 256, ... of the chunk, each node's values in s order, in FP64)
// into partial[b]; one workgroup then sums the partials in a fixed order.
constexpr int kMassBlock = 256;
constexpr int64_t kMassChunk = 16 * kMassBlock;

__device__ __forceinline__ double mass_block_sum(double v) {
  __shared__ double red[kMassBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kMassBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

// (NT is unused: the launcher instantiates every kernel for both cache policies.)
template <int Q, typename T, int LAYOUT, bool NT>
__global__ void __launch_bounds__(kMassBlock)
    k_sparse_mass_partial(const T* __restrict__ f, int64_t n, int64_t start, int64_t count,
                          double* __restrict__ partial) {
  const int64_t c0 = (int64_t)blockIdx.x * kMassChunk;
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < kMassChunk; k += kMassBlock) {
    if (c0 + k >= count) break;
    const int64_t i = start + c0 + k;
#pragma unroll
    for (int s = 0; s < Q; s++) acc += (double)f[pidx<Q, LAYOUT>(n, s, i)];
  }
  const double tot = mass_block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kMassBlock)
    k_sparse_mass_final(const double* __restrict__ partial, int64_t nb, double* __restrict__ out) {
  double acc = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += kMassBlock) acc += partial[b];
  const double tot = mass_block_sum(acc);
  if (threadIdx.x == 0) *out = tot;
}

// Partial sums of lbm_sparse_mass_noalloc: a fixed device array instead of an
// allocation, kMassScratchPartials doubles (32 KiB) on every GPU the library
// runs on, held until the process ends.  It bounds that entry point's range to
// kMassScratchPartials * kMassChunk = 16,777,216 nodes.  In host code `scratch`
// only names the array; its device address comes from hipGetSymbolAddress
// (scratch_on_current_device), once per device.  The two-pass reduction
// writes every partial it reads, so nothing needs clearing between calls;
// calls that overlap on one GPU share the array and race.
constexpr int64_t kMassScratchPartials = 4096;
static __device__ __align__(256) double scratch[kMassScratchPartials];

}  // namespace lbm

struct lbm_sparse {
  int64_t n = 0;
  int q = 19;
  int layout = LBM_LAYOUT_SOA;
  int storage = LBM_STORE_F64;
  int adjacency_mode = 1;
  double omega = 1.0;
  double force[3] = {1e-6, 0.0, 0.0};
  int32_t* adj = nullptr;  // device copy of the adjacency
  // launch shapes: the AA even step (all node-aligned) and the other steps
  unsigned even_bx = 256, step_bx = 256, push_bx = 256;
  size_t even_lds = 0, step_lds = 0, push_lds = 0;
  bool nt = true;  // non-temporal node-aligned accesses (arrays beyond the memory-side cache)
  hipStream_t stream = nullptr;  // null stream unless the caller set one
  int xcd_env = -1;              // LBM_SPARSE_XCD: node blocks per XCD run (-1: by footprint)
};

namespace {

using namespace lbm;

inline unsigned grid_for(int64_t count, unsigned bx = 256) {
  return (unsigned)((count + bx - 1) / bx);
}

struct Shape {
  unsigned grid, block;
  size_t lds;
  uint32_t group;  // node_block: 0, or node blocks per XCD run (grid a multiple of 8 group)
};

#define LBM_SPARSE_KERNEL_FN(name)                                                         \
  template <int Q, typename T, int L>                                                      \
  struct name##_fn {                                                                       \
    template <typename... A>                                                               \
    static void launch(Shape sh, bool nt, hipStream_t st, A... a) {                        \
      if (nt)                                                                              \
        hipLaunchKernelGGL((lbm::name<Q, T, L, true>), dim3(sh.grid), dim3(sh.block),      \
                           sh.lds, st, a...);                                              \
      else                                                                                 \
        hipLaunchKernelGGL((lbm::name<Q, T, L, false>), dim3(sh.grid), dim3(sh.block),     \
                           sh.lds, st, a...);                                              \
    }                                                                                      \
  };
LBM_SPARSE_KERNEL_FN(k_sparse_aa_even)
LBM_SPARSE_KERNEL_FN(k_sparse_aa_odd)
LBM_SPARSE_KERNEL_FN(k_sparse_ab_push)
LBM_SPARSE_KERNEL_FN(k_sparse_ab_pull)
LBM_SPARSE_KERNEL_FN(k_sparse_init_rest)
LBM_SPARSE_KERNEL_FN(k_sparse_mass_partial)
#undef LBM_SPARSE_KERNEL_FN

template <template <int, typename, int> class K, typename T, typename... A>
void launch_qt(const lbm_sparse* h, Shape sh, A... a) {
  if (h->q == 19) {
    if (h->layout == LBM_LAYOUT_SOA) K<19, T, LBM_LAYOUT_SOA>::launch(sh, h->nt, h->stream, a...);
    else K<19, T, LBM_LAYOUT_AOS>::launch(sh, h->nt, h->stream, a...);
  } else {
    if (h->layout == LBM_LAYOUT_SOA) K<27, T, LBM_LAYOUT_SOA>::launch(sh, h->nt, h->stream, a...);
    else K<27, T, LBM_LAYOUT_AOS>::launch(sh, h->nt, h->stream, a...);
  }
}

// Node blocks per XCD run (node_block).  Measured on the reference's decks
// (drop-in driver, MLUPS, none -> chosen):
//   arrays in the memory-side cache (simulationSize 4, 175 MB): one run per
//   XCD, AA 15.2k -> 16.2k, AB 12.8k -> 14.0k, pull 11.6k -> 11.9k;
//   streaming (simulationSize 8, 1.35 GB): runs of 8-32 blocks raise AB push
//   10.6k -> 11.2k, while AA and pull lost with long runs (14.7k -> 13.3k,
//   13.4k -> 11.4k with one run per XCD) and gained nothing with short ones.
//   With the push's 128-wide capped blocks (set_shapes), runs of 64-128
//   blocks measured best: 11.46-11.47k (16: 11.16-11.20k; 256: 11.34k).
// LBM_SPARSE_XCD (read at create) overrides: 0 off, g > 0 runs of g blocks.
uint32_t xcd_group(const lbm_sparse* h, bool push, unsigned nb) {
  if (h->xcd_env >= 0) return (uint32_t)h->xcd_env;
  if (!h->nt) return (nb + 7) / 8;
  if (!push) return 0u;
  return h->push_bx == 128 ? 64u : 16u;
}

Shape make_shape(const lbm_sparse* h, int64_t count, unsigned bx, size_t lds, bool push) {
  const unsigned nb = grid_for(count, bx);
  const uint32_t g = xcd_group(h, push, nb);
  if (!g || nb < 16) return {nb, bx, lds, 0};
  const unsigned span = 8 * g;
  return {(nb + span - 1) / span * span, bx, lds, g};
}
Shape even_shape(const lbm_sparse* h, int64_t count) {
  return make_shape(h, count, h->even_bx, h->even_lds, false);
}
Shape step_shape(const lbm_sparse* h, int64_t count, bool push = false) {
  return push ? make_shape(h, count, h->push_bx, h->push_lds, true)
              : make_shape(h, count, h->step_bx, h->step_lds, false);
}

// Launch shapes, fixed at create.  The AA even step is all node-aligned, like
// the dense engine's even step, which measured fastest with small workgroups
// at capped occupancy.  So did the AB push on streaming FP64 arrays
// (simulationSize-8 deck, MLUPS: 128 wide, <= 4 per CU with XCD runs of 64,
// 11.46k; 256 wide, no cap, runs of 16: 11.18-11.21k; 64 wide, <= 8 per CU:
// 10.9-11.3k).  The other steps keep 256-wide workgroups.  Environment
// overrides (tuning): LBM_SPARSE_EVEN_BLOCK_X, LBM_SPARSE_EVEN_BLOCKS_PER_CU,
// LBM_SPARSE_BLOCK_X, LBM_SPARSE_BLOCKS_PER_CU (all other steps, the push
// included).
void set_shapes(lbm_sparse* h) {
  auto env_int = [](const char* name, int dflt) {
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
  };
  auto clamp_bx = [](int bx) { return (unsigned)(bx < 64 ? 64 : (bx > 256 ? 256 : bx / 64 * 64)); };
  auto lds_for = [](int k) { return k > 0 ? (size_t)(160 * 1024 / k) & ~(size_t)1023 : (size_t)0; };
  // footprint of one step: populations (two arrays for AB) and adjacency,
  // against the 256 MiB memory-side cache of the MI355X (LBM_SPARSE_NT=0|1
  // overrides)
  const size_t esize = h->storage == LBM_STORE_F64 ? 8 : 4;
  const size_t nq = (size_t)h->n * h->q;
  const size_t bytes = nq * (esize * (h->adjacency_mode == 1 ? 1 : 2) + 4);
  h->nt = env_int("LBM_SPARSE_NT", bytes > ((size_t)256 << 20) ? 1 : 0) != 0;
  // the small capped even shape is for grids that stream from HBM; cached
  // grids are small enough that the cap leaves the GPU partly idle
  const bool big = h->nt && h->storage == LBM_STORE_F64;
  h->even_bx = clamp_bx(env_int("LBM_SPARSE_EVEN_BLOCK_X", big ? 128 : 256));
  h->even_lds = lds_for(env_int("LBM_SPARSE_EVEN_BLOCKS_PER_CU", big ? 4 : 0));
  h->step_bx = clamp_bx(env_int("LBM_SPARSE_BLOCK_X", 256));
  h->step_lds = lds_for(env_int("LBM_SPARSE_BLOCKS_PER_CU", 0));
  h->push_bx = clamp_bx(env_int("LBM_SPARSE_BLOCK_X", big ? 128 : 256));
  h->push_lds = lds_for(env_int("LBM_SPARSE_BLOCKS_PER_CU", big ? 4 : 0));
}

lbm_status check_range(const lbm_sparse* h, int64_t start, int64_t count) {
  LBM_REQUIRE(start >= 0 && count >= 0 && start + count <= h->n, LBM_ERR_ARG,
              "fluid range [%lld, %lld) outside [0, %lld)", (long long)start,
              (long long)(start + count), (long long)h->n);
  return LBM_OK;
}

// The two mass-reduction launches of a range into `partial` (nb entries) and
// `out`, on the handle's stream.  Shared by the allocating and the
// allocation-free entry points, so both give the same bits.
lbm_status launch_mass(const lbm_sparse* h, const void* dstrb, int64_t start, int64_t count,
                       double* partial, int64_t nb, double* out) {
  if (nb > 0) {
    const Shape g{(unsigned)nb, (unsigned)kMassBlock, 0, 0};
    if (h->storage == LBM_STORE_F64)
      launch_qt<k_sparse_mass_partial_fn, double>(h, g, (const double*)dstrb, h->n, start, count,
                                                  partial);
    else
      launch_qt<k_sparse_mass_partial_fn, float>(h, g, (const float*)dstrb, h->n, start, count,
                                                 partial);
    LBM_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_sparse_mass_final, dim3(1), dim3(kMassBlock), 0, h->stream,
                     (const double*)partial, nb, out);
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

int64_t mass_blocks(int64_t count) { return (count + kMassChunk - 1) / kMassChunk; }

// Device address of `scratch` on the current device, looked up once per device
// (at lbm_sparse_create, or at the first allocation-free call on a device).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

lbm_status scratch_on_current_device(double** out) {
  int dev = 0;
  LBM_HIP_CHECK(hipGetDevice(&dev));
  LBM_REQUIRE(dev >= 0 && dev < kMaxDevices, LBM_ERR_UNSUPPORTED, "device %d beyond %d", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    LBM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    LBM_REQUIRE(p, LBM_ERR_HIP, "no device address for the mass scratch on device %d", dev);
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return LBM_OK;
}

// true when the device can read and write *p: device, managed or pinned host memory
bool device_accessible(const void* p) {
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
         attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeUnified;
}

}  // namespace

extern "C" {

lbm_status lbm_sparse_create(const lbm_sparse_config* cfg, lbm_sparse** out) {
  LBM_REQUIRE(cfg && out, LBM_ERR_ARG, "lbm_sparse_create: null pointer");
  *out = nullptr;
  LBM_REQUIRE(cfg->q == 19 || cfg->q == 27, LBM_ERR_ARG, "lbm_sparse_create: q must be 19 or 27");
  LBM_REQUIRE(cfg->layout == LBM_LAYOUT_SOA || cfg->layout == LBM_LAYOUT_AOS, LBM_ERR_ARG,
              "lbm_sparse_create: bad layout %d", cfg->layout);
  LBM_REQUIRE(cfg->storage == LBM_STORE_F64 || cfg->storage == LBM_STORE_F32, LBM_ERR_ARG,
              "lbm_sparse_create: storage must be LBM_STORE_F64 or LBM_STORE_F32");
  LBM_REQUIRE(cfg->n_fluid > 0 && cfg->adjacency, LBM_ERR_ARG,
              "lbm_sparse_create: need n_fluid > 0 and an adjacency");
  LBM_REQUIRE(cfg->n_fluid * cfg->q < ((int64_t)1 << 31), LBM_ERR_ARG,
              "lbm_sparse_create: n_fluid * q must fit the int32 adjacency");
  const int64_t nq = cfg->n_fluid * cfg->q;
  for (int64_t k = 0; k < nq; k++)
    LBM_REQUIRE(cfg->adjacency[k] >= 0 && cfg->adjacency[k] < nq, LBM_ERR_ARG,
                "lbm_sparse_create: adjacency[%lld] = %d outside [0, %lld)", (long long)k,
                cfg->adjacency[k], (long long)nq);
  lbm_sparse* h = new lbm_sparse;
  h->n = cfg->n_fluid;
  h->q = cfg->q;
  h->layout = cfg->layout;
  h->storage = cfg->storage;
  h->adjacency_mode = cfg->adjacency_mode;
  h->omega = cfg->omega;
  for (int d = 0; d < 3; d++) h->force[d] = cfg->force[d];
  set_shapes(h);
  {
    const char* e = std::getenv("LBM_SPARSE_XCD");
    h->xcd_env = e ? std::max(0, std::min(4096, std::atoi(e))) : -1;
  }
  if (hipMalloc(&h->adj, sizeof(int32_t) * nq) != hipSuccess) {
    delete h;
    set_error("lbm_sparse_create: hipMalloc of the adjacency failed");
    return LBM_ERR_OOM;
  }
  hipError_t e = hipMemcpy(h->adj, cfg->adjacency, sizeof(int32_t) * nq, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(h->adj);
    delete h;
    set_error("lbm_sparse_create: adjacency copy failed: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  double* scr = nullptr;  // set-up of lbm_sparse_mass_noalloc's scratch on this device
  if (lbm_status st = scratch_on_current_device(&scr)) {
    (void)hipFree(h->adj);
    delete h;
    return st;
  }
  *out = h;
  return LBM_OK;
}

lbm_status lbm_sparse_destroy(lbm_sparse* h) {
  if (!h) return LBM_OK;
  // kernels still queued on the handle's stream read the adjacency
  (void)hipStreamSynchronize(h->stream);
  if (h->adj) LBM_HIP_CHECK(hipFree(h->adj));
  delete h;
  return LBM_OK;
}

lbm_status lbm_sparse_aa_even(lbm_sparse* h, void* dstrb, int64_t start, int64_t count) {
  LBM_REQUIRE(h && dstrb, LBM_ERR_ARG, "lbm_sparse_aa_even: null pointer");
  if (lbm_status st = check_range(h, start, count)) return st;
  if (count == 0) return LBM_OK;
  const Shape g = even_shape(h, count);
  if (h->storage == LBM_STORE_F64)
    launch_qt<k_sparse_aa_even_fn, double>(h, g, (double*)dstrb, h->n, start, count, h->omega,
                                          h->force[0], h->force[1], h->force[2], g.group);
  else
    launch_qt<k_sparse_aa_even_fn, float>(h, g, (float*)dstrb, h->n, start, count, h->omega,
                                         (float)h->force[0], (float)h->force[1],
                                         (float)h->force[2], g.group);
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

lbm_status lbm_sparse_aa_odd(lbm_sparse* h, void* dstrb, int64_t start, int64_t count) {
  LBM_REQUIRE(h && dstrb, LBM_ERR_ARG, "lbm_sparse_aa_odd: null pointer");
  LBM_REQUIRE(h->adjacency_mode == 1, LBM_ERR_STATE,
              "lbm_sparse_aa_odd: AA needs the x - c_s adjacency (adjacency_mode 1)");
  if (lbm_status st = check_range(h, start, count)) return st;
  if (count == 0) return LBM_OK;
  const Shape g = step_shape(h, count);
  if (h->storage == LBM_STORE_F64)
    launch_qt<k_sparse_aa_odd_fn, double>(h, g, (double*)dstrb, (const int32_t*)h->adj, h->n,
                                         start, count, h->omega, h->force[0], h->force[1],
                                         h->force[2], g.group);
  else
    launch_qt<k_sparse_aa_odd_fn, float>(h, g, (float*)dstrb, (const int32_t*)h->adj, h->n, start,
                                        count, h->omega, (float)h->force[0], (float)h->force[1],
                                        (float)h->force[2], g.group);
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

lbm_status lbm_sparse_ab(lbm_sparse* h, const void* src, void* tgt, int64_t start, int64_t count,
                         int32_t pull) {
  LBM_REQUIRE(h && src && tgt, LBM_ERR_ARG, "lbm_sparse_ab: null pointer");
  LBM_REQUIRE(src != tgt, LBM_ERR_ARG, "lbm_sparse_ab: source and target must differ");
  LBM_REQUIRE(h->adjacency_mode == (pull ? 1 : 0), LBM_ERR_STATE,
              "lbm_sparse_ab: %s needs adjacency_mode %d", pull ? "pull" : "push", pull ? 1 : 0);
  LBM_REQUIRE(!pull || h->layout == LBM_LAYOUT_SOA, LBM_ERR_UNSUPPORTED,
              "lbm_sparse_ab: the reference has no AOS pull kernel");
  if (lbm_status st = check_range(h, start, count)) return st;
  if (count == 0) return LBM_OK;
  const Shape g = step_shape(h, count, !pull);
  const int32_t* adj = h->adj;
  if (h->storage == LBM_STORE_F64) {
    if (pull)
      launch_qt<k_sparse_ab_pull_fn, double>(h, g, (const double*)src, (double*)tgt, adj, h->n,
                                            start, count, h->omega, h->force[0], h->force[1],
                                            h->force[2], g.group);
    else
      launch_qt<k_sparse_ab_push_fn, double>(h, g, (const double*)src, (double*)tgt, adj, h->n,
                                            start, count, h->omega, h->force[0], h->force[1],
                                            h->force[2], g.group);
  } else {
    const float F0 = (float)h->force[0], F1 = (float)h->force[1], F2 = (float)h->force[2];
    if (pull)
      launch_qt<k_sparse_ab_pull_fn, float>(h, g, (const float*)src, (float*)tgt, adj, h->n,
                                           start, count, h->omega, F0, F1, F2, g.group);
    else
      launch_qt<k_sparse_ab_push_fn, float>(h, g, (const float*)src, (float*)tgt, adj, h->n,
                                           start, count, h->omega, F0, F1, F2, g.group);
  }
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

lbm_status lbm_sparse_init_rest(lbm_sparse* h, void* dstrb, int32_t aa) {
  LBM_REQUIRE(h && dstrb, LBM_ERR_ARG, "lbm_sparse_init_rest: null pointer");
  const Shape g{grid_for(h->n), 256, 0};
  if (h->storage == LBM_STORE_F64)
    launch_qt<k_sparse_init_rest_fn, double>(h, g, (double*)dstrb, h->n, (int)aa);
  else
    launch_qt<k_sparse_init_rest_fn, float>(h, g, (float*)dstrb, h->n, (int)aa);
  LBM_LAUNCH_CHECK();
  return LBM_OK;
}

lbm_status lbm_sparse_mass(lbm_sparse* h, const void* dstrb, int64_t start, int64_t count,
                           double* mass) {
  LBM_REQUIRE(h && dstrb && mass, LBM_ERR_ARG, "lbm_sparse_mass: null pointer");
  if (lbm_status st = check_range(h, start, count)) return st;
  const int64_t nb = mass_blocks(count);
  double* partial = nullptr;
  LBM_HIP_CHECK(hipMalloc(&partial, sizeof(double) * (nb + 1)));
  lbm_status st = launch_mass(h, dstrb, start, count, partial, nb, partial + nb);
  hipError_t e = hipSuccess;
  if (st == LBM_OK) e = hipMemcpyAsync(mass, partial + nb, sizeof(double), hipMemcpyDeviceToHost,
                                       h->stream);
  if (st == LBM_OK && e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(partial);
  if (st != LBM_OK) return st;
  if (e != hipSuccess) {
    set_error("lbm_sparse_mass: %s", hipGetErrorString(e));
    return LBM_ERR_HIP;
  }
  LBM_REQUIRE(std::isfinite(*mass), LBM_ERR_NUMERIC, "lbm_sparse_mass: mass is not finite (%g)",
              *mass);
  return LBM_OK;
}

lbm_status lbm_sparse_mass_noalloc(lbm_sparse* h, const void* dstrb, int64_t start,
                                   int64_t count, double* mass_device) {
  LBM_REQUIRE(h && dstrb && mass_device, LBM_ERR_ARG, "lbm_sparse_mass_noalloc: null pointer");
  if (lbm_status st = check_range(h, start, count)) return st;
  const int64_t nb = mass_blocks(count);
  LBM_REQUIRE(nb <= kMassScratchPartials, LBM_ERR_ARG,
              "lbm_sparse_mass_noalloc: %lld nodes exceed the fixed scratch (at most %lld nodes "
              "per call); use lbm_sparse_mass or split the range",
              (long long)count, (long long)(kMassScratchPartials * kMassChunk));
  LBM_REQUIRE(device_accessible(dstrb) && device_accessible(mass_device), LBM_ERR_ARG,
              "lbm_sparse_mass_noalloc: the populations and the result must be device-accessible "
              "memory (device, managed or pinned host)");
  double* scr = nullptr;
  if (lbm_status st = scratch_on_current_device(&scr)) return st;
  return launch_mass(h, dstrb, start, count, scr, nb, mass_device);
}

lbm_status lbm_sparse_synchronize(lbm_sparse* h) {
  LBM_REQUIRE(h, LBM_ERR_ARG, "lbm_sparse_synchronize: null handle");
  LBM_HIP_CHECK(hipStreamSynchronize(h->stream));
  return LBM_OK;
}

// Later launches go to `stream` (NULL: the null stream).  Waits for the work
// already enqueued on the previous stream, so the order of calls is kept.
lbm_status lbm_sparse_set_stream(lbm_sparse* h, void* stream) {
  LBM_REQUIRE(h, LBM_ERR_ARG, "lbm_sparse_set_stream: null handle");
  LBM_HIP_CHECK(hipStreamSynchronize(h->stream));
  h->stream = (hipStream_t)stream;
  return LBM_OK;
}

}  // extern "C"
