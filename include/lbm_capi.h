/*
 * lbm_capi.h -- C ABI of the MI355X (gfx950) LBM collide-stream engine.
 *
 * Plain C types only (pointers, sizes, enums); every call returns an
 * lbm_status and never aborts.  All device work is ordered on the stream of
 * the handle it is issued on; the host blocks only in calls that return host
 * data (get_state, moments, mass, run_timed, sync).
 *
 * Two families of entry points:
 *
 *  1. lbm_sparse_*  -- drop-in replacement of the reference's kernel classes.
 *     The caller owns a fluid-only, indirectly addressed node list and the
 *     int32 adjacency that encodes streaming and halfway bounce-back
 *     (src/geometries/GeometryAAAB.cpp:18-124).  Each call is one launch over
 *     the fluid-index range [start, start + count), exactly like
 *       KernelAA::timestepEvenForce / timestepOddForce (src/kernels/KernelAA.hh:28-29)
 *       KernelAB::timestepForce                        (src/kernels/KernelAB.hh:28)
 *     Population arrays are device pointers owned by the caller (allocated with
 *     lbm_device_alloc, as Propagation*::setup allocates its Pdf arrays,
 *     src/propagation/PropagationAA.cpp:10-25).  Arithmetic is the reference
 *     Base kernels' (src/kernels/KernelAA_SOA_Base.cpp:185-247), bit for bit in
 *     FP64 and in FP32 (-D_PRECISION_SINGLE_ semantics).
 *
 *  2. lbm_dense_*   -- the MI355X fast path.  The whole domain of one GPU is a
 *     structured nx*ny*nz box (x fastest); populations are stored SOA as
 *     wavefront-contiguous x-rows (or AOS); neighbours are found by index
 *     arithmetic with periodic wrap in-kernel (the single-GPU replacement of the
 *     MPI halo exchange, src/Communication.cpp:379-481); solid cells and
 *     non-periodic faces bounce back through a per-cell link mask.  Storage can
 *     be FP64, FP32 deviations (f - w_s) or scaled FP16 deviations; the FP64
 *     path is the reference arithmetic bit for bit.  Replaces
 *     PropagationAA::run / PropagationAB::run (src/propagation/PropagationAA.cpp:126-215,
 *     src/propagation/PropagationAB.cpp:155-219) together with the kernels they call.
 *
 * Canonical population order used at the host boundary (set/get_state):
 *   f[s * ncell + cell], cell = x + nx * (y + ny * z), s in reference order,
 *   holding the pre-collision populations f_s(x) at the current time step.
 */
#ifndef LBM_CAPI_H
#define LBM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LBM_OK = 0,
  LBM_ERR_ARG = 1,          /* invalid argument (null pointer, bad size, bad enum) */
  LBM_ERR_HIP = 2,          /* a HIP runtime call failed */
  LBM_ERR_OOM = 3,          /* device allocation failed */
  LBM_ERR_UNSUPPORTED = 4,  /* combination not implemented */
  LBM_ERR_STATE = 5,        /* call not valid in the handle's current state */
  LBM_ERR_NUMERIC = 6       /* NaN / Inf detected by a checking reduction */
} lbm_status;

/* reference propPattern values (src/Parameters.cpp:26, "AB = 0, AA = 1") */
typedef enum { LBM_PATTERN_AB = 0, LBM_PATTERN_AA = 1 } lbm_pattern;
/* Makefile:3 -DSOA / -DAOS */
typedef enum { LBM_LAYOUT_SOA = 0, LBM_LAYOUT_AOS = 1 } lbm_layout;
typedef enum {
  LBM_STORE_F64 = 0,   /* double populations (reference Pdf = double) */
  LBM_STORE_F32S = 1,  /* float deviations f - w_s, FP32 collision, FP64 moment sums */
  LBM_STORE_F16S = 2,  /* half deviations (f - w_s) * scale, FP32 collision, FP64 moments */
  LBM_STORE_F32 = 3    /* float populations, reference -D_PRECISION_SINGLE_ arithmetic (sparse) */
} lbm_storage;

/* Last error message of the calling thread ("" if none). */
const char* lbm_last_error(void);
/* Library version string. */
const char* lbm_version(void);

/* ---------------------------------------------------------------- device memory
 * Replaces the device allocations and copies of the reference's SYCL/Kokkos
 * builds (sycl::malloc_device / q_.memcpy, src/propagation/PropagationAA.cpp:17-23,
 * :117-118, :193-199). */
lbm_status lbm_device_count(int* count);
lbm_status lbm_set_device(int device);
lbm_status lbm_device_alloc(size_t bytes, void** ptr);
lbm_status lbm_device_free(void* ptr);
lbm_status lbm_memcpy_h2d(void* dst, const void* src, size_t bytes);
lbm_status lbm_memcpy_d2h(void* dst, const void* src, size_t bytes);
lbm_status lbm_device_synchronize(void);

/* ---------------------------------------------------------------- sparse (drop-in) */
typedef struct lbm_sparse lbm_sparse;

typedef struct {
  int64_t n_fluid;           /* Geometry::getNumFluidPts (src/geometries/Geometry.hh:68) */
  int32_t q;                 /* 19 | 27 (_LATTICESIZE_, src/Parameters.hh:6-8) */
  int32_t layout;            /* lbm_layout */
  int32_t storage;           /* LBM_STORE_F64 (Pdf = double) | LBM_STORE_F32 (Pdf = float) */
  int32_t adjacency_mode;    /* 0: neighbour at x + c_s (AB push); 1: x - c_s (AA, AB pull) */
  double omega;              /* _OMEGA_ (src/Parameters.hh:26-28) */
  double force[3];           /* (_GRAVITY_, 0, 0) in the reference (src/Parameters.hh:36) */
  const int32_t* adjacency;  /* host, n_fluid * q entries, Geometry::adjacency_ layout */
} lbm_sparse_config;

/* Kernel*::setup (src/kernels/KernelAA_SOA_Base.cpp:8-57) plus the device copy of
 * the adjacency (src/geometries/GeometryAAAB.cpp:115-121). */
lbm_status lbm_sparse_create(const lbm_sparse_config* cfg, lbm_sparse** out);
lbm_status lbm_sparse_destroy(lbm_sparse* h);
/* KernelAA::timestepEvenForce (src/kernels/KernelAA.hh:28): read slot opp(s), write slot s. */
lbm_status lbm_sparse_aa_even(lbm_sparse* h, void* dstrb, int64_t start, int64_t count);
/* KernelAA::timestepOddForce (src/kernels/KernelAA.hh:29): gather/scatter through the adjacency. */
lbm_status lbm_sparse_aa_odd(lbm_sparse* h, void* dstrb, int64_t start, int64_t count);
/* KernelAB::timestepForce (src/kernels/KernelAB.hh:28); pull = 0 -> KernelAB_SOA_Base /
 * KernelAB_AOS_Base (scatter), pull = 1 -> KernelAB_SOA_Pull (gather). */
lbm_status lbm_sparse_ab(lbm_sparse* h, const void* src, void* tgt, int64_t start, int64_t count,
                         int32_t pull);
/* PropagationAA::initialize (aa = 1, w_s into slot opp(s)) / PropagationAB::initialize
 * (aa = 0, w_s into slot s) (src/propagation/PropagationAA.cpp:60-123, PropagationAB.cpp:91-152). */
lbm_status lbm_sparse_init_rest(lbm_sparse* h, void* dstrb, int32_t aa);
/* FP64 mass of the fluid-index range [start, start + count) of a population
 * array: the sum of the Q stored values of each node, accumulated in FP64 by a
 * deterministic two-pass reduction (over the whole range, the total mass in
 * either slot convention).  The reference has no mass reduction; it replaces
 * summing PropagationAA/AB::getRho over the host copy
 * (src/propagation/PropagationAA.cpp:218-257).  Allocates a partial-sum
 * workspace of ceil(count / 4096) + 1 doubles per call, waits for the result;
 * LBM_ERR_NUMERIC if it is not finite. */
lbm_status lbm_sparse_mass(lbm_sparse* h, const void* dstrb, int64_t start, int64_t count,
                           double* mass);
/* lbm_sparse_mass without any allocation.  No counterpart in the reference;
 * for callers that cannot allocate device memory: a time loop captured as a
 * HIP graph on the stream set with lbm_sparse_set_stream, serving loops,
 * harnesses that forbid hipMalloc.  Same kernels and the same bits as
 * lbm_sparse_mass, but the partial sums go to a fixed per-GPU scratch array
 * of 4096 doubles (32 KiB, held by the library until the process ends), and
 * the result is written to `mass_device` (device, managed or pinned host
 * memory) on the handle's stream without waiting and without the finiteness
 * check.  count may be at most 16,777,216 nodes (4096 chunks of 4096);
 * larger ranges return LBM_ERR_ARG.  Calls that overlap on one GPU (other
 * streams, other threads) share the scratch and race: order them. */
lbm_status lbm_sparse_mass_noalloc(lbm_sparse* h, const void* dstrb, int64_t start,
                                   int64_t count, double* mass_device);
/* Stream-order point: waits for every launch issued on the handle. */
lbm_status lbm_sparse_synchronize(lbm_sparse* h);
/* Enqueue later launches on `stream` (a hipStream_t; NULL, the default, is the
 * null stream, ordered with blocking copies as the reference's kernel calls are
 * with its host code).  No counterpart in the reference; for callers that run
 * the kernels on their own stream or capture their time loop as a HIP graph
 * (the step calls only launch kernels).  Waits for the work already enqueued on
 * the previous stream. */
lbm_status lbm_sparse_set_stream(lbm_sparse* h, void* stream);

/* ---------------------------------------------------------------- dense (fast path) */
typedef struct lbm_dense lbm_dense;

typedef struct {
  int32_t nx, ny, nz;
  int32_t q;                 /* 19 | 27 */
  int32_t pattern;           /* lbm_pattern */
  int32_t pull;              /* AB only: 0 push, 1 pull (reference deck key `pull`) */
  int32_t layout;            /* lbm_layout */
  int32_t storage;           /* LBM_STORE_F64 | LBM_STORE_F32S | LBM_STORE_F16S */
  int32_t periodic[3];       /* wrap flags; a non-periodic face bounces back */
  double omega;              /* 1 / tau */
  double force[3];           /* body force (Guo forcing) */
  const uint8_t* solid;      /* host, nx*ny*nz, nonzero = solid; NULL = all fluid */
  const double* tau;         /* host per-cell tau (extension), NULL = uniform omega */
  const double* gamma;       /* host per-cell Brinkman gamma in [0,1] (extension), NULL = none */
  double f16_scale;          /* FP16 deviation scale (power of two); 0 -> 4096 */
  /* Periodic pressure drop along x (inlet x = 0, outlet x = nx - 1): a population
   * crossing the periodic x boundary forward (x = nx - 1 -> 0) gains w_s * drho,
   * one crossing backward loses it, so the inlet density exceeds the outlet's by
   * about drho (pressure drop drho / 3).  0 = plain periodic; needs periodic[0]. */
  double density_jump_x;
} lbm_dense_config;

typedef struct {
  int64_t ncell;             /* nx * ny * nz */
  int64_t n_fluid;           /* non-solid cells (MLUPS counts these) */
  int64_t time_step;         /* completed steps since the last set_state / init */
  double bytes_per_lup;      /* algorithmic HBM bytes per fluid-node update (2 Q b) */
  size_t device_bytes;       /* device memory held by the handle */
} lbm_dense_info;

lbm_status lbm_dense_create(const lbm_dense_config* cfg, lbm_dense** out);
lbm_status lbm_dense_destroy(lbm_dense* h);
lbm_status lbm_dense_get_info(lbm_dense* h, lbm_dense_info* info);
/* Canonical pre-collision state in / out (host arrays of q * ncell doubles);
 * set_state resets time_step to 0.  Solid cells read back as 0.  The device
 * storage order (x-rows plain or interleaved, chosen at create) is internal:
 * every entry point takes and returns cell = x + nx * (y + ny * z) order. */
lbm_status lbm_dense_set_state(lbm_dense* h, const double* f);
lbm_status lbm_dense_get_state(lbm_dense* h, double* f);
/* f = f^eq(rho, u) per cell: rho (ncell, NULL -> 1) and u (3 * ncell, xyz per cell)
 * are host arrays.  Replaces the rest initialisation of PropagationAA/AB::initialize. */
lbm_status lbm_dense_init_equilibrium(lbm_dense* h, const double* rho, const double* u);
/* Device-side Taylor-Green initial state: rho = 1,
 * u = u0 sin(kx x) cos(ky y) cos(kz z), v = -u0 cos(kx x) sin(ky y) cos(kz z), w = 0,
 * k = 2 pi / n per axis, plus noise * (2 U - 1) per component with
 * U = top53(splitmix64(seed * 0x9E3779B97F4A7C15 + 3 * cell + component)) / 2^53. */
lbm_status lbm_dense_init_tgv(lbm_dense* h, double u0, double noise, uint64_t seed);
/* Advance n steps (asynchronous).  PropagationAA::run / PropagationAB::run loop body. */
lbm_status lbm_dense_step(lbm_dense* h, int32_t nsteps);
/* Advance n steps and return their device time (HIP events on the handle's stream). */
lbm_status lbm_dense_run_timed(lbm_dense* h, int32_t nsteps, double* elapsed_ms);
/* rho and u = m / rho per cell of the current pre-collision state, host outputs
 * (ncell and 3 * ncell doubles; solid cells 0).  PropagationAA::getRho / getVelocity
 * (src/propagation/PropagationAA.cpp:218-327) for all cells at once. */
lbm_status lbm_dense_moments(lbm_dense* h, double* rho, double* u);
/* FP64 sum of all fluid populations (deterministic two-pass reduction);
 * LBM_ERR_NUMERIC if the sum is not finite. */
lbm_status lbm_dense_mass(lbm_dense* h, double* mass);
lbm_status lbm_dense_synchronize(lbm_dense* h);
/* Streams and graphs.  No counterpart in the reference (whose propagation runs
 * on the host); for callers that order the engine with their own device work
 * or replay its time loop as a HIP graph.
 *
 * lbm_dense_set_stream: enqueue every later call on `stream` (a hipStream_t;
 * NULL returns to the handle's own stream).  Waits for the work already
 * enqueued on the previous stream.  The stream must outlive its use by the
 * handle.
 *
 * lbm_dense_graph_create: capture `nsteps` time steps (even, > 0) as a HIP graph
 * of kernel launches only (no memory-set or copy nodes).  Nothing runs and the
 * handle's state is unchanged.  lbm_dense_graph_launch replays it `nlaunch`
 * times on the handle's current stream and advances time_step by
 * nlaunch * nsteps; it returns LBM_ERR_ARG when the handle is an odd number of
 * steps away from where it was at capture.  A graph belongs to its handle:
 * lbm_dense_destroy returns LBM_ERR_STATE while graphs of the handle are alive. */
typedef struct lbm_dense_graph lbm_dense_graph;
lbm_status lbm_dense_set_stream(lbm_dense* h, void* stream);
lbm_status lbm_dense_graph_create(lbm_dense* h, int32_t nsteps, lbm_dense_graph** out);
lbm_status lbm_dense_graph_launch(lbm_dense_graph* g, int32_t nlaunch);
lbm_status lbm_dense_graph_destroy(lbm_dense_graph* g);

/* ---------------------------------------------------------------- discrete adjoint
 * No counterpart in the reference.  Gradients of an objective of the final
 * state after T collide-stream steps with respect to the per-cell Brinkman
 * field gamma and relaxation-time field tau.  Forward model: the dense
 * engine's AB-push step (reference BGK + Guo arithmetic, halfway bounce-back)
 * with omega = 1/tau(x) and u = (1 - gamma(x)) (m + F/2) / rho.  Objectives:
 * 0 = sum_x |m|^2 / (2 rho) (kinetic energy), 1 = sum over the plane x = outlet_x
 * of m_x (outlet flux).  The forward pass stores a checkpoint every
 * `checkpoint_every` steps; the backward pass recomputes one segment at a time
 * and runs the exact transpose of each step.  Nested levels bound the memory
 * for long runs: with checkpoint_inner = {k2, k3} a segment of k1 steps is
 * recomputed once storing a sub-checkpoint every k2 steps, each sub-segment
 * likewise every k3 steps, and only the innermost segment is held whole.
 * States held besides the two adjoint arrays: ceil(T/k1) + (ceil(k1/k2) - 1)
 * + (ceil(k2/k3) - 1) + (k_last - 1), at least 2 for the last term when nested
 * (unused levels drop out); each level costs one more forward pass.  The
 * gradients do not depend on the intervals (bit for bit). */
typedef struct lbm_adjoint lbm_adjoint;

typedef struct {
  int32_t nx, ny, nz;
  int32_t q;                 /* 19 | 27 */
  int32_t storage;           /* LBM_STORE_F64 (FP64 forward + adjoint) | LBM_STORE_F32S (FP32) */
  int32_t periodic[3];
  double omega;              /* used where tau is NULL */
  double force[3];
  const uint8_t* solid;      /* host mask or NULL */
  const double* tau;         /* host per-cell tau or NULL (then no tau gradient) */
  const double* gamma;       /* host per-cell gamma or NULL (then no gamma gradient) */
  int32_t objective;         /* 0 kinetic energy, 1 outlet flux */
  int32_t outlet_x;          /* outlet plane for objective 1 */
  int32_t checkpoint_every;  /* k1 >= 1 */
  int32_t checkpoint_inner[2]; /* {k2, k3}: 0 = level unused; k1 >= k2 >= k3 >= 1 where used */
  double density_jump_x;     /* pressure-drop boundary of the forward model (lbm_dense_config) */
} lbm_adjoint_config;

lbm_status lbm_adjoint_create(const lbm_adjoint_config* cfg, lbm_adjoint** out);
lbm_status lbm_adjoint_destroy(lbm_adjoint* h);
/* Initial canonical state (q * ncell doubles); default is the rest state. */
lbm_status lbm_adjoint_set_initial_state(lbm_adjoint* h, const double* f);
/* Forward pass of nsteps steps with checkpoints; returns J (FP64 reduction). */
lbm_status lbm_adjoint_forward(lbm_adjoint* h, int32_t nsteps, double* objective);
/* Backward pass over the last forward pass; host gradients (ncell doubles each,
 * NULL to skip) and the device time of the pass including recomputation. */
lbm_status lbm_adjoint_backward(lbm_adjoint* h, double* grad_gamma, double* grad_tau,
                                double* elapsed_ms);
/* Adjoint state lambda^0 = dJ/df^0 after a backward pass (q * ncell doubles). */
lbm_status lbm_adjoint_get_lambda0(lbm_adjoint* h, double* lambda);
/* Final state f^T of the last forward pass (canonical, q * ncell doubles). */
lbm_status lbm_adjoint_get_final_state(lbm_adjoint* h, double* f);
/* Central finite difference dJ/dparam at one cell (param 0 = gamma, 1 = tau) of
 * order 2, 4 or 6 from D_k = J(p + k h) - J(p - k h), each D_k summed node by
 * node on the device: D_1 / 2h; (8 D_1 - D_2) / 12h; (45 D_1 - 9 D_2 + D_3) / 60h.
 * The round-off of D_k divided by h sets the usable step from below, the
 * truncation (h^order) from above.  Invalidates the last forward pass. */
lbm_status lbm_adjoint_fd(lbm_adjoint* h, int32_t param, int64_t cell, double step,
                          int32_t nsteps, int32_t order, double* derivative);
/* The same central difference along a whole-field direction v (host, ncell
 * doubles; solid cells ignored): d/de J(p + e v) at e = 0, whose adjoint value
 * is sum_x grad(x) v(x).  One cell's influence on J can sit below the FP64
 * round-off of J itself on large grids; a direction over many cells keeps the
 * difference well above it.  tau + e v must stay above 1/2. */
lbm_status lbm_adjoint_fd_direction(lbm_adjoint* h, int32_t param, const double* direction,
                                    double step, int32_t nsteps, int32_t order,
                                    double* derivative);
lbm_status lbm_adjoint_device_bytes(lbm_adjoint* h, size_t* bytes);

/* ---------------------------------------------------------------- geometry and fields
 * Host-side generators (no device work).  Replace Geometry::setupGrid
 * (src/geometries/Geometry.cpp:100-151), whose only case is the cylinder.
 * Masks: solid[x + nx * (y + ny * z)] != 0 marks a solid cell. */
/* Reference cylinder for simulationSize s (src/geometries/Geometry.cpp:60-70,113-131):
 * dims = {int(84 s), 2 int(8 s) + 3, 2 int(8 s) + 3}; solid may be NULL to query dims. */
lbm_status lbm_geometry_cylinder(double simulation_size, int32_t* dims, uint8_t* solid);
/* Solid layers at y = 0 and y = ny - 1. */
lbm_status lbm_geometry_channel(int32_t nx, int32_t ny, int32_t nz, uint8_t* solid);
/* Seeded random spheres (radius in [rmin, rmax], wrapping periodic faces) added to
 * `solid` until the solid fraction reaches `fraction`; *count = spheres drawn. */
lbm_status lbm_geometry_add_spheres(int32_t nx, int32_t ny, int32_t nz, const int32_t* periodic,
                                    double rmin, double rmax, double fraction, uint64_t seed,
                                    uint8_t* solid, int64_t* count);
/* Seeded smooth field in [0, 1]: white noise, periodic Gaussian filter (sigma,
 * truncated at 4 sigma), min-max rescaled. */
lbm_status lbm_field_smooth_random(int32_t nx, int32_t ny, int32_t nz, double sigma,
                                   uint64_t seed, double* out);

/* ---------------------------------------------------------------- measurement helper */
/* Device-to-device copy kernel of `bytes` (multiple of 16) between two buffers
 * allocated by the library; used to calibrate achieved-bandwidth and
 * HBM-counter readings.  Returns the time of `reps` launches in ms. */
lbm_status lbm_copy_bandwidth(size_t bytes, int32_t reps, double* elapsed_ms);

#ifdef __cplusplus
}
#endif
#endif /* LBM_CAPI_H */
