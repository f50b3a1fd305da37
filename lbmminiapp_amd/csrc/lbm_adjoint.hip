// lbm_adjoint.hip -- discrete adjoint of the collide-stream step (forward /
// backward pair) with checkpoint-every-k recomputation.  No counterpart in the
// reference; its forward model is the reference step (BGK + Guo forcing,
// halfway bounce-back, src/kernels/KernelAB_SOA_Base.cpp:111-186) with the two
// per-node extensions of this engine: omega = 1/tau(x) and Brinkman velocity
// scaling u = (1 - gamma(x)) (m + F/2) / rho.
//
// Forward map of one step on the canonical state (f_s(x) in slot s at x):
//   g = C(f) node by node, then push streaming g_s(x) -> x + c_s (or slot opp(s)
//   at x when x + c_s is solid).  The dense engine's AB-push kernels run it.
// Objective J(f^T): 0 = sum_x |m|^2 / (2 rho) (kinetic energy of the final
// state), 1 = sum_{x: x-index == outlet} m_x (outlet flux).
// Backward step (exact transpose), per node x with lambda^{t+1} given:
//   lambda_g_s(x) = lambda^{t+1}[slot s at x + c_s]  (slot opp(s) at x if solid)
//   rho, m, u* = (m + F/2)/rho, u = (1 - gamma) u* from f^t(x)
//   a = w sum_s lambda_g_s eq_s / rho
//   B = sum_s lambda_g_s (w d eq_s/du + (1 - w/2) d F_s/du)
//     d eq_s/du = w_s rho (3 c_s + 3 cu_s c_s - 3 u),  d F_s/du = w_s (-3 F + 3 cf_s c_s)
//   lambda^t_k(x) = (1 - w) lambda_g_k + a + B . ((1 - gamma) c_k - u) / rho
//   dJ/dgamma(x) += -B . u*,   dJ/dtau(x) += -w^2 sum_s lambda_g_s (eq_s - f_s - F_s / 2)
// Memory: ceil(T/k) checkpoints + (k - 1) recomputed states of one segment +
// two adjoint arrays; every forward step is computed twice.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "lbm_dense_internal.hpp"

namespace lbm {

template <int STORE>
struct AdjType;
template <>
struct AdjType<LBM_STORE_F64> {
  using R = double;
};
template <>
struct AdjType<LBM_STORE_F32S> {
  using R = float;
};

template <int Q, int STORE>
__device__ __forceinline__ double full_value(typename Storage<STORE>::C v, int s) {
  if constexpr (STORE == LBM_STORE_F64) return v;
  else return Lattice<Q>::w(s) + (double)v;
}

template <int Q, int STORE, bool MASK, int WAVES>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 256), amdgpu_waves_per_eu(WAVES)))
    k_adj_step(const void* fv, const void* lin, void* lout, double* __restrict__ grad_gamma,
               double* __restrict__ grad_tau, DenseArgs a) {
  using T = typename Storage<STORE>::T;
  using R = typename AdjType<STORE>::R;
  const T* f = static_cast<const T*>(fv);
  const R* lam_in = static_cast<const R*>(lin);
  R* lam_out = static_cast<R*>(lout);
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.nx) return;
  const int y = blockIdx.y, z = blockIdx.z;
  const int64_t cell = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
  uint32_t lk = 0;
  const int64_t n = a.ncell;
  if constexpr (MASK) {
    lk = a.links[cell];
    if (lk & SOLID_BIT) {
      // a solid cell has no populations: its adjoint slots are zero in both
      // ping-pong arrays, so lambda^0 reads back as 0 there
#pragma unroll
      for (int k = 0; k < Q; k++) lam_out[slot<Q, LBM_LAYOUT_SOA>(n, k, cell)] = R(0);
      return;
    }
  }
  const Offsets off(x, y, z, a.nx, a.ny, a.nz);
  double fs[Q];
  R lg[Q];
  double rho = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const int cx = Lattice<Q>::c(s, 0), cy = Lattice<Q>::c(s, 1), cz = Lattice<Q>::c(s, 2);
    // the state and the row-aligned adjoint values are read once: stream them
    // past the caches; x-shifted reads share lines with the neighbouring wave
    fs[s] = full_value<Q, STORE>(ld_nt<STORE>(f, slot<Q, LBM_LAYOUT_SOA>(n, s, cell), a.inv_scale), s);
    rho += fs[s];
    m0 += cx * fs[s];
    m1 += cy * fs[s];
    m2 += cz * fs[s];
    const bool bb = MASK && ((lk >> s) & 1u);
    const int64_t li = bb ? slot<Q, LBM_LAYOUT_SOA>(n, Lattice<Q>::opp(s), cell)
                          : slot<Q, LBM_LAYOUT_SOA>(n, s, cell + off.at(cx, cy, cz));
    lg[s] = cx == 0 ? __builtin_nontemporal_load(lam_in + li) : lam_in[li];
  }
  const double om = a.omega_field ? a.omega_field[cell] : a.omega;
  const double gam = a.gamma_field ? a.gamma_field[cell] : 0.0;
  const double us0 = (m0 + 0.5 * a.F0) / rho, us1 = (m1 + 0.5 * a.F1) / rho,
               us2 = (m2 + 0.5 * a.F2) / rho;
  const R u0 = (R)((1.0 - gam) * us0), u1 = (R)((1.0 - gam) * us1), u2 = (R)((1.0 - gam) * us2);
  const R F0 = (R)a.F0, F1 = (R)a.F1, F2 = (R)a.F2;
  const R rr = (R)rho, omr = (R)om;
  const R uu = R(3) * (u0 * u0 + u1 * u1 + u2 * u2);
  const R fu = R(3) * (F0 * u0 + F1 * u1 + F2 * u2);
  R acc_a = 0, B0 = 0, B1 = 0, B2 = 0, dw = 0;
#pragma unroll
  for (int s = 0; s < Q; s++) {
    const R c0 = Lattice<Q>::c(s, 0), c1 = Lattice<Q>::c(s, 1), c2 = Lattice<Q>::c(s, 2);
    const R w = (R)Lattice<Q>::w(s);
    const R cu = R(3) * (c0 * u0 + c1 * u1 + c2 * u2);
    const R cf = R(3) * (c0 * F0 + c1 * F1 + c2 * F2);
    // eq - f without forming the O(1) weight twice: w rho (1 + P) - f
    const R P = cu + R(0.5) * cu * cu - R(0.5) * uu;
    const R eq_minus_f = (R)(w * (rr - R(1)) + w * rr * P) - (R)(fs[s] - (double)w);
    const R eq = w * rr * (R(1) + P);
    const R fd = w * (cf - fu + cu * cf);
    acc_a += lg[s] * eq;
    dw += lg[s] * (eq_minus_f - R(0.5) * fd);
    const R ge = omr * w * rr * R(3);  // omega d eq/du = ge (c + cu c - u)
    const R gf = (R(1) - R(0.5) * omr) * w * R(3);  // (1 - omega/2) d F/du = gf (cf c - F)
    B0 += lg[s] * (ge * (c0 + cu * c0 - u0) + gf * (cf * c0 - F0));
    B1 += lg[s] * (ge * (c1 + cu * c1 - u1) + gf * (cf * c1 - F1));
    B2 += lg[s] * (ge * (c2 + cu * c2 - u2) + gf * (cf * c2 - F2));
  }
  const R aa = omr * acc_a / rr;
  const R sc = R(1.0 - gam);
#pragma unroll
  for (int k = 0; k < Q; k++) {
    const R c0 = Lattice<Q>::c(k, 0), c1 = Lattice<Q>::c(k, 1), c2 = Lattice<Q>::c(k, 2);
    const R v = (R(1) - omr) * lg[k] + aa +
                (B0 * (sc * c0 - u0) + B1 * (sc * c1 - u1) + B2 * (sc * c2 - u2)) / rr;
    __builtin_nontemporal_store(v, lam_out + slot<Q, LBM_LAYOUT_SOA>(n, k, cell));
  }
  if (grad_gamma) grad_gamma[cell] += -((double)B0 * us0 + (double)B1 * us1 + (double)B2 * us2);
  if (grad_tau) grad_tau[cell] += -om * om * (double)dw;
}

// Tree sum over the workgroup, for any blockDim.x <= 256 (the seed and
// difference kernels run blocks of 64-256, e.g. 192 for rows of 129-192).
// The halving starts at the power of two p with p/2 < blockDim.x <= p; a lane
// adds its partner only when the partner exists, so every entry is summed once.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int nb = (int)blockDim.x;
  red[threadIdx.x] = v;
  __syncthreads();
  int p = 1;
  while (p < nb) p <<= 1;
  for (int w = p >> 1; w > 0; w >>= 1) {
    const int t = (int)threadIdx.x;
    if (t < w && t + w < nb) red[t] += red[t + w];
    __syncthreads();
  }
  return red[0];
}

template <int Q, int STORE>
__device__ __forceinline__ double node_objective(const typename Storage<STORE>::T* f, int64_t n,
                                                 int64_t cell, int x, int objective,
                                                 int outlet_x, float inv_scale, double& rho,
                                                 double* m) {
  rho = 0.0;
  m[0] = m[1] = m[2] = 0.0;
  for (int s = 0; s < Q; s++) {
    const double v =
        full_value<Q, STORE>(ld<STORE>(f, slot<Q, LBM_LAYOUT_SOA>(n, s, cell), inv_scale), s);
    rho += v;
    m[0] += Lattice<Q>::c(s, 0) * v;
    m[1] += Lattice<Q>::c(s, 1) * v;
    m[2] += Lattice<Q>::c(s, 2) * v;
  }
  if (objective == 0) return 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rho;
  return x == outlet_x ? m[0] : 0.0;
}

// lambda^T = dJ/df^T and per-block partial sums of J (blocks of 256 over (x, y, z)).
template <int Q, int STORE>
__global__ void __launch_bounds__(256)
    k_adj_seed(const void* fv, void* lv, DenseArgs a, int objective, int outlet_x,
               double* partial) {
  __shared__ double red[256];
  using T = typename Storage<STORE>::T;
  using R = typename AdjType<STORE>::R;
  const T* f = static_cast<const T*>(fv);
  R* lam = static_cast<R*>(lv);
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  double jn = 0.0;
  if (x < a.nx) {
    const int64_t cell = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
    const bool solid = a.links && (a.links[cell] & SOLID_BIT);
    double rho, m[3];
    if (!solid) jn = node_objective<Q, STORE>(f, a.ncell, cell, x, objective, outlet_x,
                                              a.inv_scale, rho, m);
    const double v0 = solid ? 0.0 : m[0] / rho, v1 = solid ? 0.0 : m[1] / rho,
                 v2 = solid ? 0.0 : m[2] / rho;
    const double vv = v0 * v0 + v1 * v1 + v2 * v2;
    for (int s = 0; s < Q; s++) {
      double g = 0.0;
      if (!solid) {
        if (objective == 0)
          g = Lattice<Q>::c(s, 0) * v0 + Lattice<Q>::c(s, 1) * v1 + Lattice<Q>::c(s, 2) * v2 -
              0.5 * vv;
        else
          g = x == outlet_x ? (double)Lattice<Q>::c(s, 0) : 0.0;
      }
      lam[slot<Q, LBM_LAYOUT_SOA>(a.ncell, s, cell)] = (R)g;
    }
  }
  const double tot = block_sum(jn, red);
  if (threadIdx.x == 0)
    partial[blockIdx.x + gridDim.x * ((int64_t)blockIdx.y + (int64_t)gridDim.y * blockIdx.z)] = tot;
}

// Per-block partial sums of J(fa) - J(fb), node by node (finite differences).
template <int Q, int STORE>
__global__ void __launch_bounds__(256)
    k_obj_diff(const void* fav, const void* fbv, DenseArgs a, int objective, int outlet_x,
               double* partial) {
  __shared__ double red[256];
  using T = typename Storage<STORE>::T;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  double dj = 0.0;
  if (x < a.nx) {
    const int64_t cell = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
    if (!(a.links && (a.links[cell] & SOLID_BIT))) {
      double r, m[3];
      const double ja = node_objective<Q, STORE>(static_cast<const T*>(fav), a.ncell, cell, x,
                                                 objective, outlet_x, a.inv_scale, r, m);
      const double jb = node_objective<Q, STORE>(static_cast<const T*>(fbv), a.ncell, cell, x,
                                                 objective, outlet_x, a.inv_scale, r, m);
      dj = ja - jb;
    }
  }
  const double tot = block_sum(dj, red);
  if (threadIdx.x == 0)
    partial[blockIdx.x + gridDim.x * ((int64_t)blockIdx.y + (int64_t)gridDim.y * blockIdx.z)] = tot;
}

__global__ void __launch_bounds__(256) k_adj_sum(const double* partial, int64_t n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) *out = tot;
}

}  // namespace lbm

using namespace lbm;

struct lbm_adjoint {
  lbm_dense* d = nullptr;  // AB push engine: stream, links, fields, two state arrays
  int objective = 0, outlet_x = 0, every = 1;
  std::vector<int64_t> inner;  // nested checkpoint intervals below `every` (k2, k3)
  int64_t T = -1;          // steps of the last forward pass
  std::vector<void*> ckpt;
  std::vector<std::vector<void*>> level;  // level[l]: sub-checkpoints of an inner[l] level
  void* seg = nullptr;
  int64_t seg_states = 0;
  void* lam[2] = {nullptr, nullptr};
  int lam_cur = 0;
  double* grad[2] = {nullptr, nullptr};  // d/dgamma, d/dtau (device, ncell)
  void* f0 = nullptr;      // initial state
  void* final_state = nullptr;
  void* fd[2] = {nullptr, nullptr};
  double* partial = nullptr;
  int64_t npartial = 0;
  size_t state_bytes = 0, lam_bytes = 0, extra_bytes = 0;
  std::vector<double> tau_host, gamma_host;
  std::vector<uint8_t> solid_host;
  double J = 0.0;
  int step_bx = 256;        // transposed-step workgroup width
  size_t step_lds = 0;      // its dynamic LDS floor (caps workgroups per CU)
  int step_waves = 1;       // waves per SIMD the transposed step is compiled for (1: no bound)
};

namespace {

dim3 adj_block(const lbm_dense* d) { return block_for(d->nx); }
dim3 adj_grid(const lbm_dense* d) { return grid_for(d); }

template <template <int, int> class K, typename... A>
void launch_qs(const lbm_dense* d, dim3 g, dim3 b, A... args) {
  if (d->q == 19) {
    if (d->storage == LBM_STORE_F64) K<19, LBM_STORE_F64>::launch(g, b, d->stream, args...);
    else K<19, LBM_STORE_F32S>::launch(g, b, d->stream, args...);
  } else {
    if (d->storage == LBM_STORE_F64) K<27, LBM_STORE_F64>::launch(g, b, d->stream, args...);
    else K<27, LBM_STORE_F32S>::launch(g, b, d->stream, args...);
  }
}

template <int Q, int S>
struct seed_u {
  template <typename... A>
  static void launch(dim3 g, dim3 b, hipStream_t st, A... a) {
    hipLaunchKernelGGL((k_adj_seed<Q, S>), g, b, 0, st, a...);
  }
};
template <int Q, int S>
struct objdiff_u {
  template <typename... A>
  static void launch(dim3 g, dim3 b, hipStream_t st, A... a) {
    hipLaunchKernelGGL((k_obj_diff<Q, S>), g, b, 0, st, a...);
  }
};
template <int Q, int S>
struct step_u {
  template <bool MASK, typename... A>
  static void launch_w(dim3 g, dim3 b, hipStream_t st, size_t lds, int waves, A... a) {
    switch (waves) {
      case 2: hipLaunchKernelGGL((k_adj_step<Q, S, MASK, 2>), g, b, lds, st, a...); break;
      case 3: hipLaunchKernelGGL((k_adj_step<Q, S, MASK, 3>), g, b, lds, st, a...); break;
      case 4: hipLaunchKernelGGL((k_adj_step<Q, S, MASK, 4>), g, b, lds, st, a...); break;
      default: hipLaunchKernelGGL((k_adj_step<Q, S, MASK, 1>), g, b, lds, st, a...); break;
    }
  }
  template <typename... A>
  static void launch(dim3 g, dim3 b, hipStream_t st, size_t lds, bool mask, int waves, A... a) {
    if (mask) launch_w<true>(g, b, st, lds, waves, a...);
    else launch_w<false>(g, b, st, lds, waves, a...);
  }
};

lbm_status reduce_to_host(lbm_adjoint* h, double* out) {
  hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, h->d->stream, (const double*)h->partial,
                     h->npartial, h->partial + h->npartial);
  LBM_LAUNCH_CHECK();
  LBM_HIP_CHECK(hipMemcpyAsync(out, h->partial + h->npartial, sizeof(double),
                               hipMemcpyDeviceToHost, h->d->stream));
  LBM_HIP_CHECK(hipStreamSynchronize(h->d->stream));
  return LBM_OK;
}

// T push steps from f0; returns the buffer that holds f^T.  With checkpoints,
// the state at t = i * every is copied to ckpt[i] before step t.
lbm_status run_forward(lbm_adjoint* h, int64_t T, bool checkpoints, void** final_out) {
  lbm_dense* d = h->d;
  void* src = h->f0;
  int nxt = 0;
  for (int64_t t = 0; t < T; t++) {
    if (checkpoints && t % h->every == 0)
      LBM_HIP_CHECK(hipMemcpyAsync(h->ckpt[t / h->every], src, h->state_bytes,
                                   hipMemcpyDeviceToDevice, d->stream));
    void* dst = d->f[nxt];
    nxt = 1 - nxt;
    if (lbm_status st = launch_kind(d, K_AB_PUSH, src, dst)) return st;
    src = dst;
  }
  *final_out = src;
  return LBM_OK;
}

lbm_status ensure_buffers(lbm_adjoint* h, int64_t T) {
  lbm_dense* d = h->d;
  const int64_t nseg = (T + h->every - 1) / h->every;
  while ((int64_t)h->ckpt.size() < nseg) {
    void* p = nullptr;
    if (lbm_status st = dmalloc(d, &p, h->state_bytes, "adjoint checkpoint")) return st;
    h->ckpt.push_back(p);
  }
  // A segment of an outer level spans at most `outer` steps; the level below
  // keeps its states at offsets k, 2k, ... inside it (the start is the outer
  // level's own state).
  int64_t outer = std::min<int64_t>(h->every, T);
  h->level.resize(h->inner.size());
  for (size_t l = 0; l < h->inner.size(); l++) {
    const int64_t k = std::min<int64_t>(h->inner[l], outer);
    const int64_t need = (outer + k - 1) / k - 1;
    while ((int64_t)h->level[l].size() < need) {
      void* p = nullptr;
      if (lbm_status st = dmalloc(d, &p, h->state_bytes, "adjoint sub-checkpoint")) return st;
      h->level[l].push_back(p);
    }
    outer = k;
  }
  // innermost segment, held whole; with nested levels it also serves as the
  // two-state scratch for the recomputation between sub-checkpoints
  int64_t need_seg = outer - 1;
  if (!h->inner.empty()) need_seg = std::max<int64_t>(need_seg, 2);
  if (need_seg > h->seg_states) {
    if (h->seg) {
      (void)hipFree(h->seg);
      d->device_bytes -= h->state_bytes * h->seg_states;
      h->seg = nullptr;
      h->seg_states = 0;
    }
    if (lbm_status st = dmalloc(d, &h->seg, h->state_bytes * need_seg, "adjoint segment"))
      return st;
    h->seg_states = need_seg;
  }
  return LBM_OK;
}

// The whole gamma (param 0) or tau (param 1) field p0 + e v on the device (solid
// cells keep the values create gave them); tau must stay above 1/2.
lbm_status set_field(lbm_adjoint* h, int param, const double* v, double e) {
  lbm_dense* d = h->d;
  const std::vector<double>& p0 = param == 0 ? h->gamma_host : h->tau_host;
  std::vector<double> field(d->ncell);
  for (int64_t c = 0; c < d->ncell; c++) {
    const double p = p0[c] + e * v[c];
    if (param == 0) {
      field[c] = p;
    } else {
      const bool solid = !h->solid_host.empty() && h->solid_host[c];
      LBM_REQUIRE(solid || (std::isfinite(p) && p > 0.5), LBM_ERR_ARG,
                  "lbm_adjoint_fd_direction: tau + e v = %g at cell %lld is not above 1/2", p,
                  (long long)c);
      field[c] = solid ? 1.0 : 1.0 / p;
    }
  }
  double* dst = param == 0 ? d->gamma_field : d->omega_field;
  LBM_HIP_CHECK(hipStreamSynchronize(d->stream));
  LBM_HIP_CHECK(hipMemcpy(dst, field.data(), sizeof(double) * d->ncell, hipMemcpyHostToDevice));
  return LBM_OK;
}

lbm_status set_param(lbm_adjoint* h, int param, int64_t cell, double value) {
  lbm_dense* d = h->d;
  double* field = param == 0 ? d->gamma_field : d->omega_field;
  const double v = param == 0 ? value : 1.0 / value;
  LBM_HIP_CHECK(hipStreamSynchronize(d->stream));
  LBM_HIP_CHECK(hipMemcpy(field + cell, &v, sizeof(double), hipMemcpyHostToDevice));
  return LBM_OK;
}

// Adjoint steps over the `len` steps that start from state `base`, walking
// the checkpoint levels from `l` inwards; lam[cur] holds the adjoint state at
// the end of the range on entry and at its start on return.
lbm_status backward_range(lbm_adjoint* h, const void* base, int64_t len, size_t l, int& cur) {
  lbm_dense* d = h->d;
  if (l == h->inner.size()) {
    // innermost: recompute states 1 .. len - 1 into the segment buffer
    const void* src = base;
    for (int64_t j = 1; j < len; j++) {
      void* dst = (char*)h->seg + (j - 1) * h->state_bytes;
      if (lbm_status st = launch_kind(d, K_AB_PUSH, const_cast<void*>(src), dst)) return st;
      src = dst;
    }
    const dim3 b(h->step_bx), g((d->nx + h->step_bx - 1) / h->step_bx, d->ny, d->nz);
    const bool mask = d->links != nullptr;
    for (int64_t j = len - 1; j >= 0; j--) {
      const void* state = j == 0 ? base : (char*)h->seg + (j - 1) * h->state_bytes;
      launch_qs<step_u>(d, g, b, h->step_lds, mask, h->step_waves, state, (const void*)h->lam[cur],
                        h->lam[1 - cur], h->grad[0], h->grad[1], d->args);
      LBM_LAUNCH_CHECK();
      cur = 1 - cur;
    }
    return LBM_OK;
  }
  const int64_t k = h->inner[l];
  const int64_t nsub = (len + k - 1) / k;
  std::vector<void*>& sub = h->level[l];
  // recompute once, keeping the states at offsets k, 2k, ..., (nsub - 1) k;
  // the states in between go to two scratch slots of the segment buffer
  const void* src = base;
  for (int64_t t = 1; t <= (nsub - 1) * k; t++) {
    void* dst = t % k == 0 ? sub[t / k - 1] : (char*)h->seg + (t % 2) * h->state_bytes;
    if (lbm_status st = launch_kind(d, K_AB_PUSH, const_cast<void*>(src), dst)) return st;
    src = dst;
  }
  for (int64_t j = nsub - 1; j >= 0; j--) {
    const void* start = j == 0 ? base : sub[j - 1];
    if (lbm_status st = backward_range(h, start, std::min(k, len - j * k), l + 1, cur)) return st;
  }
  return LBM_OK;
}

void free_adjoint(lbm_adjoint* h) {
  if (!h) return;
  for (void* p : h->ckpt) (void)hipFree(p);
  for (auto& lv : h->level)
    for (void* p : lv) (void)hipFree(p);
  if (h->seg) (void)hipFree(h->seg);
  for (int i = 0; i < 2; i++) {
    if (h->lam[i]) (void)hipFree(h->lam[i]);
    if (h->grad[i]) (void)hipFree(h->grad[i]);
    if (h->fd[i]) (void)hipFree(h->fd[i]);
  }
  if (h->f0) (void)hipFree(h->f0);
  if (h->partial) (void)hipFree(h->partial);
  if (h->d) lbm_dense_destroy(h->d);
  delete h;
}

}  // namespace

extern "C" {

lbm_status lbm_adjoint_create(const lbm_adjoint_config* cfg, lbm_adjoint** out) {
  LBM_REQUIRE(cfg && out, LBM_ERR_ARG, "lbm_adjoint_create: null pointer");
  *out = nullptr;
  LBM_REQUIRE(cfg->storage == LBM_STORE_F64 || cfg->storage == LBM_STORE_F32S, LBM_ERR_ARG,
              "lbm_adjoint_create: storage must be LBM_STORE_F64 or LBM_STORE_F32S");
  LBM_REQUIRE(cfg->objective == 0 || cfg->objective == 1, LBM_ERR_ARG,
              "lbm_adjoint_create: objective must be 0 (kinetic energy) or 1 (outlet flux)");
  LBM_REQUIRE(cfg->checkpoint_every >= 1, LBM_ERR_ARG,
              "lbm_adjoint_create: checkpoint_every must be >= 1");
  LBM_REQUIRE(cfg->objective == 0 || (cfg->outlet_x >= 0 && cfg->outlet_x < cfg->nx),
              LBM_ERR_ARG, "lbm_adjoint_create: outlet_x outside [0, nx)");
  const int32_t k2 = cfg->checkpoint_inner[0], k3 = cfg->checkpoint_inner[1];
  LBM_REQUIRE(k2 >= 0 && k3 >= 0 && (k2 > 0 || k3 == 0) && k2 <= cfg->checkpoint_every &&
                  k3 <= (k2 > 0 ? k2 : 0),
              LBM_ERR_ARG,
              "lbm_adjoint_create: checkpoint_inner must be {0, 0}, {k2, 0} or {k2, k3} with "
              "checkpoint_every >= k2 >= k3 >= 1");
  lbm_dense_config dc{};
  dc.nx = cfg->nx;
  dc.ny = cfg->ny;
  dc.nz = cfg->nz;
  dc.q = cfg->q;
  dc.pattern = LBM_PATTERN_AB;
  dc.pull = 0;
  dc.layout = LBM_LAYOUT_SOA;
  dc.storage = cfg->storage;
  for (int k = 0; k < 3; k++) {
    dc.periodic[k] = cfg->periodic[k];
    dc.force[k] = cfg->force[k];
  }
  dc.omega = cfg->omega;
  dc.density_jump_x = cfg->density_jump_x;  // additive on crossing links: the transpose is unchanged
  dc.solid = cfg->solid;
  dc.tau = cfg->tau;
  dc.gamma = cfg->gamma;
  lbm_adjoint* h = new lbm_adjoint;
  if (lbm_status st = dense_create(&dc, &h->d, false, false)) {
    delete h;
    return st;
  }
  lbm_dense* d = h->d;
  h->objective = cfg->objective;
  h->outlet_x = cfg->outlet_x;
  h->every = cfg->checkpoint_every;
  if (k2 > 0) h->inner.push_back(k2);
  if (k3 > 0) h->inner.push_back(k3);
  // transposed-step launch shape (LBM_ADJ_BLOCK_X, LBM_ADJ_BLOCKS_PER_CU: tuning)
  {
    const char* e = std::getenv("LBM_ADJ_BLOCK_X");
    int bx = e ? std::atoi(e) : (int)adj_block(d).x;
    bx = bx < 64 ? 64 : (bx > 256 ? 256 : bx / 64 * 64);
    h->step_bx = bx;
    const char* k = std::getenv("LBM_ADJ_BLOCKS_PER_CU");
    h->step_lds = lds_for_blocks_per_cu(k ? std::atoi(k) : 0);
    // Waves per SIMD: unbounded, the D3Q19 FP64 step needs 256 VGPRs + 6 AGPRs
    // (one wave per SIMD); bounded to 2 it spills 6 and the 128^3 channel
    // backward pass ran 154.8 instead of 169.8 ms (same gradients, bit for
    // bit).  Bounds of 3-4 spill 94+ and ran slower; FP32 deviations ran best
    // unbounded (74.9 vs 75.4 ms at 2); D3Q27 spills 90+ at 2.
    // LBM_ADJ_WAVES=1..4 overrides (tuning).
    const char* w = std::getenv("LBM_ADJ_WAVES");
    h->step_waves = w ? std::atoi(w) : (d->q == 19 && d->storage == LBM_STORE_F64 ? 2 : 1);
  }
  h->state_bytes = (size_t)d->ncell * d->q * d->esize;
  h->lam_bytes = (size_t)d->ncell * d->q * (d->storage == LBM_STORE_F64 ? 8 : 4);
  if (cfg->tau) h->tau_host.assign(cfg->tau, cfg->tau + d->ncell);
  if (cfg->solid) h->solid_host.assign(cfg->solid, cfg->solid + d->ncell);
  if (cfg->gamma) h->gamma_host.assign(cfg->gamma, cfg->gamma + d->ncell);
  const dim3 g = adj_grid(d);
  h->npartial = (int64_t)g.x * g.y * g.z;
  lbm_status st = LBM_OK;
  for (int i = 0; i < 2 && !st; i++) st = dmalloc(d, &h->lam[i], h->lam_bytes, "adjoint state");
  if (!st && cfg->gamma)
    st = dmalloc(d, (void**)&h->grad[0], sizeof(double) * d->ncell, "gamma gradient");
  if (!st && cfg->tau)
    st = dmalloc(d, (void**)&h->grad[1], sizeof(double) * d->ncell, "tau gradient");
  if (!st) st = dmalloc(d, &h->f0, h->state_bytes, "initial state");
  if (!st) st = dmalloc(d, (void**)&h->partial, sizeof(double) * (h->npartial + 1), "reduction");
  if (!st && hipMemcpy(h->f0, d->f[0], h->state_bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
    set_error("lbm_adjoint_create: initial state copy failed");
    st = LBM_ERR_HIP;
  }
  if (st) {
    free_adjoint(h);
    return st;
  }
  *out = h;
  return LBM_OK;
}

lbm_status lbm_adjoint_destroy(lbm_adjoint* h) {
  if (!h) return LBM_OK;
  (void)hipStreamSynchronize(h->d->stream);
  free_adjoint(h);
  return LBM_OK;
}

lbm_status lbm_adjoint_set_initial_state(lbm_adjoint* h, const double* f) {
  LBM_REQUIRE(h && f, LBM_ERR_ARG, "lbm_adjoint_set_initial_state: null pointer");
  if (lbm_status st = lbm_dense_set_state(h->d, f)) return st;
  LBM_HIP_CHECK(hipMemcpy(h->f0, h->d->f[0], h->state_bytes, hipMemcpyDeviceToDevice));
  h->T = -1;
  return LBM_OK;
}

lbm_status lbm_adjoint_forward(lbm_adjoint* h, int32_t nsteps, double* objective) {
  LBM_REQUIRE(h && objective && nsteps >= 1, LBM_ERR_ARG, "lbm_adjoint_forward: bad arguments");
  if (lbm_status st = ensure_buffers(h, nsteps)) return st;
  void* fin = nullptr;
  if (lbm_status st = run_forward(h, nsteps, true, &fin)) return st;
  h->final_state = fin;
  h->T = nsteps;
  lbm_dense* d = h->d;
  launch_qs<seed_u>(d, adj_grid(d), adj_block(d), (const void*)fin, h->lam[0], d->args,
                    h->objective, h->outlet_x, h->partial);
  LBM_LAUNCH_CHECK();
  h->lam_cur = 0;
  if (lbm_status st = reduce_to_host(h, &h->J)) return st;
  *objective = h->J;
  return LBM_OK;
}

lbm_status lbm_adjoint_backward(lbm_adjoint* h, double* grad_gamma, double* grad_tau,
                                double* elapsed_ms) {
  LBM_REQUIRE(h, LBM_ERR_ARG, "lbm_adjoint_backward: null handle");
  LBM_REQUIRE(h->T >= 1, LBM_ERR_STATE, "lbm_adjoint_backward: run lbm_adjoint_forward first");
  LBM_REQUIRE(!grad_gamma || h->grad[0], LBM_ERR_STATE,
              "lbm_adjoint_backward: no gamma field was given, so no gamma gradient");
  LBM_REQUIRE(!grad_tau || h->grad[1], LBM_ERR_STATE,
              "lbm_adjoint_backward: no tau field was given, so no tau gradient");
  lbm_dense* d = h->d;
  // The seed (lambda^T) is only valid once: re-seed when a backward already ran.
  launch_qs<seed_u>(d, adj_grid(d), adj_block(d), (const void*)h->final_state, h->lam[0],
                    d->args, h->objective, h->outlet_x, h->partial);
  LBM_LAUNCH_CHECK();
  int cur = 0;
  LBM_HIP_CHECK(hipEventRecord(d->e0, d->stream));
  for (int i = 0; i < 2; i++)
    if (h->grad[i]) LBM_HIP_CHECK(hipMemsetAsync(h->grad[i], 0, sizeof(double) * d->ncell, d->stream));
  const int64_t T = h->T;
  const int64_t nseg = (T + h->every - 1) / h->every;
  for (int64_t i = nseg - 1; i >= 0; i--) {
    const int64_t t0 = i * h->every;
    const int64_t len = std::min<int64_t>(h->every, T - t0);
    if (lbm_status st = backward_range(h, h->ckpt[i], len, 0, cur)) return st;
  }
  LBM_HIP_CHECK(hipEventRecord(d->e1, d->stream));
  LBM_HIP_CHECK(hipEventSynchronize(d->e1));
  float ms = 0.f;
  LBM_HIP_CHECK(hipEventElapsedTime(&ms, d->e0, d->e1));
  if (elapsed_ms) *elapsed_ms = ms;
  h->lam_cur = cur;
  if (grad_gamma)
    LBM_HIP_CHECK(hipMemcpy(grad_gamma, h->grad[0], sizeof(double) * d->ncell, hipMemcpyDeviceToHost));
  if (grad_tau)
    LBM_HIP_CHECK(hipMemcpy(grad_tau, h->grad[1], sizeof(double) * d->ncell, hipMemcpyDeviceToHost));
  return LBM_OK;
}

lbm_status lbm_adjoint_get_lambda0(lbm_adjoint* h, double* lam) {
  LBM_REQUIRE(h && lam, LBM_ERR_ARG, "lbm_adjoint_get_lambda0: null pointer");
  lbm_dense* d = h->d;
  const int64_t nq = d->ncell * d->q;
  if (d->storage == LBM_STORE_F64) {
    LBM_HIP_CHECK(hipMemcpy(lam, h->lam[h->lam_cur], sizeof(double) * nq, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(nq);
    LBM_HIP_CHECK(hipMemcpy(tmp.data(), h->lam[h->lam_cur], sizeof(float) * nq, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nq; i++) lam[i] = tmp[i];
  }
  return LBM_OK;
}

lbm_status lbm_adjoint_get_final_state(lbm_adjoint* h, double* f) {
  LBM_REQUIRE(h && f, LBM_ERR_ARG, "lbm_adjoint_get_final_state: null pointer");
  LBM_REQUIRE(h->final_state, LBM_ERR_STATE, "lbm_adjoint_get_final_state: no forward pass yet");
  lbm_dense* d = h->d;
  // f^T lives in one of the engine's two arrays: make it the current one (push state)
  d->cur = h->final_state == d->f[0] ? 0 : 1;
  if (h->final_state != d->f[0] && h->final_state != d->f[1]) {
    LBM_HIP_CHECK(hipMemcpy(d->f[0], h->final_state, h->state_bytes, hipMemcpyDeviceToDevice));
    d->cur = 0;
  }
  d->t = 0;
  return lbm_dense_get_state(d, f);
}

}  // extern "C"

namespace {

// Central difference of order 2, 4 or 6 of J along `set(e)` (which puts the
// parameter field at p0 + e v on the device): D_k = J(+kh) - J(-kh), each summed
// node by node, then the weights of the order; `set(0)` restores the field.
template <typename Set>
lbm_status central_difference(lbm_adjoint* h, Set set, double step, int32_t nsteps, int32_t order,
                              double* derivative) {
  lbm_dense* d = h->d;
  for (int i = 0; i < 2; i++)
    if (!h->fd[i])
      if (lbm_status st = dmalloc(d, &h->fd[i], h->state_bytes, "finite-difference state")) return st;
  double dj[3] = {0.0, 0.0, 0.0};
  const int npairs = order / 2;
  for (int k = 0; k < npairs; k++) {
    const double delta = (k + 1) * step;
    for (int sgn = 0; sgn < 2; sgn++) {
      if (lbm_status st = set(sgn == 0 ? delta : -delta)) {
        (void)set(0.0);
        return st;
      }
      void* fin = nullptr;
      if (lbm_status st = run_forward(h, nsteps, false, &fin)) return st;
      LBM_HIP_CHECK(hipMemcpyAsync(h->fd[sgn], fin, h->state_bytes, hipMemcpyDeviceToDevice,
                                   d->stream));
    }
    launch_qs<objdiff_u>(d, adj_grid(d), adj_block(d), (const void*)h->fd[0],
                         (const void*)h->fd[1], d->args, h->objective, h->outlet_x, h->partial);
    LBM_LAUNCH_CHECK();
    if (lbm_status st = reduce_to_host(h, &dj[k])) return st;
  }
  if (lbm_status st = set(0.0)) return st;
  h->final_state = nullptr;  // the engine's arrays were reused
  h->T = -1;
  // central-difference weights of orders 2, 4 and 6
  if (order == 2) *derivative = dj[0] / (2.0 * step);
  else if (order == 4) *derivative = (8.0 * dj[0] - dj[1]) / (12.0 * step);
  else *derivative = (45.0 * dj[0] - 9.0 * dj[1] + dj[2]) / (60.0 * step);
  return LBM_OK;
}

}  // namespace

extern "C" {

lbm_status lbm_adjoint_fd(lbm_adjoint* h, int32_t param, int64_t cell, double step,
                          int32_t nsteps, int32_t order, double* derivative) {
  LBM_REQUIRE(h && derivative && nsteps >= 1 && step > 0.0, LBM_ERR_ARG,
              "lbm_adjoint_fd: bad arguments");
  LBM_REQUIRE(order == 2 || order == 4 || order == 6, LBM_ERR_ARG,
              "lbm_adjoint_fd: order must be 2, 4 or 6");
  LBM_REQUIRE(param == 0 || param == 1, LBM_ERR_ARG, "lbm_adjoint_fd: param 0 = gamma, 1 = tau");
  lbm_dense* d = h->d;
  LBM_REQUIRE(cell >= 0 && cell < d->ncell, LBM_ERR_ARG, "lbm_adjoint_fd: cell out of range");
  LBM_REQUIRE(param == 0 ? !h->gamma_host.empty() : !h->tau_host.empty(), LBM_ERR_STATE,
              "lbm_adjoint_fd: the handle has no %s field", param == 0 ? "gamma" : "tau");
  const double p0 = param == 0 ? h->gamma_host[cell] : h->tau_host[cell];
  auto set = [&](double e) { return set_param(h, param, cell, p0 + e); };
  return central_difference(h, set, step, nsteps, order, derivative);
}

lbm_status lbm_adjoint_fd_direction(lbm_adjoint* h, int32_t param, const double* direction,
                                    double step, int32_t nsteps, int32_t order,
                                    double* derivative) {
  LBM_REQUIRE(h && direction && derivative && nsteps >= 1 && step > 0.0, LBM_ERR_ARG,
              "lbm_adjoint_fd_direction: bad arguments");
  LBM_REQUIRE(order == 2 || order == 4 || order == 6, LBM_ERR_ARG,
              "lbm_adjoint_fd_direction: order must be 2, 4 or 6");
  LBM_REQUIRE(param == 0 || param == 1, LBM_ERR_ARG,
              "lbm_adjoint_fd_direction: param 0 = gamma, 1 = tau");
  LBM_REQUIRE(param == 0 ? !h->gamma_host.empty() : !h->tau_host.empty(), LBM_ERR_STATE,
              "lbm_adjoint_fd_direction: the handle has no %s field", param == 0 ? "gamma" : "tau");
  auto set = [&](double e) { return set_field(h, param, direction, e); };
  return central_difference(h, set, step, nsteps, order, derivative);
}

lbm_status lbm_adjoint_device_bytes(lbm_adjoint* h, size_t* bytes) {
  LBM_REQUIRE(h && bytes, LBM_ERR_ARG, "lbm_adjoint_device_bytes: null pointer");
  *bytes = h->d->device_bytes;
  return LBM_OK;
}

}  // extern "C"
