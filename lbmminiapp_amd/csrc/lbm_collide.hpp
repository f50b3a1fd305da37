// lbm_collide.hpp -- BGK + Guo-forcing collision for one node, host/device.
//
// collide_ref: the reference arithmetic (src/kernels/KernelAA_SOA_Base.cpp:185-247,
// same formula in every Base kernel), fully unrolled at compile time.  It keeps
// the reference's association and evaluation order and only drops products
// with a zero lattice-velocity component, which are exact no-ops (x + (+-0) = x,
// x * +-1 = +-x), so with FP contraction off the result is the reference's,
// IEEE bit for bit.  T is the population type (the reference's `Pdf`): for
// T = float the double literals promote intermediates exactly as the
// reference's -D_PRECISION_SINGLE_ build does.
//
// collide_shifted: storage as deviations d_s = f_s - w_s (FP32 or scaled FP16),
// moment sums in FP64, collision arithmetic in FP32:
//   rho = 1 + sum d_s,  m = sum c_s d_s,  u = (m + F/2) / rho
//   d*_s = (1-w) d_s + w w_s [(rho-1) + rho (cu + cu^2/2 - uu/2)] + (1-w/2) w_s (cf - fu + cu cf)
// which is the same update as collide_ref with f = w + d, rewritten so that
// the O(1) weight never enters the FP32 arithmetic.
//
// Extensions (not in the reference): per-node omega and Brinkman velocity
// scaling u <- (1 - gamma) u ("gray" node, gamma = 0 fluid, gamma = 1 rest);
// with gamma disabled and a uniform omega both functions reduce to the above.
#pragma once
#include "lbm_lattice.hpp"

namespace lbm {

template <int Q, int S>
struct Dir {
  static constexpr int cx = Lattice<Q>::c(S, 0);
  static constexpr int cy = Lattice<Q>::c(S, 1);
  static constexpr int cz = Lattice<Q>::c(S, 2);
};

// sum of the non-zero terms c_d * a_d in d order (d = 0, 1, 2); 0 if none.
template <typename T, int C0, int C1, int C2>
__host__ __device__ inline T csum(T a0, T a1, T a2) {
  if constexpr (C0 == 0 && C1 == 0 && C2 == 0) {
    return T(0);
  } else {
    T acc{};
    bool first = true;
    if constexpr (C0 != 0) {
      acc = C0 > 0 ? a0 : -a0;
      first = false;
    }
    if constexpr (C1 != 0) {
      if (first) acc = C1 > 0 ? a1 : -a1;
      else acc = C1 > 0 ? acc + a1 : acc - a1;
      first = false;
    }
    if constexpr (C2 != 0) {
      if (first) acc = C2 > 0 ? a2 : -a2;
      else acc = C2 > 0 ? acc + a2 : acc - a2;
    }
    return acc;
  }
}

template <int Q, typename T, int S>
__host__ __device__ inline void moments_acc(const T* f, T& rho, T& v0, T& v1, T& v2) {
  rho += f[S];
  using D = Dir<Q, S>;
  if constexpr (D::cx > 0) v0 += f[S];
  if constexpr (D::cx < 0) v0 -= f[S];
  if constexpr (D::cy > 0) v1 += f[S];
  if constexpr (D::cy < 0) v1 -= f[S];
  if constexpr (D::cz > 0) v2 += f[S];
  if constexpr (D::cz < 0) v2 -= f[S];
  if constexpr (S + 1 < Q) moments_acc<Q, T, S + 1>(f, rho, v0, v1, v2);
}

template <int Q, typename T, int S, bool FORCE>
__host__ __device__ inline void relax_ref(const T* f, T* out, T rho, T v0, T v1, T v2, T vdotv,
                                          T fdotv, const T* F, double omega) {
#pragma clang fp contract(off)
  using D = Dir<Q, S>;
  const T w = T(Lattice<Q>::w(S));
  if constexpr (!FORCE) {
    // F = 0: every Guo term is an exact +0 (cdotf = fdotv = 0), so dropping
    // them leaves the reference result unchanged bit for bit.
    if constexpr (D::cx == 0 && D::cy == 0 && D::cz == 0) {
      T eq = w * rho * (1.0 - 0.5 * vdotv);
      out[S] = (1.0 - omega) * f[S] + omega * eq;
    } else {
      T cdotv = 3.0 * csum<T, D::cx, D::cy, D::cz>(v0, v1, v2);
      T eq = w * rho * (1.0 + cdotv + 0.5 * cdotv * cdotv - 0.5 * vdotv);
      out[S] = (1.0 - omega) * f[S] + omega * eq;
    }
  } else if constexpr (D::cx == 0 && D::cy == 0 && D::cz == 0) {
    // rest direction: cdotv = cdotf = 0, (1 + 0) + 0 - x == 1 - x exactly
    T eq = w * rho * (1.0 - 0.5 * vdotv);
    T fdist = w * (T(0) - fdotv);
    out[S] = (1.0 - omega) * f[S] + omega * eq + (1.0 - 0.5 * omega) * fdist;
  } else {
    T cdotv = 3.0 * csum<T, D::cx, D::cy, D::cz>(v0, v1, v2);
    T cdotf = 3.0 * csum<T, D::cx, D::cy, D::cz>(F[0], F[1], F[2]);
    T eq = w * rho * (1.0 + cdotv + 0.5 * cdotv * cdotv - 0.5 * vdotv);
    T fdist = w * (cdotf - fdotv + cdotv * cdotf);
    out[S] = (1.0 - omega) * f[S] + omega * eq + (1.0 - 0.5 * omega) * fdist;
  }
  if constexpr (S + 1 < Q)
    relax_ref<Q, T, S + 1, FORCE>(f, out, rho, v0, v1, v2, vdotv, fdotv, F, omega);
}

// In-place or out-of-place: out may alias f (each f[S] is read before out[S] is written).
// FORCE = false requires F = (0, 0, 0) and then equals FORCE = true bit for bit.
template <int Q, typename T, bool BRINKMAN, bool FORCE = true>
__host__ __device__ inline void collide_ref(const T* f, T* out, double omega, const T* F,
                                            double gamma) {
#pragma clang fp contract(off)
  T rho = 0.0;
  T v0 = 0.0, v1 = 0.0, v2 = 0.0;
  moments_acc<Q, T, 0>(f, rho, v0, v1, v2);
  if constexpr (FORCE) {
    v0 += 0.5 * F[0];
    v1 += 0.5 * F[1];
    v2 += 0.5 * F[2];
  }
  T invRho = 1.0 / rho;
  v0 *= invRho;
  v1 *= invRho;
  v2 *= invRho;
  if constexpr (BRINKMAN) {
    T scale = 1.0 - gamma;
    v0 *= scale;
    v1 *= scale;
    v2 *= scale;
  }
  T vdotv = 3.0 * (v0 * v0 + v1 * v1 + v2 * v2);
  T fdotv = FORCE ? T(3.0 * (F[0] * v0 + F[1] * v1 + F[2] * v2)) : T(0);
  relax_ref<Q, T, 0, FORCE>(f, out, rho, v0, v1, v2, vdotv, fdotv, F, omega);
}

// ---------------------------------------------------------------- shifted storage
template <int Q, int S>
__host__ __device__ inline void moments_shift(const float* d, double& drho, double& m0, double& m1,
                                              double& m2) {
  const double x = (double)d[S];
  drho += x;
  using D = Dir<Q, S>;
  if constexpr (D::cx > 0) m0 += x;
  if constexpr (D::cx < 0) m0 -= x;
  if constexpr (D::cy > 0) m1 += x;
  if constexpr (D::cy < 0) m1 -= x;
  if constexpr (D::cz > 0) m2 += x;
  if constexpr (D::cz < 0) m2 -= x;
  if constexpr (S + 1 < Q) moments_shift<Q, S + 1>(d, drho, m0, m1, m2);
}

template <int Q, int S, bool FORCE>
__host__ __device__ inline void relax_shift(float* d, float drho, float rho, float u0, float u1,
                                            float u2, float uu, float fu, float F0, float F1,
                                            float F2, float om, float omf) {
#pragma clang fp contract(fast)
  using D = Dir<Q, S>;
  constexpr float w = (float)Lattice<Q>::w(S);
  float cu = 0.0f, cf = 0.0f;
  if constexpr (!(D::cx == 0 && D::cy == 0 && D::cz == 0)) {
    cu = 3.0f * csum<float, D::cx, D::cy, D::cz>(u0, u1, u2);
    if constexpr (FORCE) cf = 3.0f * csum<float, D::cx, D::cy, D::cz>(F0, F1, F2);
  }
  const float eqdev = drho + rho * (cu + 0.5f * cu * cu - 0.5f * uu);
  if constexpr (FORCE) {
    const float fd = cf - fu + cu * cf;
    d[S] = (1.0f - om) * d[S] + w * (om * eqdev + omf * fd);
  } else {
    d[S] = (1.0f - om) * d[S] + w * (om * eqdev);
  }
  if constexpr (S + 1 < Q)
    relax_shift<Q, S + 1, FORCE>(d, drho, rho, u0, u1, u2, uu, fu, F0, F1, F2, om, omf);
}

// d[Q]: deviations f_s - w_s in FP32, updated in place.  FORCE = false requires F = 0.
template <int Q, bool BRINKMAN, bool FORCE = true>
__host__ __device__ inline void collide_shifted(float* d, double omega, const double* F,
                                                double gamma) {
  double drho = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
  moments_shift<Q, 0>(d, drho, m0, m1, m2);
  const double rho = 1.0 + drho;
  const double inv = 1.0 / rho;
  double u0 = (FORCE ? m0 + 0.5 * F[0] : m0) * inv;
  double u1 = (FORCE ? m1 + 0.5 * F[1] : m1) * inv;
  double u2 = (FORCE ? m2 + 0.5 * F[2] : m2) * inv;
  if constexpr (BRINKMAN) {
    u0 *= (1.0 - gamma);
    u1 *= (1.0 - gamma);
    u2 *= (1.0 - gamma);
  }
  const float fu0 = (float)u0, fu1 = (float)u1, fu2 = (float)u2;
  const float F0 = (float)F[0], F1 = (float)F[1], F2 = (float)F[2];
  const float uu = 3.0f * (fu0 * fu0 + fu1 * fu1 + fu2 * fu2);
  const float fu = 3.0f * (F0 * fu0 + F1 * fu1 + F2 * fu2);
  const float om = (float)omega;
  const float omf = 1.0f - 0.5f * om;
  relax_shift<Q, 0, FORCE>(d, (float)drho, (float)rho, fu0, fu1, fu2, uu, fu, F0, F1, F2, om,
                           omf);
}

}  // namespace lbm
