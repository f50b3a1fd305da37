/*
 * lbm_oracle.c -- CPU restatement of the lbmminiapp collide-stream path.
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never the product.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp; no FMA, no
 * fast-math, so every operation is one IEEE-rounded double or float op as in
 * the reference's plain g++ build).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Velocity sets in the reference's direction order.
 * D3Q19: src/lattices/LatticeD3Q19.cpp:12-86; D3Q27: src/lattices/LatticeD3Q27.cpp:12-118. */
static const int DIRS27[27][3] = {
    {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},  {1, 1, 0},
    {1, -1, 0},  {1, 0, 1},   {1, 0, -1},  {-1, 1, 0},  {-1, -1, 0}, {-1, 0, 1},  {-1, 0, -1},
    {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},   {1, 1, -1},  {1, -1, 1},
    {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}, {0, 0, 0}};

int orc_lattice(int q, int32_t* dirs, int32_t* opp, double* w) {
  if (q != 19 && q != 27) return -1;
  for (int s = 0; s < q; s++) {
    const int* d;
    if (q == 19 && s == 18)
      d = DIRS27[26];
    else
      d = DIRS27[s];
    for (int k = 0; k < 3; k++) dirs[3 * s + k] = d[k];
    int nz = (d[0] != 0) + (d[1] != 0) + (d[2] != 0);
    if (q == 19)
      w[s] = nz == 0 ? 1.0 / 3.0 : (nz == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
    else
      w[s] = nz == 0 ? 8.0 / 27.0 : (nz == 1 ? 2.0 / 27.0 : (nz == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
  }
  for (int s = 0; s < q; s++) {
    opp[s] = -1;
    for (int r = 0; r < q; r++)
      if (dirs[3 * r] == -dirs[3 * s] && dirs[3 * r + 1] == -dirs[3 * s + 1] &&
          dirs[3 * r + 2] == -dirs[3 * s + 2])
        opp[s] = r;
  }
  return 0;
}

/* Adjacency with halfway bounce-back, src/geometries/GeometryAAAB.cpp:18-124:
 * neighbour x + c_s (push) or x - c_s (AA / pull); a neighbour outside the box
 * or solid maps to the node's own opposite slot.  The reference wraps x only
 * (GeometryAAAB.cpp:56-66); here each dimension wraps when periodic[d] != 0. */
int orc_build_adjacency(int q, int layout, int mode, int nx, int ny, int nz,
                        const int32_t* periodic, const int32_t* fluid_of_cell,
                        const int32_t* cell_of_fluid, int64_t n, int32_t* adj) {
  int32_t dirs[81], opp[27];
  double w[27];
  if (orc_lattice(q, dirs, opp, w)) return -1;
  const int dims[3] = {nx, ny, nz};
  for (int64_t i = 0; i < n; i++) {
    int64_t cell = cell_of_fluid[i];
    int p[3] = {(int)(cell % nx), (int)((cell / nx) % ny), (int)(cell / ((int64_t)nx * ny))};
    for (int s = 0; s < q; s++) {
      int sd = mode == ORC_ADJ_PUSH ? s : opp[s];
      int nb[3];
      int inside = 1;
      for (int k = 0; k < 3; k++) {
        nb[k] = p[k] + dirs[3 * sd + k];
        if (periodic[k]) {
          if (nb[k] < 0) nb[k] += dims[k];
          else if (nb[k] >= dims[k]) nb[k] -= dims[k];
        }
        if (nb[k] < 0 || nb[k] >= dims[k]) inside = 0;
      }
      int64_t nbf = -1;
      if (inside) nbf = fluid_of_cell[nb[0] + (int64_t)nx * (nb[1] + (int64_t)ny * nb[2])];
      int64_t a;
      if (nbf < 0)
        a = layout == ORC_SOA ? n * opp[s] + i : (int64_t)q * i + opp[s];
      else
        a = layout == ORC_SOA ? n * s + nbf : (int64_t)q * nbf + s;
      adj[layout == ORC_SOA ? n * s + i : (int64_t)q * i + s] = (int32_t)a;
    }
  }
  return 0;
}

#define T double
#define SUF f64
#include "oracle_kernels.inc"
#undef T
#undef SUF

#define T float
#define SUF f32
#include "oracle_kernels.inc"
#undef T
#undef SUF

/* f^eq with the collision's association (KernelAA_SOA_Base.cpp:227-238 with F = 0). */
void orc_equilibrium(int q, double rho, const double* u, double* feq) {
  int32_t dirs[81], opp[27];
  double w[27];
  orc_lattice(q, dirs, opp, w);
  double vdotv = 3.0 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int s = 0; s < q; s++) {
    double c0 = dirs[3 * s], c1 = dirs[3 * s + 1], c2 = dirs[3 * s + 2];
    double cdotv = 3.0 * (c0 * u[0] + c1 * u[1] + c2 * u[2]);
    feq[s] = w[s] * rho * (1.0 + cdotv + 0.5 * cdotv * cdotv - 0.5 * vdotv);
  }
}

/* rho and u = m / rho from canonical SOA populations (f_s(i) at n*s + i), as
 * PropagationAB::getRho / getVelocity (src/propagation/PropagationAB.cpp:222-281). */
void orc_moments(int q, int64_t n, const double* f, double* rho, double* u) {
  int32_t dirs[81], opp[27];
  double w[27];
  orc_lattice(q, dirs, opp, w);
  for (int64_t i = 0; i < n; i++) {
    double r = 0.0, v[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < q; s++) {
      double x = f[n * s + i];
      r += x;
      for (int d = 0; d < 3; d++) v[d] += x * dirs[3 * s + d];
    }
    double inv = 1.0 / r;
    rho[i] = r;
    for (int d = 0; d < 3; d++) u[3 * i + d] = v[d] * inv;
  }
}

double orc_mass(int64_t n, const double* f) {
  double m = 0.0;
  for (int64_t i = 0; i < n; i++) m += f[i];
  return m;
}

/* ---------------- discrete adjoint (double, canonical push map) ----------------
 * Forward map of one step: g = C(f) per node (collide_base_f64 with per-node
 * omega = 1/tau and Brinkman scaling u = (1 - gamma) u*), then push streaming
 * g_s(i) -> adj_push[n*s + i].  The adjoint is the exact transpose: gather
 * lambda_g_s(i) = lambda^{t+1}[adj_push[n*s+i]], then
 *   lambda^t_k = (1-w) lambda_g_k + a + B . ((1-gamma) c_k - u) / rho
 * with a = w sum_s lambda_g_s eq_s / rho and
 *   B = sum_s lambda_g_s (w d eq_s/du + (1 - w/2) d F_s/du).
 * Parameter sensitivities: dg_s/dw = eq_s - f_s - F_s/2 (dtau = -w^2 dw),
 * dg/dgamma = -B . u*.  No reference counterpart: pinned by finite
 * differences of the forward oracle in tests/. */
static void adj_node(int q, const int32_t* dirs, const double* w, const double* f_i,
                     const double* lam_g, double omega, const double* F, double gamma_i,
                     double* lam_out, double* dgamma, double* domega) {
  double rho = 0.0, m[3] = {0, 0, 0};
  for (int s = 0; s < q; s++) {
    rho += f_i[s];
    for (int d = 0; d < 3; d++) m[d] += f_i[s] * dirs[3 * s + d];
  }
  double us[3], u[3];
  for (int d = 0; d < 3; d++) {
    us[d] = (m[d] + 0.5 * F[d]) / rho;
    u[d] = (1.0 - gamma_i) * us[d];
  }
  double uu = 3.0 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double fu = 3.0 * (F[0] * u[0] + F[1] * u[1] + F[2] * u[2]);
  double a = 0.0, B[3] = {0, 0, 0}, dw = 0.0;
  for (int s = 0; s < q; s++) {
    double c[3] = {dirs[3 * s], dirs[3 * s + 1], dirs[3 * s + 2]};
    double cu = 3.0 * (c[0] * u[0] + c[1] * u[1] + c[2] * u[2]);
    double cf = 3.0 * (c[0] * F[0] + c[1] * F[1] + c[2] * F[2]);
    double eq = w[s] * rho * (1.0 + cu + 0.5 * cu * cu - 0.5 * uu);
    double fd = w[s] * (cf - fu + cu * cf);
    a += lam_g[s] * eq / rho;
    dw += lam_g[s] * (eq - f_i[s] - 0.5 * fd);
    for (int d = 0; d < 3; d++) {
      double deq = w[s] * rho * (3.0 * c[d] + 3.0 * cu * c[d] - 3.0 * u[d]);
      double dfd = w[s] * (-3.0 * F[d] + 3.0 * cf * c[d]);
      B[d] += lam_g[s] * (omega * deq + (1.0 - 0.5 * omega) * dfd);
    }
  }
  a *= omega;
  for (int k = 0; k < q; k++) {
    double acc = (1.0 - omega) * lam_g[k] + a;
    for (int d = 0; d < 3; d++) acc += B[d] * ((1.0 - gamma_i) * dirs[3 * k + d] - u[d]) / rho;
    lam_out[k] = acc;
  }
  *dgamma = -(B[0] * us[0] + B[1] * us[1] + B[2] * us[2]);
  *domega = dw;
}

double orc_objective(int q, int objective, int64_t n, const double* f,
                     const int32_t* cell_of_fluid, int nx, int outlet_x) {
  int32_t dirs[81], opp[27];
  double w[27];
  orc_lattice(q, dirs, opp, w);
  double J = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double rho = 0.0, m[3] = {0, 0, 0};
    for (int s = 0; s < q; s++) {
      rho += f[n * s + i];
      for (int d = 0; d < 3; d++) m[d] += f[n * s + i] * dirs[3 * s + d];
    }
    if (objective == 0)
      J += 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rho;
    else if ((int)(cell_of_fluid[i] % nx) == outlet_x)
      J += m[0];
  }
  return J;
}

int orc_adjoint(int q, int64_t n, const int32_t* adj, const double* f0, int nsteps,
                double omega, const double* F, const double* tau_node, const double* gamma_node,
                int objective, const int32_t* cell_of_fluid, int nx, int outlet_x,
                double* f_final, double* j_out, double* grad_gamma, double* grad_tau,
                double* lam0) {
  int32_t dirs[81], opp[27];
  double w[27];
  if (orc_lattice(q, dirs, opp, w)) return -1;
  const int64_t nq = n * (int64_t)q;
  double* states = (double*)malloc(sizeof(double) * nq * (size_t)(nsteps + 1));
  double* omega_node = (double*)malloc(sizeof(double) * n);
  double* lam = (double*)malloc(sizeof(double) * nq);
  double* lam2 = (double*)malloc(sizeof(double) * nq);
  if (!states || !omega_node || !lam || !lam2) return -2;
  for (int64_t i = 0; i < n; i++) omega_node[i] = tau_node ? 1.0 / tau_node[i] : omega;
  memcpy(states, f0, sizeof(double) * nq);
  for (int t = 0; t < nsteps; t++)
    orc_ab_push_f64(q, ORC_SOA, n, states + nq * t, states + nq * (t + 1), adj, 0, n, omega, F,
                    omega_node, gamma_node);
  const double* fT = states + nq * nsteps;
  if (f_final) memcpy(f_final, fT, sizeof(double) * nq);
  *j_out = orc_objective(q, objective, n, fT, cell_of_fluid, nx, outlet_x);
  /* dJ/df at the final state */
  for (int64_t i = 0; i < n; i++) {
    double rho = 0.0, m[3] = {0, 0, 0};
    for (int s = 0; s < q; s++) {
      rho += fT[n * s + i];
      for (int d = 0; d < 3; d++) m[d] += fT[n * s + i] * dirs[3 * s + d];
    }
    double v[3] = {m[0] / rho, m[1] / rho, m[2] / rho};
    double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    int at_outlet = (int)(cell_of_fluid[i] % nx) == outlet_x;
    for (int s = 0; s < q; s++) {
      if (objective == 0)
        lam[n * s + i] = dirs[3 * s] * v[0] + dirs[3 * s + 1] * v[1] + dirs[3 * s + 2] * v[2] - 0.5 * vv;
      else
        lam[n * s + i] = at_outlet ? (double)dirs[3 * s] : 0.0;
    }
  }
  for (int64_t i = 0; i < n; i++) {
    if (grad_gamma) grad_gamma[i] = 0.0;
    if (grad_tau) grad_tau[i] = 0.0;
  }
  for (int t = nsteps - 1; t >= 0; t--) {
    const double* ft = states + nq * t;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double fi[27], lg[27], lo[27], dg, dwm;
      for (int s = 0; s < q; s++) {
        fi[s] = ft[n * s + i];
        lg[s] = lam[adj[n * s + i]];
      }
      double gi = gamma_node ? gamma_node[i] : 0.0;
      adj_node(q, dirs, w, fi, lg, omega_node[i], F, gi, lo, &dg, &dwm);
      for (int s = 0; s < q; s++) lam2[n * s + i] = lo[s];
      if (grad_gamma) grad_gamma[i] += dg;
      if (grad_tau) grad_tau[i] += -omega_node[i] * omega_node[i] * dwm;
    }
    double* tmp = lam;
    lam = lam2;
    lam2 = tmp;
  }
  if (lam0) memcpy(lam0, lam, sizeof(double) * nq);
  free(states);
  free(omega_node);
  free(lam);
  free(lam2);
  return 0;
}
