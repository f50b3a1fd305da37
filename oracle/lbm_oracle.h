/*
 * lbm_oracle.h -- CPU restatement of the lbmminiapp collide-stream path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it; the product path
 * (lbmminiapp_amd/) never does.
 *
 * Every function restates reference arithmetic with the reference's operation
 * order, so that IEEE results are reproducible bit for bit:
 *   - velocity sets ........ src/lattices/LatticeD3Q19.cpp:5-88, LatticeD3Q27.cpp:5-120
 *   - BGK + Guo collision .. src/kernels/KernelAA_SOA_Base.cpp:185-247
 *   - AA even / odd ........ src/kernels/KernelAA_SOA_Base.cpp:173-254 / :257-342
 *                            (AOS: src/kernels/KernelAA_AOS_Base.cpp:174-342)
 *   - AB push .............. src/kernels/KernelAB_SOA_Base.cpp:111-186 (AOS :108-182)
 *   - AB pull .............. src/kernels/KernelAB_SOA_Pull.cpp:113-194
 *   - unrolled Q19 variants  src/kernels/KernelAB_SOA_Unroll.cpp:223-414,
 *                            KernelAB_SOA_Unroll_Pull.cpp:222-414,
 *                            KernelAA_SOA_Unroll.cpp:415-799
 *   - adjacency / bounce-back src/geometries/GeometryAAAB.cpp:18-124
 *   - rest initialisation .. src/propagation/PropagationAA.cpp:60-123,
 *                            PropagationAB.cpp:91-152
 *
 * Extensions that the reference does not have (runtime omega and force,
 * per-node omega, Brinkman velocity scaling gamma, periodic wrap in every
 * dimension, the discrete adjoint) reduce to the reference arithmetic when
 * they are switched off (NULL field pointers, force = (1e-6, 0, 0)).
 *
 * Index conventions
 *   cell   = x + nx * (y + ny * z)                   (dense grid, x fastest)
 *   SOA    = nfluid * s + i                          (Makefile:3 -DSOA)
 *   AOS    = q * i + s                               (-DAOS)
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SOA 0
#define ORC_AOS 1

/* adjacency modes: 0 = neighbour at x + c_s (AB push), 1 = neighbour at x - c_s (AA, AB pull) */
#define ORC_ADJ_PUSH 0
#define ORC_ADJ_PULL 1

int orc_lattice(int q, int32_t* dirs, int32_t* opp, double* w);

int orc_build_adjacency(int q, int layout, int mode, int nx, int ny, int nz,
                        const int32_t* periodic, const int32_t* fluid_of_cell,
                        const int32_t* cell_of_fluid, int64_t nfluid, int32_t* adj);

/* ---- Base kernels (double and float storage) ---- */
#define ORC_DECL(T, SUF)                                                                      \
  void orc_aa_even_##SUF(int q, int layout, int64_t nfluid, T* f, int64_t start,             \
                         int64_t count, double omega, const double* force,                    \
                         const double* omega_node, const double* gamma_node);                 \
  void orc_aa_odd_##SUF(int q, int layout, int64_t nfluid, T* f, const int32_t* adj,         \
                        int64_t start, int64_t count, double omega, const double* force,      \
                        const double* omega_node, const double* gamma_node);                  \
  void orc_ab_push_##SUF(int q, int layout, int64_t nfluid, const T* src, T* tgt,            \
                         const int32_t* adj, int64_t start, int64_t count, double omega,      \
                         const double* force, const double* omega_node,                       \
                         const double* gamma_node);                                           \
  void orc_ab_pull_##SUF(int q, int layout, int64_t nfluid, const T* src, T* tgt,            \
                         const int32_t* adj, int64_t start, int64_t count, double omega,      \
                         const double* force, const double* omega_node,                       \
                         const double* gamma_node);                                           \
  void orc_init_rest_aa_##SUF(int q, int layout, int64_t nfluid, T* f);                      \
  void orc_init_rest_ab_##SUF(int q, int layout, int64_t nfluid, T* f);                      \
  void orc_ab_push_unroll19_##SUF(int64_t nfluid, const T* src, T* tgt, const int32_t* adj,  \
                                  int64_t start, int64_t count, double omega);                \
  void orc_ab_pull_unroll19_##SUF(int64_t nfluid, const T* src, T* tgt, const int32_t* adj,  \
                                  int64_t start, int64_t count, double omega);                \
  void orc_aa_even_unroll19_##SUF(int64_t nfluid, T* f, int64_t start, int64_t count,        \
                                  double omega);                                              \
  void orc_aa_odd_unroll19_##SUF(int64_t nfluid, T* f, const int32_t* adj, int64_t start,    \
                                 int64_t count, double omega);

ORC_DECL(double, f64)
ORC_DECL(float, f32)
#undef ORC_DECL

/* equilibrium f^eq_s(rho, u), written with the collision's operation order */
void orc_equilibrium(int q, double rho, const double* u, double* feq);

/* dense helpers (double): fluid-index order given by cell_of_fluid */
void orc_moments(int q, int64_t nfluid, const double* f_canon, double* rho, double* u);
double orc_mass(int64_t n, const double* f);

/* ---- discrete adjoint of the canonical push map f^{n+1} = S(C(f^n)) (double) ----
 * objective 0: J = sum_x 0.5 |m|^2 / rho   (kinetic energy of the final state)
 * objective 1: J = sum_{x: x_coord == outlet_x} m_x                (outlet flux)
 * Gradients are written per fluid node for gamma (Brinkman) and tau (= 1/omega);
 * lam0 (nullable, q * nfluid, canonical order) receives lambda^0 = dJ/df^0.
 */
double orc_objective(int q, int objective, int64_t nfluid, const double* f_canon,
                     const int32_t* cell_of_fluid, int nx, int outlet_x);
int orc_adjoint(int q, int64_t nfluid, const int32_t* adj_push, const double* f0_canon,
                int nsteps, double omega, const double* force, const double* tau_node,
                const double* gamma_node, int objective, const int32_t* cell_of_fluid,
                int nx, int outlet_x, double* f_final, double* j_out, double* grad_gamma,
                double* grad_tau, double* lam0);

#ifdef __cplusplus
}
#endif
#endif
