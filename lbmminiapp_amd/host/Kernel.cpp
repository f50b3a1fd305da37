// Kernel.cpp -- KernelAA_HIP / KernelAB_HIP over the C ABI (lbm_sparse_*).
#include "Kernel.hh"

#include <stdexcept>
#include <string>

#include "lbm_capi.h"

namespace {
void check(lbm_status st, const char* what) {
  if (st != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

lbm_sparse* make_sparse(Lattice& lattice, Geometry& geometry, int mode) {
  Parameters& p = geometry.getParameters();
  lbm_sparse_config cfg{};
  cfg.n_fluid = geometry.getNumFluidPts();
  cfg.q = lattice.size();
  cfg.layout = p.getLayout() == 0 ? LBM_LAYOUT_SOA : LBM_LAYOUT_AOS;
  cfg.storage = sizeof(Pdf) == 8 ? LBM_STORE_F64 : LBM_STORE_F32;
  cfg.adjacency_mode = mode;
  cfg.omega = p.getOmega();
  for (int d = 0; d < 3; d++) cfg.force[d] = p.getForce()[d];
  cfg.adjacency = geometry.adjacency_;
  lbm_sparse* h = nullptr;
  check(lbm_sparse_create(&cfg, &h), "lbm_sparse_create");
  return h;
}
}  // namespace

KernelAA_HIP::~KernelAA_HIP(void) { lbm_sparse_destroy(h_); }

void KernelAA_HIP::setup() {
  if (geometry_.getAdjacencyMode() != 1)
    throw std::runtime_error("KernelAA_HIP: the geometry's adjacency is not the AA (x - c_s) one");
  h_ = make_sparse(lattice_, geometry_, 1);
}

void KernelAA_HIP::timestepEvenForce(Pdf* dstrb, int startIdx, int countIdx) {
  check(lbm_sparse_aa_even(h_, dstrb, startIdx, countIdx), "timestepEvenForce");
}

void KernelAA_HIP::timestepOddForce(Pdf* dstrb, int startIdx, int countIdx) {
  check(lbm_sparse_aa_odd(h_, dstrb, startIdx, countIdx), "timestepOddForce");
}

void KernelAA_HIP::initRest(Pdf* dstrb) { check(lbm_sparse_init_rest(h_, dstrb, 1), "initRest"); }

void KernelAA_HIP::synchronize() { check(lbm_sparse_synchronize(h_), "synchronize"); }

KernelAB_HIP::~KernelAB_HIP(void) { lbm_sparse_destroy(h_); }

void KernelAB_HIP::setup() {
  pull_ = geometry_.getParameters().getPullSetting() ? 1 : 0;
  if (geometry_.getAdjacencyMode() != pull_)
    throw std::runtime_error("KernelAB_HIP: the geometry's adjacency does not match `pull`");
  h_ = make_sparse(lattice_, geometry_, pull_);
}

void KernelAB_HIP::timestepForce(Pdf* src, Pdf* tgt, int startIdx, int countIdx) {
  check(lbm_sparse_ab(h_, src, tgt, startIdx, countIdx, pull_), "timestepForce");
}

void KernelAB_HIP::initRest(Pdf* dstrb) { check(lbm_sparse_init_rest(h_, dstrb, 0), "initRest"); }

void KernelAB_HIP::synchronize() { check(lbm_sparse_synchronize(h_), "synchronize"); }
