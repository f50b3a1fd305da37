// Propagation.cpp -- time loops on one GPU, readout, validation, VTK output.
#include "Propagation.hh"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "lbm_capi.h"

namespace {

void check(lbm_status st, const char* what) {
  if (st != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void write_header(std::ofstream& out, int32_t kind, int32_t q, int32_t layout, int32_t pdf_bytes,
                  int64_t n, int64_t steps) {
  out.write("LBMD", 4);
  const int32_t version = 1;
  out.write((const char*)&version, 4);
  out.write((const char*)&kind, 4);
  out.write((const char*)&q, 4);
  out.write((const char*)&layout, 4);
  out.write((const char*)&pdf_bytes, 4);
  out.write((const char*)&n, 8);
  out.write((const char*)&steps, 8);
}

}  // namespace

Propagation::Propagation(Lattice& lattice, Geometry& geometry, Communication& comm,
                         Parameters& parameters)
    : lattice_(lattice), geometry_(geometry), comm_(comm), parameters_(parameters) {}

Propagation::~Propagation(void) {
  delete[] dstrb_;
  delete[] dstrb2_;
  if (dstrb_d_) lbm_device_free(dstrb_d_);
  if (dstrb2_d_) lbm_device_free(dstrb2_d_);
}

// MFLUPS = fluid nodes * steps / (1e6 * t), MFEUPS = MFLUPS * (Q - 1)
// (src/propagation/Propagation.cpp:200-212).
void Propagation::printMFLUPS(double runtime) {
  const double mflups =
      (geometry_.getGlobalNumFluidPts() * (double)parameters_.getTimesteps()) / (1000000 * runtime);
  const double mfeups = mflups * (lattice_.size() - 1);
  std::cout << "Overall performance" << std::endl;
  std::cout << "Total " << mfeups << " MFEUPS (million fluid elementary updates per second)"
            << std::endl;
  std::cout << "Total " << mflups << " MFLUPS (million fluid lattice updates per second)"
            << std::endl;
  std::cout << "Avg " << mflups / nTasks_ << " MFLUPS per rank" << std::endl;
}

// Centre-line Poiseuille check of the reference (src/propagation/Propagation.cpp:164-197):
// u_max = g D^2 / (16 rho nu), nu = (1/omega - 1/2) / 3, one-sided 25 % tolerance.
void Propagation::checkResult(int currentStep) {
  const int ci = (geometry_.getTotalLengthByDim(0) - 1) / 2;
  const int cj = (geometry_.getTotalLengthByDim(1) - 1) / 2;
  const int ck = (geometry_.getTotalLengthByDim(2) - 1) / 2;
  const int idx = geometry_.getGridIdx(ci, cj, ck);
  if (idx < 0) return;
  const int f = geometry_.getGridElement(idx);
  if (f < 0) return;
  const double diameter = 16.0 * parameters_.getSimulationSize();
  const double nu = (1.0 / 3.0) * (1.0 / parameters_.getOmega() - 0.5);
  const double rho = getRho(f, currentStep);
  const double u = getVelocity(f, currentStep)[0];
  const double expected = (parameters_.getForce()[0] * diameter * diameter) / (16.0 * rho * nu);
  std::cout << "Centre-line velocity " << u << " (Poiseuille " << expected << ")" << std::endl;
  if ((u - expected) / expected < 0.25) std::cout << "Passed baseline validation" << std::endl;
  else std::cout << "Failed baseline validation" << std::endl;
  // The reference's test is one-sided (any u below 1.25 u_P passes, a flow at
  // rest included); the relative deviation states how close the run really is.
  const double dev = std::fabs(u - expected) / expected;
  std::cout << "Two-sided deviation |u - u_P| / u_P = " << dev
            << (dev < 0.25 ? " (within 25%)" : " (outside 25%)") << std::endl;
}

namespace {

// Legacy VTK binary sections are big-endian.
template <typename T>
void put_be(std::ofstream& out, T v) {
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T); i++) out.put((char)b[sizeof(T) - 1 - i]);
}

}  // namespace

// Legacy VTK, one file per process: rho, point type, velocity and coordinates
// over the box (the fields of src/propagation/Propagation.cpp:36-161).
// writevtk 1 writes it as ASCII, as the reference does; writevtk 2 writes the
// same fields in the legacy binary encoding (big-endian), about 4x smaller.
void Propagation::writeVTK(int currentStep) {
  const bool binary = parameters_.getWriteVtk() == 2;
  char name[64];
  std::snprintf(name, sizeof(name), "outs.%05d.vtk", myRank_);
  std::ofstream out(name, binary ? std::ios::binary : std::ios::out);
  const int nx = geometry_.getDimsByDim(0), ny = geometry_.getDimsByDim(1),
            nz = geometry_.getDimsByDim(2);
  const int x0 = geometry_.getMinByDim(0), y0 = geometry_.getMinByDim(1),
            z0 = geometry_.getMinByDim(2);
  out << "# vtk DataFile Version 3.0\nlbmminiapp_amd\n" << (binary ? "BINARY" : "ASCII")
      << "\nDATASET STRUCTURED_POINTS\n";
  out << "DIMENSIONS " << nx << " " << ny << " " << nz << "\n";
  out << "ORIGIN " << x0 << " " << y0 << " " << z0 << "\nSPACING 1 1 1\n";
  out << "POINT_DATA " << (int64_t)nx * ny * nz << "\n";
  auto each = [&](auto fn) {
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) fn(x0 + i, y0 + j, z0 + k);
  };
  auto fluid_of = [&](int i, int j, int k) {
    const int idx = geometry_.getGridIdx(i, j, k);
    return idx >= 0 ? geometry_.getGridElement(idx) : -1;
  };
  auto put_d = [&](double v, const char* sep) {
    if (binary) put_be<double>(out, v);
    else out << v << sep;
  };
  auto put_i = [&](int32_t v, const char* sep) {
    if (binary) put_be<int32_t>(out, v);
    else out << v << sep;
  };
  out << "SCALARS rho double\nLOOKUP_TABLE default\n";
  each([&](int i, int j, int k) {
    const int f = fluid_of(i, j, k);
    put_d(f >= 0 ? getRho(f, currentStep) : 0.0, "\n");
  });
  if (binary) out << "\n";
  out << "SCALARS id int\nLOOKUP_TABLE default\n";
  each([&](int i, int j, int k) { put_i(geometry_.getGridElement(geometry_.getGridIdx(i, j, k)), "\n"); });
  if (binary) out << "\n";
  out << "VECTORS velocity double\n";
  each([&](int i, int j, int k) {
    const int f = fluid_of(i, j, k);
    std::vector<Pdf> v(3, 0.0);
    if (f >= 0) v = getVelocity(f, currentStep);
    put_d(v[0], " ");
    put_d(v[1], " ");
    put_d(v[2], "\n");
  });
  if (binary) out << "\n";
  out << "VECTORS nodes int\n";
  each([&](int i, int j, int k) {
    put_i(i, " ");
    put_i(j, " ");
    put_i(k, "\n");
  });
  if (binary) out << "\n";
}

// Binary dump (little-endian): "LBMD", version, kind (0 AA, 1 AB, 2 dense), q,
// layout, bytes per population, n, steps; then for the sparse engines the fluid
// coordinates (n x 3 int32) and adjacency (n q int32), then the n q populations
// in storage order.
void Propagation::dumpPopulations(const std::string& path, int currentStep) {
  std::ofstream out(path, std::ios::binary);
  const int q = lattice_.size();
  const int64_t n = geometry_.getNumFluidPts();
  const int32_t kind = parameters_.getPropPattern() == 1 ? 0 : 1;
  write_header(out, kind, q, parameters_.getLayout(), (int32_t)sizeof(Pdf), n, currentStep);
  for (int64_t f = 0; f < n; f++) out.write((const char*)geometry_.fluidPts_[f].data(), 12);
  out.write((const char*)geometry_.adjacency_, sizeof(int) * n * q);
  out.write((const char*)dstrb_, sizeof(Pdf) * n * q);
}

// ---------------------------------------------------------------- AA
PropagationAA::~PropagationAA(void) { delete kernel_; }

void PropagationAA::setup(void) {
  commOffset_ = comm_.getCommLength();
  const size_t n = (size_t)geometry_.getNumFluidPts() * lattice_.size();
  dstrb_ = new Pdf[n]();
  void* p = nullptr;
  check(lbm_device_alloc(sizeof(Pdf) * n, &p), "lbm_device_alloc");
  dstrb_d_ = (Pdf*)p;
  kernel_ = new KernelAA_HIP(lattice_, geometry_);
  kernel_->setup();
}

void PropagationAA::initialize(void) { kernel_->initRest(dstrb_d_); }

void PropagationAA::run(void) {
  initialize();
  const int steps = parameters_.getTimesteps();
  kernel_->synchronize();
  const double t0 = now_s();
  const int first = geometry_.getInletStart();
  const int nfirst = geometry_.getInletCount() + geometry_.getOutletCount() + geometry_.getBorderCount();
  for (int t = 0; t < steps; t++) {
    if (t % 2 == 0) {
      kernel_->timestepEvenForce(dstrb_d_, first, nfirst);
      kernel_->timestepEvenForce(dstrb_d_, geometry_.getBulkStart(), geometry_.getBulkCount());
    } else {
      kernel_->timestepOddForce(dstrb_d_, first, nfirst);
      kernel_->timestepOddForce(dstrb_d_, geometry_.getBulkStart(), geometry_.getBulkCount());
    }
    comm_.exchange(dstrb_d_);
  }
  kernel_->synchronize();
  runtime_ = now_s() - t0;
  const size_t n = (size_t)geometry_.getNumFluidPts() * lattice_.size();
  check(lbm_memcpy_d2h(dstrb_, dstrb_d_, sizeof(Pdf) * n), "lbm_memcpy_d2h");
  printMFLUPS(runtime_);
  checkResult(steps);
  if (parameters_.getWriteVtk()) writeVTK(steps);
}

// Even step count: f_s in slot opp(s); odd: through the adjacency
// (src/propagation/PropagationAA.cpp:218-327).
Pdf PropagationAA::getRho(int fluidIdx, int currentStep) {
  return getVelocityRho(fluidIdx, currentStep, nullptr);
}

std::vector<Pdf> PropagationAA::getVelocity(int fluidIdx, int currentStep) {
  std::vector<Pdf> v(3, 0.0);
  getVelocityRho(fluidIdx, currentStep, v.data());
  return v;
}

Pdf PropagationAA::getVelocityRho(int fluidIdx, int currentStep, Pdf* vel) {
  const int q = lattice_.size();
  const int64_t n = geometry_.getNumFluidPts();
  const bool soa = parameters_.getLayout() == 0;
  Pdf rho = 0.0;
  Pdf v[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < q; s++) {
    int64_t loc;
    if (currentStep % 2 == 0) {
      const int o = lattice_.getOpp(s);
      loc = soa ? n * o + fluidIdx : (int64_t)q * fluidIdx + o;
    } else {
      loc = geometry_.adjacency_[soa ? n * s + fluidIdx : (int64_t)q * fluidIdx + s];
    }
    const Pdf x = dstrb_[loc];
    rho += x;
    for (int d = 0; d < 3; d++) v[d] += x * lattice_.stencil_[s].dir[d];
  }
  if (vel) {
    const Pdf inv = 1.0 / rho;
    for (int d = 0; d < 3; d++) vel[d] = v[d] * inv;
  }
  return rho;
}

// ---------------------------------------------------------------- AB
PropagationAB::~PropagationAB(void) { delete kernel_; }

void PropagationAB::setup(void) {
  commOffset_ = comm_.getCommLength();
  const size_t n = (size_t)geometry_.getNumFluidPts() * lattice_.size();
  dstrb_ = new Pdf[n]();
  dstrb2_ = new Pdf[n]();
  void* p = nullptr;
  check(lbm_device_alloc(sizeof(Pdf) * n, &p), "lbm_device_alloc");
  dstrb_d_ = (Pdf*)p;
  check(lbm_device_alloc(sizeof(Pdf) * n, &p), "lbm_device_alloc");
  dstrb2_d_ = (Pdf*)p;
  kernel_ = new KernelAB_HIP(lattice_, geometry_);
  kernel_->setup();
}

void PropagationAB::initialize(void) {
  kernel_->initRest(dstrb_d_);
  kernel_->initRest(dstrb2_d_);
}

void PropagationAB::run(void) {
  initialize();
  const int steps = parameters_.getTimesteps();
  kernel_->synchronize();
  const double t0 = now_s();
  for (int t = 0; t < steps; t++) {
    kernel_->timestepForce(dstrb_d_, dstrb2_d_, geometry_.getBorderStart(), geometry_.getBorderCount());
    kernel_->timestepForce(dstrb_d_, dstrb2_d_, geometry_.getBulkStart(), geometry_.getBulkCount());
    comm_.exchange(dstrb2_d_);
    std::swap(dstrb_d_, dstrb2_d_);
  }
  kernel_->synchronize();
  runtime_ = now_s() - t0;
  const size_t n = (size_t)geometry_.getNumFluidPts() * lattice_.size();
  check(lbm_memcpy_d2h(dstrb_, dstrb_d_, sizeof(Pdf) * n), "lbm_memcpy_d2h");
  printMFLUPS(runtime_);
  checkResult(steps - 1);
  if (parameters_.getWriteVtk()) writeVTK(steps - 1);
}

// Local slots (src/propagation/PropagationAB.cpp:222-281).
Pdf PropagationAB::getRho(int fluidIdx, int) {
  const int q = lattice_.size();
  const int64_t n = geometry_.getNumFluidPts();
  const bool soa = parameters_.getLayout() == 0;
  Pdf rho = 0.0;
  for (int s = 0; s < q; s++) rho += dstrb_[soa ? n * s + fluidIdx : (int64_t)q * fluidIdx + s];
  return rho;
}

std::vector<Pdf> PropagationAB::getVelocity(int fluidIdx, int) {
  const int q = lattice_.size();
  const int64_t n = geometry_.getNumFluidPts();
  const bool soa = parameters_.getLayout() == 0;
  Pdf rho = 0.0;
  std::vector<Pdf> vel(3, 0.0);
  for (int s = 0; s < q; s++) {
    const Pdf v = dstrb_[soa ? n * s + fluidIdx : (int64_t)q * fluidIdx + s];
    rho += v;
    for (int d = 0; d < 3; d++) vel[d] += v * lattice_.stencil_[s].dir[d];
  }
  const Pdf inv = 1.0 / rho;
  for (int d = 0; d < 3; d++) vel[d] *= inv;
  return vel;
}

// ---------------------------------------------------------------- dense engine
PropagationDense::~PropagationDense(void) {
  if (h_) lbm_dense_destroy(h_);
}

void PropagationDense::setup(void) {
  lbm_dense_config cfg{};
  cfg.nx = geometry_.getDimsByDim(0);
  cfg.ny = geometry_.getDimsByDim(1);
  cfg.nz = geometry_.getDimsByDim(2);
  cfg.q = lattice_.size();
  cfg.pattern = parameters_.getPropPattern() == 1 ? LBM_PATTERN_AA : LBM_PATTERN_AB;
  cfg.pull = parameters_.getPullSetting();
  cfg.layout = parameters_.getLayout() == 0 ? LBM_LAYOUT_SOA : LBM_LAYOUT_AOS;
  const std::string& st = parameters_.getStorage();
  cfg.storage = st == "f32s" ? LBM_STORE_F32S : (st == "f16s" ? LBM_STORE_F16S : LBM_STORE_F64);
  for (int d = 0; d < 3; d++) {
    cfg.periodic[d] = geometry_.getPeriodic(d);
    cfg.force[d] = parameters_.getForce()[d];
  }
  cfg.omega = parameters_.getOmega();
  cfg.density_jump_x = parameters_.getDensityJump();
  bool any_solid = false;
  for (uint8_t v : geometry_.getSolidMask()) any_solid |= v != 0;
  cfg.solid = any_solid ? geometry_.getSolidMask().data() : nullptr;
  check(lbm_dense_create(&cfg, &h_), "lbm_dense_create");
}

void PropagationDense::initialize(void) {
  if (parameters_.getInit() == "tgv")
    check(lbm_dense_init_tgv(h_, parameters_.getU0(), parameters_.getNoise(), parameters_.getSeed()),
          "lbm_dense_init_tgv");
  else {
    const int64_t ncell = (int64_t)geometry_.getDimsByDim(0) * geometry_.getDimsByDim(1) *
                          geometry_.getDimsByDim(2);
    std::vector<double> u(3 * ncell, 0.0);
    check(lbm_dense_init_equilibrium(h_, nullptr, u.data()), "lbm_dense_init_equilibrium");
  }
}

void PropagationDense::run(void) {
  initialize();
  const int steps = parameters_.getTimesteps();
  double m0 = 0.0, m1 = 0.0, ms = 0.0;
  check(lbm_dense_mass(h_, &m0), "lbm_dense_mass");
  check(lbm_dense_synchronize(h_), "sync");
  const double t0 = now_s();
  check(lbm_dense_run_timed(h_, steps, &ms), "lbm_dense_run_timed");
  runtime_ = now_s() - t0;
  check(lbm_dense_mass(h_, &m1), "lbm_dense_mass");
  printMFLUPS(runtime_);
  lbm_dense_info info{};
  check(lbm_dense_get_info(h_, &info), "lbm_dense_get_info");
  const double gbs = info.bytes_per_lup * info.n_fluid * (double)steps / (ms * 1e-3) / 1e9;
  std::cout << "Device time " << ms << " ms, " << gbs << " GB/s of algorithmic traffic ("
            << 100.0 * gbs / 8000.0 << " % of 8 TB/s), mass drift " << (m1 - m0) / m0 << std::endl;
  const int64_t ncell = info.ncell;
  rho_.resize(ncell);
  u_.resize(3 * ncell);
  check(lbm_dense_moments(h_, rho_.data(), u_.data()), "lbm_dense_moments");
  checkResult(steps);
  if (parameters_.getWriteVtk()) writeVTK(steps);
}

int64_t PropagationDense::cellOf(int fluidIdx) {
  const auto& p = geometry_.fluidPts_[fluidIdx];
  const int64_t nx = geometry_.getDimsByDim(0), ny = geometry_.getDimsByDim(1);
  return (p[0] - geometry_.getMinByDim(0)) + nx * ((p[1] - geometry_.getMinByDim(1)) +
                                                   ny * (p[2] - geometry_.getMinByDim(2)));
}

Pdf PropagationDense::getRho(int fluidIdx, int) { return (Pdf)rho_[cellOf(fluidIdx)]; }

std::vector<Pdf> PropagationDense::getVelocity(int fluidIdx, int) {
  const int64_t c = cellOf(fluidIdx);
  return {(Pdf)u_[3 * c], (Pdf)u_[3 * c + 1], (Pdf)u_[3 * c + 2]};
}

void PropagationDense::dumpPopulations(const std::string& path, int currentStep) {
  lbm_dense_info info{};
  check(lbm_dense_get_info(h_, &info), "lbm_dense_get_info");
  std::vector<double> f((size_t)info.ncell * lattice_.size());
  check(lbm_dense_get_state(h_, f.data()), "lbm_dense_get_state");
  std::ofstream out(path, std::ios::binary);
  write_header(out, 2, lattice_.size(), 0, 8, info.ncell, currentStep);
  out.write((const char*)f.data(), sizeof(double) * f.size());
}
