// Propagation.hh -- time loops with the reference's interface
// (src/propagation/Propagation.hh:9-55, PropagationAA.hh, PropagationAB.hh) on
// one GPU, and the single-process Communication stub.
//
// PropagationAA / PropagationAB drive the drop-in kernels on the reference's
// fluid-only node list (two launches per step, border then bulk, as
// src/propagation/PropagationAA.cpp:169-183 and PropagationAB.cpp:182-188).
// PropagationDense runs the structured-grid engine (lbm_dense_*) on the same
// geometry's solid mask.
#ifndef LBM_HOST_PROPAGATION_H
#define LBM_HOST_PROPAGATION_H

#include <string>
#include <vector>

#include "Geometry.hh"
#include "Kernel.hh"
#include "Lattice.hh"
#include "Parameters.hh"

struct lbm_dense;

// One process owns the whole domain, so the halo exchange of
// src/Communication.cpp:379-481 has nothing to send (commLength_ == 0 there too).
class Communication {
 public:
  Communication(Lattice&, Geometry&, Parameters&) {}
  void identifyNeighbors() {}
  void identifyPatternAB() {}
  void identifyPatternAA() {}
  void kokkosSyclNoNeighborSetup() {}
  void exchange(Pdf*) {}
  int getCommLength() { return 0; }
};

class Propagation {
 public:
  Propagation(Lattice& lattice, Geometry& geometry, Communication& comm, Parameters& parameters);
  virtual ~Propagation(void);
  virtual void setup(void) = 0;
  virtual void initialize(void) = 0;
  virtual void run(void) = 0;
  virtual Pdf getRho(int fluidIdx, int currentStep) = 0;
  virtual std::vector<Pdf> getVelocity(int fluidIdx, int currentStep) = 0;

  void writeVTK(int currentStep);
  void checkResult(int currentStep);
  void printMFLUPS(double runtime);

  // extensions
  double lastRuntime() const { return runtime_; }
  virtual void dumpPopulations(const std::string& path, int currentStep);

 protected:
  Lattice& lattice_;
  Geometry& geometry_;
  Communication& comm_;
  Parameters& parameters_;
  int myRank_ = 0, nTasks_ = 1;
  Pdf *dstrb_ = nullptr, *dstrb2_ = nullptr;      // host copies for readout
  Pdf *dstrb_d_ = nullptr, *dstrb2_d_ = nullptr;  // device populations
  int commOffset_ = 0;
  double runtime_ = 0.0;
};

class PropagationAA : public Propagation {
 public:
  PropagationAA(Lattice& lattice, Geometry& geometry, Communication& comm, Parameters& parameters)
      : Propagation(lattice, geometry, comm, parameters) {}
  ~PropagationAA(void);
  void setup(void);
  void initialize(void);
  void run(void);
  Pdf getRho(int fluidIdx, int currentStep);
  std::vector<Pdf> getVelocity(int fluidIdx, int currentStep);

 protected:
  KernelAA_HIP* kernel_ = nullptr;
  Pdf getVelocityRho(int fluidIdx, int currentStep, Pdf* vel);
};

class PropagationAB : public Propagation {
 public:
  PropagationAB(Lattice& lattice, Geometry& geometry, Communication& comm, Parameters& parameters)
      : Propagation(lattice, geometry, comm, parameters) {}
  ~PropagationAB(void);
  void setup(void);
  void initialize(void);
  void run(void);
  Pdf getRho(int fluidIdx, int currentStep);
  std::vector<Pdf> getVelocity(int fluidIdx, int currentStep);

 protected:
  KernelAB_HIP* kernel_ = nullptr;
};

class PropagationDense : public Propagation {
 public:
  PropagationDense(Lattice& lattice, Geometry& geometry, Communication& comm,
                   Parameters& parameters)
      : Propagation(lattice, geometry, comm, parameters) {}
  ~PropagationDense(void);
  void setup(void);
  void initialize(void);
  void run(void);
  Pdf getRho(int fluidIdx, int currentStep);
  std::vector<Pdf> getVelocity(int fluidIdx, int currentStep);
  void dumpPopulations(const std::string& path, int currentStep);

 protected:
  lbm_dense* h_ = nullptr;
  std::vector<double> rho_, u_;
  int64_t cellOf(int fluidIdx);
};

#endif
