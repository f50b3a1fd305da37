// Parameters.hh -- input deck of the LBM proxy app (src/Parameters.hh:46-70).
//
// Same whitespace "key value" format and keys as the reference
// (src/Parameters.cpp:50-77): densitychange, proppattern, pull, unroll,
// simulationsize, timesteps, writevtk (camelCase or lowercase), with the
// reference's defaults (src/Parameters.cpp:23-30).  Differences: there is
// one process (no MPI broadcast), a deck that cannot be opened is an error
// (the reference silently keeps the defaults), unknown keys draw a warning,
// and the reference's compile-time physics (_LATTICESIZE_, _OMEGA_, _GRAVITY_,
// SOA/AOS, _PRECISION_SINGLE_; src/Parameters.hh:6-42) are runtime keys:
//   lattice 19|27        tau T | omega W        force fx fy fz
//   layout soa|aos        engine sparse|dense    storage f64|f32s|f16s (dense)
//   geometry cylinder|box|channel|spheres|porous|file   size nx ny nz   periodic px py pz
//   solidfraction f  radiusmin r  radiusmax r  seed s  init rest|tgv  u0 u  noise a
//   densityjump d (engine dense: inlet/outlet pressure drop d / 3 on the periodic x axis)
//   dump path            (binary dump of the final populations, for checks)
//   maskfile path        (solid mask from a mask file: geometry file, dimensions from the file)
//   writemask path       (write the solid mask the run uses to a mask file)
// Mask file (little-endian): the 8 bytes "LBMMASK1", int32 nx, ny, nz, then
// nx*ny*nz bytes, cell x + nx*(y + ny*z), 0 fluid and 1 solid.
#ifndef LBM_HOST_PARAMETERS_H
#define LBM_HOST_PARAMETERS_H

#include <cstdint>
#include <string>

#ifdef _PRECISION_SINGLE_
typedef float Pdf;
#else
typedef double Pdf;
#endif

class Parameters {
 public:
  Parameters(void);
  ~Parameters(void) {}
  void readInput(std::string inputFile);

  double getDensityChange() { return densityChange_; }
  int getGeometryLibID() { return geometryLibID_; }
  int getPropPattern() { return propPattern_; }
  int getPullSetting() { return pullSetting_; }
  int getUnrollSetting() { return unrollSetting_; }
  double getSimulationSize() { return simulationSize_; }
  int getTimesteps() { return timesteps_; }
  int getWriteVtk() { return writeVtk_; }

  // runtime forms of the reference's compile-time settings, and extensions
  int getLatticeSize() const { return latticeSize_; }
  double getOmega() const { return omega_; }
  const double* getForce() const { return force_; }
  int getLayout() const { return layout_; }          // 0 SOA, 1 AOS
  const std::string& getEngine() const { return engine_; }
  const std::string& getStorage() const { return storage_; }
  const std::string& getGeometry() const { return geometry_; }
  int getSize(int dim) const { return size_[dim]; }
  int getPeriodic(int dim) const { return periodic_[dim]; }
  bool periodicGiven() const { return periodicGiven_; }
  double getSolidFraction() const { return solidFraction_; }
  double getDensityJump() const { return densityJump_; }
  double getRadiusMin() const { return radiusMin_; }
  double getRadiusMax() const { return radiusMax_; }
  uint64_t getSeed() const { return seed_; }
  const std::string& getInit() const { return init_; }
  double getU0() const { return u0_; }
  double getNoise() const { return noise_; }
  const std::string& getDumpPath() const { return dump_; }
  const std::string& getMaskFile() const { return maskFile_; }
  const std::string& getWriteMask() const { return writeMask_; }

 private:
  double densityChange_, simulationSize_;
  int geometryLibID_, propPattern_, pullSetting_;
  int unrollSetting_, timesteps_, writeVtk_;

  int latticeSize_ = 19;
  double omega_ = 1.0;
  double force_[3] = {1e-6, 0.0, 0.0};
  int layout_ = 0;
  std::string engine_ = "sparse", storage_ = "f64", geometry_ = "cylinder", init_ = "rest", dump_;
  std::string maskFile_, writeMask_;
  int size_[3] = {64, 64, 64};
  int periodic_[3] = {1, 0, 0};
  bool periodicGiven_ = false;
  double solidFraction_ = 0.05, radiusMin_ = 4.0, radiusMax_ = 12.0;
  double densityJump_ = 0.0;  // pressure-drop boundary of the dense engine (drho across x)
  uint64_t seed_ = 1;
  double u0_ = 0.05, noise_ = 0.0;
};
#endif
