// Geometry.hh -- fluid-node list, grid map and adjacency with the reference's
// interface (src/geometries/Geometry.hh:46-127, GeometryAAAB.hh:8-30), for one
// process: the whole box is this process's subdomain.
//
// The solid mask comes from the engine's generators (include/lbm_capi.h):
// geometry "cylinder" is the reference's pipe (src/geometries/Geometry.cpp:60-70,
// 113-131); "box", "channel", "spheres" (channel + spheres) and "porous"
// (spheres in a periodic box) are sized by the deck key `size`.
#ifndef LBM_HOST_GEOMETRY_H
#define LBM_HOST_GEOMETRY_H

#include <cstdint>
#include <vector>

#include "Lattice.hh"
#include "Parameters.hh"

enum LAT_PT_TYPES_ {
  LAT_PT_WALL = -10,
  LAT_PT_INLET = -11,
  LAT_PT_OUTLET = -12
};

class Geometry {
 public:
  Geometry(Lattice& lattice, Parameters& parameters);
  virtual ~Geometry(void);
  void decomposeDomain();
  void setupGrid();
  void assignFluidIndices();
  virtual void setup() = 0;
  virtual void setupAdjacency() = 0;

  int getGridIdx(int i, int j, int k);
  int64_t getGlobalNumFluidPts() { return nFluidGlobal_; }
  int getNumFluidPts() { return nFluid_; }
  int getTotalLengthByDim(int dim) { return totalLength_[dim]; }
  int getMinByDim(int dim) { return myMin_[dim]; }
  int getMaxByDim(int dim) { return myMax_[dim]; }
  int getDimsByDim(int dim) { return myDims_[dim]; }
  int getGridElement(int idx) { return grid_[idx]; }

  int getInletStart() { return startInlet_; }
  int getInletCount() { return countInlet_; }
  int getOutletStart() { return startOutlet_; }
  int getOutletCount() { return countOutlet_; }
  int getBorderStart() { return startBorder_; }
  int getBorderCount() { return countBorder_; }
  int getBulkStart() { return startBulk_; }
  int getBulkCount() { return countBulk_; }

  // extensions
  Parameters& getParameters() { return parameters_; }
  Lattice& getLattice() { return lattice_; }
  int getPeriodic(int dim) const { return periodic_[dim]; }
  // solid mask in the engine's cell order x + nx * (y + ny * z)
  const std::vector<uint8_t>& getSolidMask() const { return solid_; }
  // adjacency mode: 0 = neighbour at x + c_s (AB push), 1 = x - c_s (AA, AB pull)
  int getAdjacencyMode() const { return adjacencyMode_; }

  int* adjacency_ = nullptr;
  std::vector<std::vector<int> > fluidPts_;

 protected:
  Lattice& lattice_;
  Parameters& parameters_;
  std::vector<int> totalLength_, myMin_, myMax_, myDims_;
  std::vector<int> grid_;
  std::vector<uint8_t> solid_;
  int periodic_[3] = {1, 0, 0};
  int adjacencyMode_ = 1;
  int64_t nFluidGlobal_ = 0;
  int nFluid_ = 0;
  int startInlet_ = 0, countInlet_ = 0, startOutlet_ = 0, countOutlet_ = 0;
  int startBorder_ = 0, countBorder_ = 0, startBulk_ = 0, countBulk_ = 0;
};

class GeometryAAAB : public Geometry {
 public:
  GeometryAAAB(Lattice& lattice, Parameters& parameters) : Geometry(lattice, parameters) {}
  ~GeometryAAAB(void) {}
  void setup();
  void setupAdjacency();
  void sortFluidPts();

 private:
  bool checkPtOnTaskBoundary(const std::vector<int>& fluidPt);
  int64_t getSortIdxFromPt(const std::vector<int>& fluidPt);
};

#endif
