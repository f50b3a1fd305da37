// Lattice.hh -- velocity sets with the reference's interface
// (src/lattices/Lattice.hh:12-35, LatticeD3Q19.hh, LatticeD3Q27.hh).
#ifndef LBM_HOST_LATTICE_H
#define LBM_HOST_LATTICE_H

#include <vector>

typedef int LatticeT;

typedef struct {
  int opp;
  int dir[3];
  double c;  // weight
} Stencil;

class Lattice {
 public:
  Lattice(void) {}
  virtual ~Lattice(void) {}
  int getVel(int velIdx, int dim) { return stencil_[velIdx].dir[dim]; }
  int getOpp(int velIdx) { return stencil_[velIdx].opp; }
  int size() const { return (int)stencil_.size(); }

  std::vector<Stencil> stencil_;
  double inv_cs2, cs2;
  int interactDistance;

 protected:
  void build(int q);
};

// D3Q19: order +-x, +-y, +-z, xy, xz, yz edges, rest (src/lattices/LatticeD3Q19.cpp:12-86)
class LatticeD3Q19 : public Lattice {
 public:
  LatticeD3Q19(void) { build(19); }
};

// D3Q27: D3Q19's 18 moving directions, 8 corners, rest (src/lattices/LatticeD3Q27.cpp:12-118)
class LatticeD3Q27 : public Lattice {
 public:
  LatticeD3Q27(void) { build(27); }
};

#endif
