// Kernel.hh -- the reference's kernel interfaces (src/kernels/KernelAA.hh:7-58,
// src/kernels/KernelAB.hh:8-50) and their gfx950 implementations.
//
// KernelAA_HIP / KernelAB_HIP replace every reference kernel class
// (KernelAA_SOA_Base, _SOA_Unroll, _AOS_Base, KernelAB_SOA_Base, _SOA_Pull,
// _SOA_Unroll, _SOA_Unroll_Pull, _AOS_Base): one work-item per fluid node over
// [startIdx, startIdx + countIdx), Base-kernel arithmetic (bit-for-bit), the
// geometry's adjacency copied to the device once in setup().  Pdf pointers are
// device pointers (as the reference's SYCL/Kokkos builds pass device memory).
#ifndef LBM_HOST_KERNEL_H
#define LBM_HOST_KERNEL_H

#include "Geometry.hh"
#include "Lattice.hh"
#include "Parameters.hh"

struct lbm_sparse;

class KernelAA {
 public:
  KernelAA(Lattice& lattice, Geometry& geometry) : lattice_(lattice), geometry_(geometry) {}
  virtual ~KernelAA(void) {}
  virtual void setup() = 0;
  virtual void timestepEvenForce(Pdf* dstrb, int startIdx, int countIdx) = 0;
  virtual void timestepOddForce(Pdf* dstrb, int startIdx, int countIdx) = 0;

 protected:
  Lattice& lattice_;
  Geometry& geometry_;
};

class KernelAB {
 public:
  KernelAB(Lattice& lattice, Geometry& geometry) : lattice_(lattice), geometry_(geometry) {}
  virtual ~KernelAB(void) {}
  virtual void setup() = 0;
  virtual void timestepForce(Pdf* dstrb_src, Pdf* dstrb_tgt, int startIdx, int countIdx) = 0;

 protected:
  Lattice& lattice_;
  Geometry& geometry_;
};

class KernelAA_HIP : public KernelAA {
 public:
  KernelAA_HIP(Lattice& lattice, Geometry& geometry) : KernelAA(lattice, geometry) {}
  ~KernelAA_HIP(void);
  void setup();
  void timestepEvenForce(Pdf* dstrb, int startIdx, int countIdx);
  void timestepOddForce(Pdf* dstrb, int startIdx, int countIdx);
  // sets every population to w_s in slot opp(s) (PropagationAA::initialize)
  void initRest(Pdf* dstrb);
  void synchronize();

 private:
  lbm_sparse* h_ = nullptr;
};

class KernelAB_HIP : public KernelAB {
 public:
  KernelAB_HIP(Lattice& lattice, Geometry& geometry) : KernelAB(lattice, geometry) {}
  ~KernelAB_HIP(void);
  void setup();
  void timestepForce(Pdf* dstrb_src, Pdf* dstrb_tgt, int startIdx, int countIdx);
  void initRest(Pdf* dstrb);
  void synchronize();

 private:
  lbm_sparse* h_ = nullptr;
  int pull_ = 0;
};

#endif
