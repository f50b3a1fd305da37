// ref_harness.cpp -- drives the reference's own collide-stream kernels
// (compiled unmodified from /root/reference/src by oracle/Makefile.ref) on an
// adjacency and population arrays supplied by the caller.
//
// TEST INFRASTRUCTURE ONLY: used to pin the C restatement (lbm_oracle.c), to
// write golden fixtures (tests/golden/gen_golden.py) and, optionally, as the
// timed CPU baseline.  Never linked into the product.
//
// The reference's Geometry constructor calls MPI_Comm_rank
// (src/geometries/Geometry.cpp:18-19) and Parameters' constructor does too
// (src/Parameters.cpp:13-17), so MPI is initialised once (singleton init).
// HarnessGeometry fills the protected nFluid_ and the public adjacency_ that the
// kernels read (src/kernels/KernelAA_SOA_Base.cpp:174,273); everything else of
// Geometry stays unused.
#include <mpi.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "Geometry.hh"
#include "GeometryAAAB.hh"
#include "KernelAA_AOS_Base.hh"
#include "KernelAA_SOA_Base.hh"
#include "KernelAA_SOA_Unroll.hh"
#include "KernelAB_AOS_Base.hh"
#include "KernelAB_SOA_Base.hh"
#include "KernelAB_SOA_Pull.hh"
#include "KernelAB_SOA_Unroll.hh"
#include "KernelAB_SOA_Unroll_Pull.hh"
#include "LatticeD3Q19.hh"
#include "LatticeD3Q27.hh"
#include "Parameters.hh"

namespace {

void ensure_mpi() {
  int inited = 0;
  MPI_Initialized(&inited);
  if (!inited) MPI_Init(nullptr, nullptr);
}

class HarnessGeometry : public Geometry {
 public:
  HarnessGeometry(Lattice& lattice, Parameters& params, int nfluid, int* adj)
      : Geometry(lattice, params) {
    nFluid_ = nfluid;
    nFluidGlobal_ = nfluid;
    adjacency_ = adj;
  }
  void setup() override {}
  void setupAdjacency() override {}
};

enum Variant {
  AA_SOA_BASE = 0,
  AA_SOA_UNROLL = 1,
  AA_AOS_BASE = 2,
  AB_SOA_BASE = 3,
  AB_SOA_PULL = 4,
  AB_SOA_UNROLL = 5,
  AB_SOA_UNROLL_PULL = 6,
  AB_AOS_BASE = 7,
};

}  // namespace

extern "C" {

int ref_lattice_size() { return _LATTICESIZE_; }
int ref_pdf_bytes() { return (int)sizeof(Pdf); }
double ref_omega() { return _OMEGA_; }

// Runs `steps` time steps of one reference kernel variant on the whole node
// range [0, nfluid) (one launch per step, as a single-rank run with no border
// split).  AA: `f` is updated in place, step t even -> timestepEvenForce,
// t odd -> timestepOddForce (PropagationAA.cpp:169-183), starting at parity
// `t0`.  AB: `f` and `f2` alternate as source and target
// (PropagationAB.cpp:182-188); after the call `f` holds the latest state.
// Returns the wall time of the step loop in seconds, or a negative value on
// error.
double ref_run(int variant, int nfluid, int* adj, Pdf* f, Pdf* f2, int steps, int t0) {
  ensure_mpi();
  Parameters params;
  Lattice* lattice = nullptr;
  if (_LATTICESIZE_ == 19) lattice = new LatticeD3Q19;
  else lattice = new LatticeD3Q27;
  HarnessGeometry geom(*lattice, params, nfluid, adj);
  KernelAA* aa = nullptr;
  KernelAB* ab = nullptr;
  switch (variant) {
    case AA_SOA_BASE: aa = new KernelAA_SOA_Base(*lattice, geom); break;
    case AA_SOA_UNROLL: aa = new KernelAA_SOA_Unroll(*lattice, geom); break;
    case AA_AOS_BASE: aa = new KernelAA_AOS_Base(*lattice, geom); break;
    case AB_SOA_BASE: ab = new KernelAB_SOA_Base(*lattice, geom); break;
    case AB_SOA_PULL: ab = new KernelAB_SOA_Pull(*lattice, geom); break;
    case AB_SOA_UNROLL: ab = new KernelAB_SOA_Unroll(*lattice, geom); break;
    case AB_SOA_UNROLL_PULL: ab = new KernelAB_SOA_Unroll_Pull(*lattice, geom); break;
    case AB_AOS_BASE: ab = new KernelAB_AOS_Base(*lattice, geom); break;
    default: return -1.0;
  }
  if (aa) aa->setup();
  if (ab) ab->setup();
  auto t_start = std::chrono::steady_clock::now();
  Pdf* src = f;
  Pdf* tgt = f2;
  for (int t = t0; t < t0 + steps; t++) {
    if (aa) {
      if (t % 2 == 0) aa->timestepEvenForce(f, 0, nfluid);
      else aa->timestepOddForce(f, 0, nfluid);
    } else {
      ab->timestepForce(src, tgt, 0, nfluid);
      Pdf* tmp = src;
      src = tgt;
      tgt = tmp;
    }
  }
  auto t_end = std::chrono::steady_clock::now();
  if (ab && src != f) std::memcpy(f, src, sizeof(Pdf) * (size_t)nfluid * _LATTICESIZE_);
  delete aa;
  delete ab;
  delete lattice;
  return std::chrono::duration<double>(t_end - t_start).count();
}

// Lattice tables exactly as the reference constructs them
// (src/lattices/LatticeD3Q19.cpp:5-88, LatticeD3Q27.cpp:5-120).
int ref_lattice(int* dirs, int* opp, double* w) {
  Lattice* lattice = nullptr;
  if (_LATTICESIZE_ == 19) lattice = new LatticeD3Q19;
  else lattice = new LatticeD3Q27;
  for (int s = 0; s < _LATTICESIZE_; s++) {
    for (int d = 0; d < 3; d++) dirs[3 * s + d] = lattice->getVel(s, d);
    opp[s] = lattice->getOpp(s);
    w[s] = lattice->stencil_[s].c;
  }
  delete lattice;
  return _LATTICESIZE_;
}

// The reference's own geometry pipeline on a deck file (src/Main.cpp:63-72):
// GeometryAAAB -> decomposeDomain -> setupGrid -> setup -> assignFluidIndices
// -> setupAdjacency, for one MPI task.  With pts == NULL only the fluid count
// and the border/bulk ranges are returned; otherwise pts (n x 3) and adj
// (n x Q, the layout the library was compiled for: SOA) are filled.
int ref_geometry(const char* deck, int* nfluid, int* ranges, int* pts, int* adj) {
  ensure_mpi();
  Parameters params;
  params.readInput(deck);
  Lattice* lattice = nullptr;
  if (_LATTICESIZE_ == 19) lattice = new LatticeD3Q19;
  else lattice = new LatticeD3Q27;
  GeometryAAAB geom(*lattice, params);
  geom.decomposeDomain();
  geom.setupGrid();
  geom.setup();
  geom.assignFluidIndices();
  geom.setupAdjacency();
  const int n = geom.getNumFluidPts();
  *nfluid = n;
  ranges[0] = geom.getBorderStart();
  ranges[1] = geom.getBorderCount();
  ranges[2] = geom.getBulkStart();
  ranges[3] = geom.getBulkCount();
  if (pts) {
    for (int f = 0; f < n; f++)
      for (int d = 0; d < 3; d++) pts[3 * f + d] = geom.fluidPts_[f][d];
    std::memcpy(adj, geom.adjacency_, sizeof(int) * (size_t)n * _LATTICESIZE_);
  }
  delete lattice;
  return 0;
}

}  // extern "C"
