// main.cpp -- deck-driven driver with the structure of the reference's
// src/Main.cpp:22-145 (one process, one GPU):
//   Parameters::readInput -> Lattice -> GeometryAAAB (decomposeDomain, setupGrid,
//   setup, assignFluidIndices, setupAdjacency) -> Communication -> Propagation
//   (AB for propPattern 0, AA for 1; `engine dense` selects the structured-grid
//   engine) -> setup -> run.
//
//   lbm-proxy-app <deck> [--dump-geometry <file>]
// --dump-geometry writes the fluid coordinates and adjacency (binary, see
// Propagation::dumpPopulations for the header) and exits without touching a GPU.
#include <chrono>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

#include "Geometry.hh"
#include "Lattice.hh"
#include "Parameters.hh"
#include "Propagation.hh"

static void dump_geometry(Geometry& g, Lattice& lat, Parameters& p, const std::string& path) {
  std::ofstream out(path, std::ios::binary);
  out.write("LBMG", 4);
  const int32_t hdr[4] = {1, lat.size(), p.getLayout(), g.getAdjacencyMode()};
  out.write((const char*)hdr, sizeof(hdr));
  const int64_t n = g.getNumFluidPts();
  const int32_t dims[3] = {g.getDimsByDim(0), g.getDimsByDim(1), g.getDimsByDim(2)};
  const int32_t ranges[4] = {g.getBorderStart(), g.getBorderCount(), g.getBulkStart(),
                             g.getBulkCount()};
  out.write((const char*)&n, 8);
  out.write((const char*)dims, sizeof(dims));
  out.write((const char*)ranges, sizeof(ranges));
  for (int64_t f = 0; f < n; f++) out.write((const char*)g.fluidPts_[f].data(), 12);
  out.write((const char*)g.adjacency_, sizeof(int) * n * lat.size());
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <input deck> [--dump-geometry <file>]" << std::endl;
    return 2;
  }
  try {
    std::cout << "Run started on 1 tasks" << std::endl;
    const auto t_start = std::chrono::steady_clock::now();
    Parameters parameters;
    parameters.readInput(argv[1]);
    std::unique_ptr<Lattice> lattice;
    if (parameters.getLatticeSize() == 19) lattice.reset(new LatticeD3Q19);
    else lattice.reset(new LatticeD3Q27);
    std::unique_ptr<Geometry> geometry(new GeometryAAAB(*lattice, parameters));
    geometry->decomposeDomain();
    geometry->setupGrid();
    geometry->setup();
    geometry->assignFluidIndices();
    geometry->setupAdjacency();
    if (argc > 3 && std::string(argv[2]) == "--dump-geometry") {
      dump_geometry(*geometry, *lattice, parameters, argv[3]);
      return 0;
    }
    Communication comm(*lattice, *geometry, parameters);
    comm.kokkosSyclNoNeighborSetup();
    std::unique_ptr<Propagation> prop;
    if (parameters.getEngine() == "dense")
      prop.reset(new PropagationDense(*lattice, *geometry, comm, parameters));
    else if (parameters.getPropPattern() == 0)
      prop.reset(new PropagationAB(*lattice, *geometry, comm, parameters));
    else
      prop.reset(new PropagationAA(*lattice, *geometry, comm, parameters));
    prop->setup();
    prop->run();
    if (!parameters.getDumpPath().empty())
      prop->dumpPopulations(parameters.getDumpPath(), parameters.getTimesteps());
    const double total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::cout << "Total runtime (rank 0): " << total << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
