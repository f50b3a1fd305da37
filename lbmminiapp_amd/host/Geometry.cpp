// Geometry.cpp -- one-process geometry: box, solid mask, fluid ordering and
// adjacency (the reference's GeometryAAAB semantics).
#include "Geometry.hh"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "lbm_capi.h"

namespace {
void check(lbm_status st, const char* what) {
  if (st != LBM_OK)
    throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

// Mask file: "LBMMASK1", int32 nx, ny, nz (little-endian), then one byte per
// cell, x fastest, 0 fluid and 1 solid.
const char kMaskMagic[8] = {'L', 'B', 'M', 'M', 'A', 'S', 'K', '1'};

void read_mask_header(std::ifstream& in, const std::string& path, int32_t dims[3]) {
  char magic[8];
  in.read(magic, 8);
  in.read(reinterpret_cast<char*>(dims), 3 * sizeof(int32_t));
  if (!in || std::memcmp(magic, kMaskMagic, 8) != 0)
    throw std::runtime_error("'" + path + "' is not a mask file (LBMMASK1)");
  for (int d = 0; d < 3; d++)
    if (dims[d] <= 0) throw std::runtime_error("mask file '" + path + "': bad dimensions");
}

void mask_file_dims(const std::string& path, int32_t dims[3]) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open mask file '" + path + "'");
  read_mask_header(in, path, dims);
}

void read_mask_file(const std::string& path, const int32_t want[3], std::vector<uint8_t>& solid) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open mask file '" + path + "'");
  int32_t dims[3];
  read_mask_header(in, path, dims);
  for (int d = 0; d < 3; d++)
    if (dims[d] != want[d]) throw std::runtime_error("mask file '" + path + "': dimensions changed");
  const int64_t ncell = (int64_t)dims[0] * dims[1] * dims[2];
  solid.assign(ncell, 0);
  in.read(reinterpret_cast<char*>(solid.data()), ncell);
  if (in.gcount() != ncell) throw std::runtime_error("mask file '" + path + "' is truncated");
  for (auto& v : solid) v = v ? 1 : 0;
}

void write_mask_file(const std::string& path, const int32_t dims[3],
                     const std::vector<uint8_t>& solid) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write mask file '" + path + "'");
  std::vector<uint8_t> bytes(solid.size());
  for (size_t c = 0; c < solid.size(); c++) bytes[c] = solid[c] ? 1 : 0;
  out.write(kMaskMagic, 8);
  out.write(reinterpret_cast<const char*>(dims), 3 * sizeof(int32_t));
  out.write(reinterpret_cast<const char*>(bytes.data()), (std::streamsize)bytes.size());
  if (!out) throw std::runtime_error("writing mask file '" + path + "' failed");
}
}  // namespace

Geometry::Geometry(Lattice& lattice, Parameters& parameters)
    : lattice_(lattice), parameters_(parameters) {}

Geometry::~Geometry(void) { delete[] adjacency_; }

// One process owns the whole box (src/geometries/Geometry.cpp:23-97 with one task).
void Geometry::decomposeDomain() {
  const std::string& kind = parameters_.getGeometry();
  totalLength_.assign(3, 0);
  if (kind == "cylinder") {
    int32_t dims[3];
    check(lbm_geometry_cylinder(parameters_.getSimulationSize(), dims, nullptr), "cylinder");
    for (int d = 0; d < 3; d++) totalLength_[d] = dims[d];
    periodic_[0] = 1;  // the reference wraps x when there is one task
    periodic_[1] = periodic_[2] = 0;
  } else if (kind == "file") {
    int32_t dims[3];
    mask_file_dims(parameters_.getMaskFile(), dims);
    for (int d = 0; d < 3; d++) totalLength_[d] = dims[d];
    periodic_[0] = periodic_[1] = periodic_[2] = 1;  // the deck's `periodic` sets the walls
  } else {
    for (int d = 0; d < 3; d++) totalLength_[d] = parameters_.getSize(d);
    periodic_[0] = 1;
    periodic_[1] = (kind == "box" || kind == "porous") ? 1 : 0;
    periodic_[2] = 1;
  }
  if (parameters_.periodicGiven())
    for (int d = 0; d < 3; d++) periodic_[d] = parameters_.getPeriodic(d);
  myMin_.assign(3, 0);
  myMax_.resize(3);
  for (int d = 0; d < 3; d++) myMax_[d] = totalLength_[d] - 1;
}

// Solid mask from the generators; fluid points in x-major order, z fastest
// (the loop order of src/geometries/Geometry.cpp:113-131).
void Geometry::setupGrid() {
  myDims_.resize(3);
  for (int d = 0; d < 3; d++) myDims_[d] = myMax_[d] - myMin_[d] + 1;
  const int nx = myDims_[0], ny = myDims_[1], nz = myDims_[2];
  const int64_t ncell = (int64_t)nx * ny * nz;
  solid_.assign(ncell, 0);
  const std::string& kind = parameters_.getGeometry();
  if (kind == "cylinder") {
    int32_t dims[3];
    check(lbm_geometry_cylinder(parameters_.getSimulationSize(), dims, solid_.data()), "cylinder");
  } else if (kind == "channel" || kind == "spheres") {
    check(lbm_geometry_channel(nx, ny, nz, solid_.data()), "channel");
  } else if (kind == "file") {
    const int32_t want[3] = {nx, ny, nz};
    read_mask_file(parameters_.getMaskFile(), want, solid_);
  } else if (kind != "box" && kind != "porous") {
    throw std::runtime_error("unknown geometry '" + kind + "'");
  }
  if (kind == "spheres" || kind == "porous") {
    int64_t count = 0;
    check(lbm_geometry_add_spheres(nx, ny, nz, periodic_, parameters_.getRadiusMin(),
                                   parameters_.getRadiusMax(), parameters_.getSolidFraction(),
                                   parameters_.getSeed(), solid_.data(), &count),
          "spheres");
  }
  if (!parameters_.getWriteMask().empty()) {
    const int32_t dims[3] = {nx, ny, nz};
    write_mask_file(parameters_.getWriteMask(), dims, solid_);
  }
  grid_.assign(ncell, LAT_PT_WALL);
  fluidPts_.clear();
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        if (!solid_[i + (int64_t)nx * (j + (int64_t)ny * k)])
          fluidPts_.push_back({myMin_[0] + i, myMin_[1] + j, myMin_[2] + k});
  if ((int64_t)fluidPts_.size() * lattice_.size() >= ((int64_t)1 << 31))
    throw std::runtime_error("geometry too large for the int32 adjacency of the sparse engine");
  nFluid_ = (int)fluidPts_.size();
  nFluidGlobal_ = nFluid_;
  std::cout << "Simulation has " << nFluidGlobal_ << " fluid points" << std::endl;
}

void Geometry::assignFluidIndices() {
  for (int f = 0; f < nFluid_; f++)
    grid_[getGridIdx(fluidPts_[f][0], fluidPts_[f][1], fluidPts_[f][2])] = f;
}

int Geometry::getGridIdx(int i, int j, int k) {
  const int ri = i - myMin_[0], rj = j - myMin_[1], rk = k - myMin_[2];
  if (ri < 0 || rj < 0 || rk < 0) return -1;
  if (ri >= myDims_[0] || rj >= myDims_[1] || rk >= myDims_[2]) return -1;
  return myDims_[2] * myDims_[1] * ri + myDims_[2] * rj + rk;
}

// ---------------------------------------------------------------- GeometryAAAB
void GeometryAAAB::setup() { sortFluidPts(); }

// Border points (within interactDistance of the box faces) first, then bulk,
// each in x-major order (src/geometries/GeometryAAAB.cpp:127-205 for one task).
void GeometryAAAB::sortFluidPts() {
  std::vector<std::vector<int> > border, bulk;
  for (auto& p : fluidPts_) (checkPtOnTaskBoundary(p) ? border : bulk).push_back(p);
  auto by_key = [&](const std::vector<int>& a, const std::vector<int>& b) {
    return getSortIdxFromPt(a) < getSortIdxFromPt(b);
  };
  std::sort(border.begin(), border.end(), by_key);
  std::sort(bulk.begin(), bulk.end(), by_key);
  startInlet_ = countInlet_ = startOutlet_ = countOutlet_ = 0;
  startBorder_ = 0;
  countBorder_ = (int)border.size();
  startBulk_ = countBorder_;
  countBulk_ = (int)bulk.size();
  fluidPts_ = border;
  fluidPts_.insert(fluidPts_.end(), bulk.begin(), bulk.end());
}

bool GeometryAAAB::checkPtOnTaskBoundary(const std::vector<int>& p) {
  const int d = lattice_.interactDistance;
  for (int k = 0; k < 3; k++)
    if (p[k] - d < myMin_[k] || p[k] + d > myMax_[k]) return true;
  return false;
}

int64_t GeometryAAAB::getSortIdxFromPt(const std::vector<int>& p) {
  return (int64_t)myDims_[2] * myDims_[1] * (p[0] - myMin_[0]) +
         (int64_t)myDims_[2] * (p[1] - myMin_[1]) + (p[2] - myMin_[2]);
}

// Neighbour x + c_s (AB push) or x - c_s (AA, AB pull); periodic faces wrap;
// a neighbour outside the box or solid bounces back into the node's own
// opposite slot (src/geometries/GeometryAAAB.cpp:18-124).
void GeometryAAAB::setupAdjacency() {
  const int q = lattice_.size();
  const int n = nFluid_;
  const int propPattern = parameters_.getPropPattern();
  const int pull = parameters_.getPullSetting();
  adjacencyMode_ = (propPattern == 0 && pull == 0) ? 0 : 1;
  const bool soa = parameters_.getLayout() == 0;
  delete[] adjacency_;
  adjacency_ = new int[(size_t)n * q];
  for (int f = 0; f < n; f++) {
    for (int s = 0; s < q; s++) {
      const int sd = adjacencyMode_ == 0 ? s : lattice_.getOpp(s);
      int nb[3];
      bool inside = true;
      for (int d = 0; d < 3; d++) {
        nb[d] = fluidPts_[f][d] + lattice_.getVel(sd, d);
        if (periodic_[d]) {
          if (nb[d] < myMin_[d]) nb[d] += myDims_[d];
          else if (nb[d] > myMax_[d]) nb[d] -= myDims_[d];
        }
        if (nb[d] < myMin_[d] || nb[d] > myMax_[d]) inside = false;
      }
      const int idx = inside ? getGridIdx(nb[0], nb[1], nb[2]) : -1;
      const int nf = idx < 0 ? -1 : grid_[idx];
      const int o = lattice_.getOpp(s);
      int value;
      if (nf < 0) value = soa ? n * o + f : q * f + o;
      else value = soa ? n * s + nf : q * nf + s;
      adjacency_[soa ? (size_t)n * s + f : (size_t)q * f + s] = value;
    }
  }
}
