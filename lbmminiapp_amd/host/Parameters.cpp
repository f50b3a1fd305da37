// Parameters.cpp -- deck reader (src/Parameters.cpp:20-92 without MPI).
#include "Parameters.hh"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <vector>

Parameters::Parameters() {
  densityChange_ = 0.01;
  geometryLibID_ = 1;
  propPattern_ = 0;
  pullSetting_ = 0;
  unrollSetting_ = 0;
  simulationSize_ = 1.0;
  timesteps_ = 1;
  writeVtk_ = 0;
}

void Parameters::readInput(std::string inputFile) {
  std::ifstream readFile(inputFile.c_str(), std::ifstream::in);
  if (!readFile.is_open()) throw std::runtime_error("cannot open input deck '" + inputFile + "'");
  std::string line;
  while (std::getline(readFile, line)) {
    std::cout << line << std::endl;
    std::istringstream iss(line);
    std::vector<std::string> p;
    std::string tok;
    while (iss >> tok) p.push_back(tok);
    if (p.empty() || p[0][0] == '#') continue;
    const std::string& k = p[0];
    auto num = [&](size_t i) {
      if (p.size() <= i) throw std::runtime_error("deck key '" + k + "' needs a value");
      return p[i];
    };
    if (k == "densitychange" || k == "densityChange") densityChange_ = std::atof(num(1).c_str());
    else if (k == "proppattern" || k == "propPattern") propPattern_ = std::atoi(num(1).c_str());
    else if (k == "pull") pullSetting_ = std::atoi(num(1).c_str());
    else if (k == "unroll") unrollSetting_ = std::atoi(num(1).c_str());
    else if (k == "simulationsize" || k == "simulationSize") simulationSize_ = std::atof(num(1).c_str());
    else if (k == "timesteps") timesteps_ = std::atoi(num(1).c_str());
    else if (k == "writevtk") writeVtk_ = std::atoi(num(1).c_str());
    else if (k == "lattice") latticeSize_ = std::atoi(num(1).c_str());
    else if (k == "tau") omega_ = 1.0 / std::atof(num(1).c_str());
    else if (k == "omega") omega_ = std::atof(num(1).c_str());
    else if (k == "force") {
      for (int d = 0; d < 3; d++) force_[d] = std::atof(num(1 + d).c_str());
    } else if (k == "layout") layout_ = num(1) == "aos" ? 1 : 0;
    else if (k == "precision") storage_ = num(1) == "single" ? "f32" : "f64";
    else if (k == "engine") engine_ = num(1);
    else if (k == "storage") storage_ = num(1);
    else if (k == "geometry") geometry_ = num(1);
    else if (k == "size") {
      for (int d = 0; d < 3; d++) size_[d] = std::atoi(num(1 + d).c_str());
    } else if (k == "periodic") {
      for (int d = 0; d < 3; d++) periodic_[d] = std::atoi(num(1 + d).c_str());
      periodicGiven_ = true;
    } else if (k == "solidfraction") solidFraction_ = std::atof(num(1).c_str());
    else if (k == "densityjump") densityJump_ = std::atof(num(1).c_str());
    else if (k == "radiusmin") radiusMin_ = std::atof(num(1).c_str());
    else if (k == "radiusmax") radiusMax_ = std::atof(num(1).c_str());
    else if (k == "seed") seed_ = std::strtoull(num(1).c_str(), nullptr, 10);
    else if (k == "init") init_ = num(1);
    else if (k == "u0") u0_ = std::atof(num(1).c_str());
    else if (k == "noise") noise_ = std::atof(num(1).c_str());
    else if (k == "dump") dump_ = num(1);
    else if (k == "maskfile") maskFile_ = num(1);
    else if (k == "writemask") writeMask_ = num(1);
    else std::cerr << "warning: unknown deck key '" << k << "' ignored" << std::endl;
  }
  if (geometry_ == "file" && maskFile_.empty())
    throw std::runtime_error("geometry 'file' needs the deck key 'maskfile'");
  if (!maskFile_.empty() && geometry_ != "file") geometry_ = "file";
  // The reference's decks name a third pattern (MR = 2) that its Main.cpp
  // never constructs (src/Main.cpp:86-121 leave the propagation unset).
  if (propPattern_ != 0 && propPattern_ != 1)
    throw std::runtime_error("propPattern must be 0 (AB) or 1 (AA); pattern 2 (MR) has no "
                             "implementation");
  if (latticeSize_ != 19 && latticeSize_ != 27)
    throw std::runtime_error("deck key 'lattice' must be 19 or 27");
  if (!(omega_ > 0.0 && omega_ < 2.0)) throw std::runtime_error("omega must be in (0, 2)");
}
