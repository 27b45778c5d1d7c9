// Host-side simulation inputs, mirroring the reference's InputData
// (include/Input/InputData.h:64-655) and Material (include/Input/Input.h:35-50).
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <ostream>
#include <string>
#include <vector>

#include "../../../include/cyrsoxs.h"

namespace cyrhost {

struct Material {
  std::complex<double> npara{0, 0};
  std::complex<double> nperp{0, 0};
};

enum MorphologyOrder : int { ZYX = 0, XYZ = 1, ORDER_INVALID = -1 };

struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct InputData {
  std::vector<double> energies;
  double startAngle = 0, endAngle = 0, incrementAngle = 1;
  uint32_t num_threads = 4;
  uint32_t voxelDims[3] = {0, 0, 0};  // X, Y, Z
  double physSize = 0;
  bool writeHDF5 = true;
  bool writeVTI = false;
  bool rotMask = false;
  int ewaldsInterpolation = 1;        // LINEAR (include/Input/InputData.h:207)
  uint32_t windowingType = 0;
  std::vector<std::array<double, 3>> kVectors;
  uint32_t scatterApproach = 0;
  std::string VTIDirName = "VTI";
  std::string HDF5DirName = "HDF5";
  uint32_t caseType = 0;
  uint32_t morphologyType = 0;
  bool referenceFrame = true;         // LAB
  std::array<double, 3> detectorCoordinates{0, 0, 1};
  int algorithmType = 0;
  int numMaxStreams = 1;
  int morphologyOrder = ORDER_INVALID;
  bool dumpMorphology = false;
  int NUM_MATERIAL = 0;
  // runtime switches that are compile-time in the reference
  int precision = CYR_FP32;
  int pipeline = CYR_PIPE_AUTO;

  bool if2DComputation() const { return voxelDims[2] == 1; }
  uint64_t numVoxels() const { return (uint64_t)voxelDims[0] * voxelDims[1] * voxelDims[2]; }
  uint32_t numAngles() const;

  // config.txt constructor semantics (include/Input/InputData.h:258-336); throws InputError.
  void readConfig(const std::string &filename);
  // Material<m+1>.txt in `dir` for every energy (include/Input/InputData.h:338-379).
  void readRefractiveIndexData(std::vector<Material> &out, const std::string &dir = ".") const;
  // enum range checks (include/Input/InputData.h:380-390); throws InputError
  void validateEnums() const;
  void print(std::ostream &os) const;
  void printToFile(std::ostream &os) const;

  // The C-ABI descriptor (arrays it points to live in this object).
  cyr_sim_desc desc() const;

 private:
  mutable std::vector<double> kflat_;
};

// Fills cyr materials (double[nE][M][4]) from Material structs ([nE][M]).
std::vector<double> flatten_materials(const std::vector<Material> &m);

// Writes CyRSoXS.log (include/utils.h:207-233).
void printMetaData(const InputData &in, const std::string &path = "CyRSoXS.log");

// NaN scan of a morphology (include/utils.h:236-246).
bool checkMorphology(const double *voxels, uint64_t n);

}  // namespace cyrhost
