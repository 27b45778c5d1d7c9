// CyRSoXS command-line executable (reference: src/main.cpp:107-173).
//
//   CyRSoXS <morphology.h5> [HDF5OutputDir] [--precision fp32|fp64] [--pipeline auto|ebasis|mbasis|plain]
//
// Reads config.txt and Material<N>.txt from the working directory, the
// morphology from the HDF5 file, runs the pipeline on the current HIP device
// through the C ABI and writes HDF5/Energy_<E>.h5 plus CyRSoXS.log.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/cyrsoxs.h"
#include "h5io.hpp"
#include "input_data.hpp"

using namespace cyrhost;

int main(int argc, char **argv) {
  std::vector<std::string> pos;
  InputData in;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--precision" && i + 1 < argc) {
      std::string p = argv[++i];
      in.precision = (p == "fp64" || p == "double") ? CYR_FP64 : CYR_FP32;
    } else if (a == "--pipeline" && i + 1 < argc) {
      std::string p = argv[++i];
      in.pipeline = p == "plain" ? CYR_PIPE_PLAIN : p == "mbasis" ? CYR_PIPE_MBASIS : p == "ebasis" ? CYR_PIPE_EBASIS : CYR_PIPE_AUTO;
    } else {
      pos.push_back(a);
    }
  }
  if (pos.empty()) {
    std::cout << "Usage : " << argv[0] << " HDF5FileName HDF5OutputDirname [optional]"
              << " [--precision fp32|fp64] [--pipeline auto|ebasis|mbasis|plain]\n";
    return 1;
  }
  const std::string fname = pos[0];
  try {
    in.readConfig("config.txt");
    in.NUM_MATERIAL = h5_number_of_materials(fname);
    std::vector<Material> mats;
    in.readRefractiveIndexData(mats, ".");
    in.validateEnums();
    if (pos.size() > 1) in.HDF5DirName = pos[1];
    h5_dimension_and_order(fname, (int)in.morphologyType, in.voxelDims, in.physSize, in.morphologyOrder);
    if (in.if2DComputation()) std::cout << "[INFO] 2D computation enabled\n";
    in.print(std::cout);
    if (in.caseType != 0) std::cout << "This is an experimental feature which is not tested. \n";
    std::vector<double> vox;
    h5_read_morphology(fname, in.voxelDims, (int)in.morphologyType, in.morphologyOrder, in.NUM_MATERIAL, vox);
    if (in.dumpMorphology) h5_write_morphology("Morphology.h5", in, vox.data());
    if (!checkMorphology(vox.data(), vox.size())) throw std::runtime_error("Nan detected in the morphology");
    const cyr_sim_desc d = in.desc();
    const uint64_t P = (uint64_t)in.voxelDims[0] * in.voxelDims[1];
    const size_t n_out = in.energies.size() * in.kVectors.size() * P;
    const std::vector<double> m = flatten_materials(mats);
    std::vector<unsigned char> out(n_out * (in.precision == CYR_FP64 ? 8 : 4));
    cyr_status st;
    if (in.precision == CYR_FP64) {
      st = cyr_simulate(&d, vox.data(), m.data(), out.data());
    } else {
      std::vector<float> vf(vox.begin(), vox.end());
      st = cyr_simulate(&d, vf.data(), m.data(), out.data());
    }
    if (st != CYR_OK) {
      std::cerr << "[GPU error] " << cyr_status_string(st) << ": " << cyr_last_error() << "\n";
      return 2;
    }
    if (in.writeHDF5) h5_write_patterns(in, out.data(), in.HDF5DirName);
    printMetaData(in);
  } catch (const std::exception &e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
  std::cout << "Complete. Exiting \n";
  return 0;
}
