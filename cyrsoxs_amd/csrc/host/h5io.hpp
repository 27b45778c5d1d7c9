// HDF5 morphology reader and scattering-pattern writer (reference:
// include/Input/readH5.h:45-539, include/utils.h:38-87, include/Output/writeH5.h:40-68).
//
// libhdf5 / libhdf5_hl are loaded at run time (dlopen) from CYRSOXS_HDF5_DIR,
// /opt/conda/lib or the default search path, so neither the GPU library nor the
// Python module carries a link-time dependency on HDF5.  Every function throws
// H5Error when HDF5 is unavailable or the file is malformed.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "input_data.hpp"

namespace cyrhost {

struct H5Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

bool h5_available();

// /Morphology_Parameters/NumMaterial
int h5_number_of_materials(const std::string &file);
// PhysSize and dims/order from the first material's dataset labels.
void h5_dimension_and_order(const std::string &file, int morphology_type, uint32_t dims[3], double &phys_size,
                            int &order);
// Reads [M][V][4] (double) in ZYX (X-fastest) order; Euler S == 0 zeroes theta/psi.
void h5_read_morphology(const std::string &file, const uint32_t dims[3], int morphology_type, int order,
                        int num_materials, std::vector<double> &voxels);
// Writes <dir>/Energy_<E>.h5 per energy with KIDList/KVec and K<i>/projection.
void h5_write_patterns(const InputData &in, const void *patterns, const std::string &dir);
// Writes a morphology file in the reference's layout (used by tests and VoxelData.writeToH5).
void h5_write_morphology(const std::string &file, const InputData &in, const double *voxels);
// Writes one complex volume [Nz][Ny][Nx][2] (Polarization.writeToHDF5).
void h5_write_polarization(const std::string &file, const std::string &name, const InputData &in,
                           const void *data);

void create_directory(const std::string &dir);

}  // namespace cyrhost
