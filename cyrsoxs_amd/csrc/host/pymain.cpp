// pybind11 module "CyRSoXS" with the reference's Python API (src/pymain.cpp:128-243,
// include/PyClass/*.h): InputData, RefractiveIndex, VoxelData, ScatteringPattern,
// Polarization, launch, cleanup, computePolarization and the enums.  Compute goes
// through the C ABI (include/cyrsoxs.h); there is no CPU path.
//
// Additions (no reference counterpart): InputData.precision (Precision.FP32/FP64,
// the reference's compile-time DOUBLE_PRECISION switch) and InputData.pipeline
// (Pipeline.Auto default = E-basis or material basis by cost, Pipeline.EBasis,
// Pipeline.MBasis = material-basis spectra, Pipeline.Plain = per-angle reference
// structure).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <bitset>
#include <cmath>
#include <complex>
#include <cstring>
#include <iostream>
#include <vector>

#include "../../../include/cyrsoxs.h"
#include "h5io.hpp"
#include "input_data.hpp"

namespace py = pybind11;
using namespace cyrhost;

namespace {

bool fequals(double a, double b) { return std::fabs(a - b) < 1E-6; }

enum ParamIdx { ENERGY = 0, DIMENSION, PHYSSIZE, EANGLE, KVECTORS, MORPHOLOGY_TYPE, CASE_TYPE, DETECTOR_COORD,
                NUM_MATERIAL, PARAM_MAX };
const char *param_names[] = {"ENERGY", "DIMENSION", "PHYSSIZE", "EANGLE", "KVectors", "MorphologyType",
                             "CaseType", "DetectorCoord", "NUM_MATERIAL"};

enum class Interp : int { NearestNeighour = 0, Linear = 1 };
enum class Windowing : int { NoPadding = 0, Hanning = 1 };
enum class Scatter : int { Partial = 0, Full = 1 };
enum class Morph : int { EulerAngles = 0, VectorMorphology = 1 };
enum class Case : int { Default = 0, BeamDivergence = 1, GrazingIncidence = 2 };
enum class Frame : int { Material = 0, Lab = 1 };
enum class Order : int { ZYX = 0, XYZ = 1 };
enum class Precision : int { FP32 = CYR_FP32, FP64 = CYR_FP64 };
enum class Pipeline : int { Plain = CYR_PIPE_PLAIN, EBasis = CYR_PIPE_EBASIS, MBasis = CYR_PIPE_MBASIS, Auto = CYR_PIPE_AUTO };

int to_int(const py::object &o) { return py::int_(o).cast<int>(); }

// InputData with the pybind setters of include/Input/InputData.h:430-615.
struct PyInputData {
  InputData d;
  std::bitset<PARAM_MAX> checker;
  explicit PyInputData(int nmat) {
    d.writeHDF5 = false;
    d.NUM_MATERIAL = nmat;
    checker.set(NUM_MATERIAL);
  }
  void setEnergies(std::vector<double> e) {
    std::sort(e.begin(), e.end());
    if (std::adjacent_find(e.begin(), e.end()) != e.end()) {
      py::print("energies has duplicate entries . Can not add");
      return;
    }
    d.energies = e;
    checker.set(ENERGY);
  }
  void setDimensions(const std::array<uint32_t, 3> &shape, Order order) {
    d.morphologyOrder = (int)order;
    if (order == Order::XYZ) {
      d.voxelDims[0] = shape[0];
      d.voxelDims[1] = shape[1];
      d.voxelDims[2] = shape[2];
    } else {
      d.voxelDims[2] = shape[0];
      d.voxelDims[1] = shape[1];
      d.voxelDims[0] = shape[2];
    }
    if (d.voxelDims[2] == 1) py::print("[INFO] 2D computation enabled");
    checker.set(DIMENSION);
  }
  void setPhysSize(double p) {
    d.physSize = p;
    checker.set(PHYSSIZE);
  }
  // pybind order is (start, end, increment) (include/Input/InputData.h:490)
  void setEAngles(double start, double end, double inc) {
    d.startAngle = start;
    d.endAngle = end;
    d.incrementAngle = inc;
    if (start > end) {
      py::print("[ERROR]: Start EAngle cannot be greater than end EAngle");
      return;
    }
    if (fequals(start, end)) {
      d.incrementAngle = 1.0;
    } else if (fequals(inc, 0.0)) {
      py::print("[ERROR] :  Increment angle for Electric field rotation cannot be 0");
      return;
    }
    checker.set(EANGLE);
  }
  void setMorphologyType(Morph m) {
    d.morphologyType = (uint32_t)m;
    checker.set(MORPHOLOGY_TYPE);
  }
  void setCaseType(Case c) {
    d.caseType = (uint32_t)c;
    checker.set(CASE_TYPE);
    if (c == Case::Default) {
      d.kVectors.assign(1, {0, 0, 1});
      checker.set(KVECTORS);
      checker.set(DETECTOR_COORD);
    }
  }
  void setKVectors(const std::array<double, 3> &k) {
    if (d.caseType == 0) {
      py::print("Cannot add kVectors for default case type");
      return;
    }
    const double n = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    d.kVectors.push_back({k[0] / n, k[1] / n, k[2] / n});
    checker.set(KVECTORS);
    if (d.caseType == 1) checker.set(DETECTOR_COORD);
  }
  void setDetectorCoordinates(const std::array<double, 3> &v) {
    if (d.caseType != 2) {
      py::print("Cannot add Detector for default or Beam Divergence type");
      return;
    }
    d.detectorCoordinates = v;
    checker.set(DETECTOR_COORD);
  }
  void setAlgorithm(int id, int streams) {
    if (id >= 2 || id < 0) {
      py::print("Incorrect AlgID.");
      return;
    }
    d.algorithmType = id;
    d.numMaxStreams = streams;
  }
  bool validate() const {
    if (!checker.all())
      for (int i = 0; i < PARAM_MAX; i++)
        if (!checker.test(i)) py::print("The value for ", param_names[i], "is not set");
    return checker.all();
  }
  void print() const {
    std::ostringstream os;
    d.print(os);
    py::print(os.str());
  }
};

struct PyRefractiveIndex {
  PyInputData *in;
  std::vector<Material> idx;
  std::vector<bool> valid;
  explicit PyRefractiveIndex(PyInputData &i) : in(&i) {
    idx.assign(i.d.energies.size() * (size_t)i.d.NUM_MATERIAL, Material{});
    valid.assign(i.d.energies.size(), false);
  }
  // values: [NUM_MATERIAL][4] = (DeltaPara, BetaPara, DeltaPerp, BetaPerp)
  void addData(const std::vector<std::vector<double>> &values, double energy) {
    const int M = in->d.NUM_MATERIAL;
    if ((int)values.size() != M) {
      py::print("Wrong input for Energy. Number not matching with number of Materials");
      return;
    }
    for (const auto &v : values)
      if (v.size() != 4) {
        py::print("Wrong number of input parameters. Parameters must be in the order of "
                  "(DeltaPara, BetaPara, DeltaPerp, , BetaPerp)");
        return;
      }
    const auto &e = in->d.energies;
    const size_t c = std::lower_bound(e.begin(), e.end(), energy) - e.begin();
    if (c >= e.size() || !fequals(e[c], energy)) {
      py::print("Bad Energy. Trying to add parameters for non-existent energy.");
      return;
    }
    if (idx.size() != e.size() * (size_t)M) {
      idx.assign(e.size() * (size_t)M, Material{});
      valid.assign(e.size(), false);
    }
    for (int i = 0; i < M; i++) {
      idx[c * M + i].npara = {1 - values[i][0], values[i][1]};
      idx[c * M + i].nperp = {1 - values[i][2], values[i][3]};
    }
    valid[c] = true;
  }
  bool validate() const {
    bool ok = !valid.empty() && std::all_of(valid.begin(), valid.end(), [](bool x) { return x; });
    if (!ok)
      for (size_t i = 0; i < valid.size(); i++)
        if (!valid[i]) py::print("Optical constants data missing for energy = ", in->d.energies[i], "eV");
    return ok;
  }
  void clear() { std::vector<Material>().swap(idx); }
  void print() const {
    const int M = in->d.NUM_MATERIAL;
    for (size_t i = 0; i < in->d.energies.size(); i++) {
      py::print("Energy = ", in->d.energies[i]);
      for (int m = 0; m < M && i * M + m < idx.size(); m++)
        py::print("Material = ", m, "npara = ", idx[i * M + m].npara, "nperp = ", idx[i * M + m].nperp);
    }
  }
};

using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;

struct PyVoxelData {
  PyInputData *in;
  std::vector<double> vox;  // [M][V][4]
  std::vector<bool> valid;
  explicit PyVoxelData(PyInputData &i) : in(&i) {
    vox.assign(i.d.numVoxels() * 4 * (size_t)i.d.NUM_MATERIAL, 0.0);
    valid.assign(i.d.NUM_MATERIAL, false);
  }
  uint64_t V() const { return in->d.numVoxels(); }
  bool check_mat(uint32_t id) {
    if (id > (uint32_t)in->d.NUM_MATERIAL || id == 0)
      throw std::logic_error(
          "Number of material does not match with the compiled version. matID must range from 1 to NUM_MATERIAL");
    if (valid[id - 1]) {
      py::print("The material is already set. Please first reset to add the entries. Returning.");
      return false;
    }
    if (vox.size() != V() * 4 * (size_t)in->d.NUM_MATERIAL) vox.assign(V() * 4 * (size_t)in->d.NUM_MATERIAL, 0.0);
    return true;
  }
  // XYZ-ordered numpy input -> internal ZYX (include/PyClass/VoxelData.h:101-104)
  std::vector<double> to_zyx(const DArr &a, int nc) {
    const uint64_t n = V() * nc;
    if ((uint64_t)a.size() != n) throw std::runtime_error("array size does not match the dimensions");
    std::vector<double> v(a.data(), a.data() + n);
    if (in->d.morphologyOrder == XYZ) {
      std::vector<double> o(n);
      const uint64_t X = in->d.voxelDims[0], Y = in->d.voxelDims[1], Z = in->d.voxelDims[2];
      for (uint64_t k = 0; k < Z; k++)
        for (uint64_t j = 0; j < Y; j++)
          for (uint64_t i = 0; i < X; i++)
            for (int c = 0; c < nc; c++) o[(k * X * Y + j * X + i) * nc + c] = v[(i * Y * Z + j * Z + k) * nc + c];
      v.swap(o);
    }
    return v;
  }
  void addVector(const DArr &aligned, const DArr &unaligned, uint32_t id) {
    if (in->d.morphologyType == 0) {
      py::print("Error: [Expected]: Vector Morphology [Found:] Euler Angles. Returning\n");
      return;
    }
    if (!check_mat(id)) return;
    const auto al = to_zyx(aligned, 3), un = to_zyx(unaligned, 1);
    double *vm = &vox[(uint64_t)(id - 1) * V() * 4];
    for (uint64_t i = 0; i < V(); i++) {
      vm[4 * i + 0] = al[3 * i + 0];
      vm[4 * i + 1] = al[3 * i + 1];
      vm[4 * i + 2] = al[3 * i + 2];
      vm[4 * i + 3] = un[i];
    }
    valid[id - 1] = true;
  }
  void addEuler(const DArr &S, const DArr &theta, const DArr &psi, const DArr &vfrac, uint32_t id) {
    if (in->d.morphologyType != 0) {
      py::print("Error: [Expected]: Euler Angles / Spherical Coordinates. [Found:] VectorMorphology\n");
      return;
    }
    if (!check_mat(id)) return;
    const auto s = to_zyx(S, 1), t = to_zyx(theta, 1), p = to_zyx(psi, 1), v = to_zyx(vfrac, 1);
    double *vm = &vox[(uint64_t)(id - 1) * V() * 4];
    for (uint64_t i = 0; i < V(); i++) {  // S == 0 zeroes the angles (include/PyClass/VoxelData.h:154-160)
      vm[4 * i + 0] = s[i];
      vm[4 * i + 1] = s[i] != 0 ? t[i] : 0.0;
      vm[4 * i + 2] = s[i] != 0 ? p[i] : 0.0;
      vm[4 * i + 3] = v[i];
    }
    valid[id - 1] = true;
  }
  void addVfrac(const DArr &vfrac, uint32_t id) {
    if (in->d.morphologyType != 0) {
      py::print("Error: [Expected]: Euler Angles / Spherical Coordinates. [Found:] VectorMorphology\n");
      return;
    }
    if (!check_mat(id)) return;
    const auto v = to_zyx(vfrac, 1);
    double *vm = &vox[(uint64_t)(id - 1) * V() * 4];
    for (uint64_t i = 0; i < V(); i++) {
      vm[4 * i + 0] = vm[4 * i + 1] = vm[4 * i + 2] = 0;
      vm[4 * i + 3] = v[i];
    }
    valid[id - 1] = true;
  }
  void readFromH5(const std::string &fname) {
    if (std::any_of(valid.begin(), valid.end(), [](bool b) { return b; })) {
      py::print("Some of the material is already set. Please first reset to continue. Returning.");
      return;
    }
    if (h5_number_of_materials(fname) != in->d.NUM_MATERIAL) {
      py::print("[HDF5 Error]: Number of material.");
      return;
    }
    uint32_t dims[3];
    double phys;
    int order;
    h5_dimension_and_order(fname, (int)in->d.morphologyType, dims, phys, order);
    if (order != in->d.morphologyOrder) {
      py::print("[HDF5 Error]: Morphology Order mismatch.");
      return;
    }
    if (!fequals(phys, in->d.physSize)) {
      py::print("[HDF5 Error]: PhysSize mismatch.");
      return;
    }
    h5_read_morphology(fname, in->d.voxelDims, (int)in->d.morphologyType, order, in->d.NUM_MATERIAL, vox);
    std::fill(valid.begin(), valid.end(), true);
  }
  void reset() { std::fill(valid.begin(), valid.end(), false); }
  void clear() { std::vector<double>().swap(vox); }
  bool validate() const {
    for (int i = 0; i < in->d.NUM_MATERIAL; i++)
      if (!valid[i]) {
        py::print("Voxel Data missing / corrupt for material = ", i + 1);
        return false;
      }
    if (!checkMorphology(vox.data(), vox.size())) {
      py::print("Nan Present in the morphology");
      return false;
    }
    return true;
  }
  void writeToH5() const {
    if (!validate()) return;
    h5_write_morphology("Morphology.h5", in->d, vox.data());
  }
};

struct PyScatteringPattern {
  PyInputData *in;
  std::vector<unsigned char> data;
  int precision;
  explicit PyScatteringPattern(PyInputData &i) : in(&i), precision(i.d.precision) {
    data.assign(i.d.energies.size() * i.d.kVectors.size() * (size_t)i.d.voxelDims[0] * i.d.voxelDims[1] *
                    (precision == CYR_FP64 ? 8 : 4),
                0);
  }
  size_t es() const { return precision == CYR_FP64 ? 8 : 4; }
  void clear() { std::vector<unsigned char>().swap(data); }
  void writeToHDF5(const std::string &dir) const {
    InputData d = in->d;
    d.precision = precision;
    h5_write_patterns(d, data.data(), dir);
  }
  void writeToVTI(const std::string &) const { py::print("Not supported"); }  // include/utils.h:95-96
  py::array view(std::vector<py::ssize_t> shape, size_t offset_bytes, py::handle base) const {
    std::vector<py::ssize_t> strides(shape.size());
    py::ssize_t s = (py::ssize_t)es();
    for (int i = (int)shape.size() - 1; i >= 0; i--) {
      strides[i] = s;
      s *= shape[i];
    }
    const void *p = data.data() + offset_bytes;
    if (precision == CYR_FP64) return py::array(py::dtype::of<double>(), shape, strides, p, base);
    return py::array(py::dtype::of<float>(), shape, strides, p, base);
  }
  // zero-copy (Ny, Nx) view of one energy / k (include/PyClass/ScatteringPattern.h:98-133)
  py::object dataToNumpy(py::object self, double energy, uint32_t kID) const {
    const auto &d = in->d;
    if (d.caseType == 0 && kID > 0) {
      py::print("kID cannot be greater than 0 for Default");
      return py::none();
    }
    if (kID >= d.kVectors.size()) {
      py::print("[ERROR] kID = ", kID, " must be smaller than kVector Size = ", d.kVectors.size());
      return py::none();
    }
    const auto &e = d.energies;
    if (e.empty() || energy < e.front() || energy > e.back()) {
      py::print("[ERROR]: Wrong EnergyID");
      return py::none();
    }
    const size_t id = std::lower_bound(e.begin(), e.end(), energy) - e.begin();
    if (!fequals(e[id], energy)) {
      py::print("[ERROR]: Wrong EnergyID");
      return py::none();
    }
    const size_t P = (size_t)d.voxelDims[0] * d.voxelDims[1];
    return view({(py::ssize_t)d.voxelDims[1], (py::ssize_t)d.voxelDims[0]}, (id * d.kVectors.size() + kID) * P * es(),
                self);
  }
  // (nE, Ny, Nx) view for k vector kID
  py::object writeAllToNumpy(py::object self, uint32_t kID) const {
    const auto &d = in->d;
    if (d.caseType == 0 && kID > 0) {
      py::print("kID cannot be greater than 0 for Default");
      return py::none();
    }
    if (kID >= d.kVectors.size()) {
      py::print("[ERROR] kID = ", kID, " must be smaller than kVector Size = ", d.kVectors.size());
      return py::none();
    }
    const size_t P = (size_t)d.voxelDims[0] * d.voxelDims[1];
    const size_t nK = d.kVectors.size();
    std::vector<py::ssize_t> shape = {(py::ssize_t)d.energies.size(), (py::ssize_t)d.voxelDims[1],
                                      (py::ssize_t)d.voxelDims[0]};
    std::vector<py::ssize_t> strides = {(py::ssize_t)(nK * P * es()), (py::ssize_t)(d.voxelDims[0] * es()),
                                        (py::ssize_t)es()};
    const void *p = data.data() + kID * P * es();
    if (precision == CYR_FP64) return py::array(py::dtype::of<double>(), shape, strides, p, self);
    return py::array(py::dtype::of<float>(), shape, strides, p, self);
  }
};

struct PyPolarization {
  PyInputData *in;
  int precision;
  std::vector<unsigned char> p[3];
  explicit PyPolarization(PyInputData &i) : in(&i), precision(i.d.precision) {
    const size_t b = i.d.numVoxels() * (precision == CYR_FP64 ? 16 : 8);
    for (auto &v : p) v.assign(b, 0);
  }
  void clear() {
    for (auto &v : p) std::vector<unsigned char>().swap(v);
  }
  void writeToHDF5() const {
    create_directory("Polarization");
    InputData d = in->d;
    d.precision = precision;
    const char *names[] = {"Polarization/PolarizationX.h5", "Polarization/PolarizationY.h5",
                           "Polarization/PolarizationZ.h5"};
    const char *ds[] = {"pX", "pY", "pZ"};
    for (int i = 0; i < 3; i++) h5_write_polarization(names[i], ds[i], d, p[i].data());
  }
  py::object writeToNumpy(py::object self, int id) const {
    if (id < 0 || id >= 3) {
      py::print("Wrong id. Must be less than 2.");
      return py::none();
    }
    const auto &d = in->d;
    std::vector<py::ssize_t> shape = {(py::ssize_t)d.voxelDims[2], (py::ssize_t)d.voxelDims[1],
                                      (py::ssize_t)d.voxelDims[0]};
    if (precision == CYR_FP64)
      return py::array_t<std::complex<double>>(shape, reinterpret_cast<const std::complex<double> *>(p[id].data()),
                                               self);
    return py::array_t<std::complex<float>>(shape, reinterpret_cast<const std::complex<float> *>(p[id].data()), self);
  }
};

void raise_status(cyr_status st) {
  if (st != CYR_OK)
    throw std::runtime_error(std::string("CyRSoXS GPU error: ") + cyr_status_string(st) + ": " + cyr_last_error());
}

// launch (src/pymain.cpp:78-119)
void launch(PyInputData &in, PyRefractiveIndex &ri, PyVoxelData &vd, PyScatteringPattern &sp, bool meta) {
  if (!in.validate()) {
    py::print("Issues with Input Data");
    return;
  }
  if (in.d.caseType != 0) py::print("This is an experimental feature which is not tested");
  if (!ri.validate()) {
    py::print("Issues with Energy data");
    return;
  }
  if (!vd.validate()) {
    py::print("Issues with Voxel Data input");
    return;
  }
  const size_t need = in.d.energies.size() * in.d.kVectors.size() * (size_t)in.d.voxelDims[0] * in.d.voxelDims[1] *
                      (in.d.precision == CYR_FP64 ? 8 : 4);
  if (sp.data.size() != need || sp.precision != in.d.precision) {
    py::print("ScatteringPattern does not match the InputData (create it after setting the inputs)");
    return;
  }
  cyr_status st;
  {
    py::gil_scoped_release release;
    const cyr_sim_desc d = in.d.desc();
    const std::vector<double> m = flatten_materials(ri.idx);
    if (in.d.precision == CYR_FP64) {
      st = cyr_simulate(&d, vd.vox.data(), m.data(), sp.data.data());
    } else {
      std::vector<float> vf(vd.vox.begin(), vd.vox.end());
      st = cyr_simulate(&d, vf.data(), m.data(), sp.data.data());
    }
    if (st == CYR_OK && meta) printMetaData(in.d);
  }
  raise_status(st);
}

void cleanup(PyRefractiveIndex &ri, PyVoxelData &vd, PyScatteringPattern &sp) {
  vd.clear();
  ri.clear();
  sp.clear();
}

// computePolarization (src/pymain.cpp:47-67)
void computePolarizationDebug(PyInputData &in, PyRefractiveIndex &ri, PyVoxelData &vd, PyPolarization &pol,
                              double energy, double eangle) {
  const auto &e = in.d.energies;
  if (e.empty() || energy < e.front() || energy > e.back()) {
    py::print("[ERROR]: Wrong Energy");
    return;
  }
  const size_t id = std::lower_bound(e.begin(), e.end(), energy) - e.begin();
  if (!fequals(e[id], energy)) {
    py::print("[ERROR]: Wrong EnergyID");
    return;
  }
  if (pol.precision != in.d.precision) {
    py::print("Polarization does not match the InputData precision");
    return;
  }
  cyr_status st;
  {
    py::gil_scoped_release release;
    const cyr_sim_desc d = in.d.desc();
    const std::vector<double> m = flatten_materials(ri.idx);
    if (in.d.precision == CYR_FP64) {
      st = cyr_polarization(&d, vd.vox.data(), m.data(), (uint32_t)id, eangle, pol.p[0].data(), pol.p[1].data(),
                            pol.p[2].data());
    } else {
      std::vector<float> vf(vd.vox.begin(), vd.vox.end());
      st = cyr_polarization(&d, vf.data(), m.data(), (uint32_t)id, eangle, pol.p[0].data(), pol.p[1].data(),
                            pol.p[2].data());
    }
  }
  raise_status(st);
}

}  // namespace

PYBIND11_MODULE(CyRSoXS, module) {
  module.doc() = "MI355X (HIP) CyRSoXS: Born-approximation RSoXS simulation";
  py::print("CyRSoXS (MI355X / gfx950 HIP build)");
  py::print("============================================================================");
  py::print("Library version            :", cyr_version());

  py::enum_<Interp>(module, "InterpolationType")
      .value("Linear", Interp::Linear)
      .value("NearestNeighour", Interp::NearestNeighour)
      .export_values();
  py::enum_<Windowing>(module, "FFTWindowing")
      .value("NoPadding", Windowing::NoPadding)
      .value("Hanning", Windowing::Hanning)
      .export_values();
  py::enum_<Scatter>(module, "ScatterApproach").value("Partial", Scatter::Partial).value("Full", Scatter::Full).export_values();
  py::enum_<Morph>(module, "MorphologyType")
      .value("EulerAngles", Morph::EulerAngles)
      .value("VectorMorphology", Morph::VectorMorphology)
      .export_values();
  py::enum_<Case>(module, "CaseType")
      .value("Default", Case::Default)
      .value("BeamDivergence", Case::BeamDivergence)
      .value("GrazingIncidence", Case::GrazingIncidence)
      .export_values();
  py::enum_<Frame>(module, "ReferenceFrame").value("Lab", Frame::Lab).value("Material", Frame::Material).export_values();
  py::enum_<Order>(module, "MorphologyOrder").value("XYZ", Order::XYZ).value("ZYX", Order::ZYX).export_values();
  py::enum_<Precision>(module, "Precision").value("FP32", Precision::FP32).value("FP64", Precision::FP64).export_values();
  py::enum_<Pipeline>(module, "Pipeline").value("Plain", Pipeline::Plain).value("EBasis", Pipeline::EBasis).value("MBasis", Pipeline::MBasis).value("Auto", Pipeline::Auto).export_values();

  py::class_<PyInputData>(module, "InputData")
      .def(py::init<int>(), "Constructor", py::arg("NumMaterial"))
      .def("setEnergies", &PyInputData::setEnergies, "Set the energy data", py::arg("energies"))
      .def("print", &PyInputData::print, "Print the input data")
      .def("setERotationAngle", &PyInputData::setEAngles, "Set the rotation for Electric field", py::arg("StartAngle"),
           py::arg("EndAngle"), py::arg("IncrementAngle"))
      .def("setPhysSize", &PyInputData::setPhysSize, "Set the Physical size (in nm)", py::arg("PhysSize"))
      .def("setDimensions", &PyInputData::setDimensions, "Set the Dimensions", py::arg("Dims"),
           py::arg("MorphologyOrder"))
      .def("validate", &PyInputData::validate, "Validate the input data")
      .def("setCaseType", &PyInputData::setCaseType, "case Type")
      .def("setMorphologyType", &PyInputData::setMorphologyType, "morphology Type")
      .def("setKVectors", &PyInputData::setKVectors, "set K vectors")
      .def("setDetectorCoordinates", &PyInputData::setDetectorCoordinates,
           "set Detector coordinates for Grazing incidence")
      .def("setAlgorithm", &PyInputData::setAlgorithm, "set Algorithm Type", py::arg("AlgorithmID"),
           py::arg("MaxStreams"))
      .def_property(
          "interpolationType", [](const PyInputData &s) { return s.d.ewaldsInterpolation; },
          [](PyInputData &s, py::object v) { s.d.ewaldsInterpolation = to_int(v); }, "Ewalds interpolation type")
      .def_property(
          "windowingType", [](const PyInputData &s) { return s.d.windowingType; },
          [](PyInputData &s, py::object v) { s.d.windowingType = (uint32_t)to_int(v); }, "Windowing type")
      .def_property(
          "rotMask", [](const PyInputData &s) { return s.d.rotMask; },
          [](PyInputData &s, py::object v) { s.d.rotMask = py::bool_(v); }, "Rotation Mask")
      .def_property(
          "openMP", [](const PyInputData &s) { return s.d.num_threads; },
          [](PyInputData &s, py::object v) { s.d.num_threads = (uint32_t)to_int(v); }, "number of OpenMP threads")
      .def_property(
          "scatterApproach", [](const PyInputData &s) { return s.d.scatterApproach; },
          [](PyInputData &s, py::object v) { s.d.scatterApproach = (uint32_t)to_int(v); }, "sets the scatter approach")
      .def_property(
          "referenceFrame", [](const PyInputData &s) { return s.d.referenceFrame; },
          [](PyInputData &s, py::object v) { s.d.referenceFrame = to_int(v) != 0; }, "sets the reference frame")
      .def_property(
          "precision", [](const PyInputData &s) { return s.d.precision; },
          [](PyInputData &s, py::object v) { s.d.precision = to_int(v) == CYR_FP64 ? CYR_FP64 : CYR_FP32; },
          "compute precision (Precision.FP32 / Precision.FP64)")
      .def_property(
          "pipeline", [](const PyInputData &s) { return s.d.pipeline; },
          [](PyInputData &s, py::object v) {
            const int p = to_int(v);
            s.d.pipeline = (p >= CYR_PIPE_PLAIN && p <= CYR_PIPE_AUTO) ? p : CYR_PIPE_AUTO;
          },
          "execution strategy (Pipeline.Auto default / EBasis / MBasis / Plain)");

  py::class_<PyRefractiveIndex>(module, "RefractiveIndex")
      .def(py::init<PyInputData &>(), "Constructor", py::keep_alive<1, 2>())
      .def("addData", &PyRefractiveIndex::addData, "Add Optical constants data ", py::arg("OpticalConstants"),
           py::arg("Energy"))
      .def("validate", &PyRefractiveIndex::validate, "Validate the refractive Index data")
      .def("clear", &PyRefractiveIndex::clear, "Clears the memory")
      .def("print", &PyRefractiveIndex::print, "Prints the refractive index data");

  py::class_<PyVoxelData>(module, "VoxelData")
      .def(py::init<PyInputData &>(), "Constructor", py::arg("InputData"), py::keep_alive<1, 2>())
      .def("addVoxelData", &PyVoxelData::addVector, "Adds the Alignment and unaligned component to the given material",
           py::arg("AlignedData"), py::arg("UnalignedData"), py::arg("MaterialID"))
      .def("addVoxelData", &PyVoxelData::addEuler, "Adds the EulerAngle components to the given material",
           py::arg("S"), py::arg("Theta"), py::arg("Psi"), py::arg("Vfrac"), py::arg("MaterialID"))
      .def("addVoxelData", &PyVoxelData::addVfrac,
           "Adds the EulerAngle components to the given material. S/Theta/Phi = 0", py::arg("Vfrac"),
           py::arg("MaterialID"))
      .def("clear", &PyVoxelData::clear, "Clears the voxel data")
      .def("validate", &PyVoxelData::validate, "validate the Voxel data")
      .def("readFromH5", &PyVoxelData::readFromH5, "Reads from HDF5", py::arg("Filename"))
      .def("reset", &PyVoxelData::reset, "Resets the voxel data")
      .def("writeToH5", &PyVoxelData::writeToH5, "Writes voxel data to HDF5 file")
      .def(
          "dataToNumpy",
          [](const PyVoxelData &v) {
            py::array_t<double> a({(py::ssize_t)v.in->d.NUM_MATERIAL, (py::ssize_t)v.V(), (py::ssize_t)4});
            std::memcpy(a.mutable_data(), v.vox.data(), v.vox.size() * sizeof(double));
            return a;
          },
          "Copy of the morphology as [NumMaterial][Nz*Ny*Nx][4] (X fastest)");

  py::class_<PyScatteringPattern>(module, "ScatteringPattern")
      .def(py::init<PyInputData &>(), "Constructor", py::arg("InputData"), py::keep_alive<1, 2>())
      .def("clear", &PyScatteringPattern::clear, "Clears the memory")
      .def("writeToHDF5", &PyScatteringPattern::writeToHDF5, "Dumps data in  HDF5 file format",
           py::arg("dirName") = "HDF5")
      .def("writeToVTI", &PyScatteringPattern::writeToVTI, "Dumps data in  VTI file format", py::arg("dirName") = "VTI")
      .def(
          "dataToNumpy",
          [](py::object self, double e, uint32_t k) { return self.cast<PyScatteringPattern &>().dataToNumpy(self, e, k); },
          "Returns data in numpy array", py::arg("Energy"), py::arg("kID") = 0)
      .def(
          "writeAllToNumpy",
          [](py::object self, uint32_t k) { return self.cast<PyScatteringPattern &>().writeAllToNumpy(self, k); },
          "Returns all data at once in numpy array", py::arg("kID") = 0);

  py::class_<PyPolarization>(module, "Polarization")
      .def(py::init<PyInputData &>(), "Constructor", py::arg("InputData"), py::keep_alive<1, 2>())
      .def("clear", &PyPolarization::clear, "Clears the memory")
      .def("writeToHDF5", &PyPolarization::writeToHDF5, "write to HDF5")
      .def(
          "writeToNumpy", [](py::object self, int id) { return self.cast<PyPolarization &>().writeToNumpy(self, id); },
          "returns numpy array", py::arg("id"));

  module.def("launch", &launch, "GPU computation", py::arg("InputData"), py::arg("RefractiveIndexData"),
             py::arg("VoxelData"), py::arg("ScatteringPattern"), py::arg("WriteMetaData") = true);
  module.def("cleanup", &cleanup, "Cleanup", py::arg("RefractiveIndex"), py::arg("VoxelData"),
             py::arg("ScatteringPattern"));
  module.def("computePolarization", &computePolarizationDebug, "Computes the polarization vector for debugging",
             py::arg("InputData"), py::arg("RefractiveIndex"), py::arg("VoxelData"), py::arg("Polarization"),
             py::arg("Energy"), py::arg("EAngle"));
  module.def("hdf5Available", &h5_available, "True if libhdf5 could be loaded");
  module.def(
      "readConfig",
      [](const std::string &path, int num_material, const std::string &material_dir) {
        InputData d;
        d.readConfig(path);
        d.NUM_MATERIAL = num_material;
        d.validateEnums();
        py::dict r;
        r["CaseType"] = d.caseType;
        r["Energies"] = d.energies;
        r["EAngleRotation"] = std::vector<double>{d.startAngle, d.incrementAngle, d.endAngle};
        r["NumAngles"] = d.numAngles();
        r["MorphologyType"] = d.morphologyType;
        r["RotMask"] = d.rotMask;
        r["EwaldsInterpolation"] = d.ewaldsInterpolation;
        r["WindowingType"] = d.windowingType;
        r["ScatterApproach"] = d.scatterApproach;
        r["Algorithm"] = d.algorithmType;
        r["ReferenceFrame"] = (int)d.referenceFrame;
        r["HDF5DirName"] = d.HDF5DirName;
        r["NumThreads"] = d.num_threads;
        r["kVectors"] = d.kVectors;
        r["DetectorCoordinates"] = d.detectorCoordinates;
        if (num_material > 0) {
          std::vector<Material> m;
          d.readRefractiveIndexData(m, material_dir);
          std::vector<std::vector<std::complex<double>>> mm;
          for (const auto &x : m) mm.push_back({x.npara, x.nperp});
          r["Materials"] = mm;
        }
        return r;
      },
      "Parses config.txt (and Material<N>.txt when num_material > 0) as the CLI does", py::arg("path"),
      py::arg("num_material") = 0, py::arg("material_dir") = ".");
  py::register_exception<InputError>(module, "InputError");
  py::register_exception<H5Error>(module, "HDF5Error");
}
