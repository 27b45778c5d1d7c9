// Host-side InputData: config.txt / Material<N>.txt reading, validation, logging.
#include "input_data.hpp"

#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "../host_math.hpp"
#include "config.hpp"

namespace cyrhost {

namespace {

bool fequals(double a, double b) { return std::fabs(a - b) < 1E-6; }  // include/Datatypes.h:143

const char *interp_name[] = {"Nearest Neighbour", "Trilinear interpolation"};
const char *window_name[] = {"NONE", "HANNING"};
const char *case_name[] = {"Default", "BeamDivergence", "GrazingIncidence"};
const char *morph_name[] = {"EulerAngles", "VectorMorphology"};
const char *order_name[] = {"ZYX", "XYZ"};
const char *frame_name[] = {"MATERIAL", "LAB"};
const char *scatter_name[] = {"Partial", "Full"};
const char *alg_name[] = {"CommunicationMinimizing", "MemoryMinimizing"};

template <typename T> bool read_opt(const Config &cfg, const std::string &key, T &v);

template <> bool read_opt<double>(const Config &cfg, const std::string &key, double &v) {
  const Setting *s = cfg.lookup(key);
  if (!s) return false;
  v = s->as_double();
  return true;
}
template <> bool read_opt<int>(const Config &cfg, const std::string &key, int &v) {
  const Setting *s = cfg.lookup(key);
  if (!s) return false;
  v = (int)s->as_int();
  return true;
}
template <> bool read_opt<uint32_t>(const Config &cfg, const std::string &key, uint32_t &v) {
  const Setting *s = cfg.lookup(key);
  if (!s) return false;
  const int64_t x = s->as_int();
  if (x < 0) throw InputError("[Input Error] " + key + " must be non-negative");
  v = (uint32_t)x;
  return true;
}
template <> bool read_opt<bool>(const Config &cfg, const std::string &key, bool &v) {
  const Setting *s = cfg.lookup(key);
  if (!s) return false;
  v = s->as_bool();
  return true;
}
template <> bool read_opt<std::string>(const Config &cfg, const std::string &key, std::string &v) {
  const Setting *s = cfg.lookup(key);
  if (!s) return false;
  if (s->kind != Setting::STRING) throw InputError("[Input Error] " + key + " must be a string");
  v = s->s;
  return true;
}

template <typename T> void read_req(const Config &cfg, const std::string &key, T &v) {
  if (!read_opt<T>(cfg, key, v))
    throw InputError("[Input Error] No value corresponding to " + key + " found. Exiting");
}

std::vector<double> read_array(const Config &cfg, const std::string &key, size_t size = 0) {
  const Setting *s = cfg.lookup(key);
  if (!s) throw InputError("[Input Error] No value corresponding to " + key + " found. Exiting");
  if (s->kind != Setting::ARRAY && s->kind != Setting::LIST)
    throw InputError("expected array input but found single value (key = " + key + ")");
  if (size && s->size() != size)
    throw InputError("[Input Error]: The array corresponding to " + key + " must be of length " +
                     std::to_string(size) + ". But found of length " + std::to_string(s->size()));
  std::vector<double> out;
  for (size_t i = 0; i < s->size(); i++) out.push_back(s->at(i).as_double());
  return out;
}

std::array<double, 3> normalised(const std::vector<double> &v) {
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (!(n > 0)) throw InputError("[Input Error] zero-length vector");
  return {v[0] / n, v[1] / n, v[2] / n};
}

}  // namespace

uint32_t InputData::numAngles() const {
  return (uint32_t)std::round((endAngle - startAngle) / incrementAngle + 1);
}

void InputData::readConfig(const std::string &filename) {
  Config cfg;
  try {
    cfg.read_file(filename);
  } catch (const ConfigError &e) {
    throw InputError(e.what());
  }
  try {
    read_req(cfg, "CaseType", caseType);
    energies = read_array(cfg, "Energies");
    const std::vector<double> t = read_array(cfg, "EAngleRotation", 3);
    // config order is [start, increment, end] (include/Input/InputData.h:274-276)
    startAngle = t[0];
    incrementAngle = t[1];
    endAngle = t[2];
    if (startAngle > endAngle) throw InputError("Start EAngle can not be greater than end EAngle");
    if (incrementAngle == 0) {
      if (!fequals(startAngle, endAngle)) throw InputError("Increment angle = 0");
      incrementAngle = 1.0;
    }
    read_req(cfg, "MorphologyType", morphologyType);
    read_opt(cfg, "NumThreads", num_threads);
    read_opt(cfg, "RotMask", rotMask);
    read_opt(cfg, "EwaldsInterpolation", ewaldsInterpolation);
    read_opt(cfg, "HDF5DirName", HDF5DirName);
    read_opt(cfg, "WindowingType", windowingType);
    if (!read_opt(cfg, "Algorithm", algorithmType)) read_opt(cfg, "AlgorithmType", algorithmType);
    read_opt(cfg, "ScatterApproach", scatterApproach);
    read_opt(cfg, "DumpMorphology", dumpMorphology);
    read_opt(cfg, "MaxStreams", numMaxStreams);
    uint32_t rf;
    if (read_opt(cfg, "ReferenceFrame", rf)) {
      if (rf >= 2) throw InputError("[Error] : Reference Frame Wrong value. Max acceptable value 1");
      referenceFrame = rf != 0;
    }
    kVectors.clear();
    if (caseType == 0) {
      kVectors.push_back({0, 0, 1});
    } else {
      const Setting *lk = cfg.lookup("listKVectors");
      if (!lk) throw InputError("[Input Error] No value corresponding to listKVectors found. Exiting");
      for (size_t i = 0; i < lk->size(); i++) {
        const Setting *k = lk->at(i).find("k");
        if (!k || k->size() != 3) throw InputError("[Input Error] listKVectors entries need k = [x, y, z]");
        kVectors.push_back(normalised({k->at(0).as_double(), k->at(1).as_double(), k->at(2).as_double()}));
      }
      if (kVectors.empty()) throw InputError("[Input Error] listKVectors is empty");
    }
    if (caseType == 2) detectorCoordinates = normalised(read_array(cfg, "DetectorCoordinates", 3));
  } catch (const ConfigError &e) {
    throw InputError(std::string("[Input Error] ") + e.what());
  }
}

void InputData::readRefractiveIndexData(std::vector<Material> &out, const std::string &dir) const {
  out.assign(energies.size() * (size_t)NUM_MATERIAL, Material{});
  for (int m = 0; m < NUM_MATERIAL; m++) {
    const std::string fname = dir + "/Material" + std::to_string(m + 1) + ".txt";
    Config cfg;
    try {
      cfg.read_file(fname);
    } catch (const ConfigError &e) {
      throw InputError(e.what());
    }
    for (size_t i = 0; i < energies.size(); i++) {
      const std::string g = "EnergyData" + std::to_string(i);
      const Setting *grp = cfg.lookup(g);
      if (!grp) throw InputError("[Input Error] No energy found for " + std::to_string(energies[i]));
      auto get = [&](const char *k) {
        const Setting *s = grp->find(k);
        if (!s) throw InputError(std::string("[Input Error] ") + g + "." + k + " missing in " + fname);
        return s->as_double();
      };
      const double e = get("Energy");
      if (std::fabs(e - energies[i]) > 1E-3)  // include/Input/InputData.h:356-364
        throw InputError("[Input Error] No energy found for " + std::to_string(energies[i]));
      Material &mt = out[i * NUM_MATERIAL + m];
      mt.npara = {1 - get("DeltaPara"), get("BetaPara")};
      mt.nperp = {1 - get("DeltaPerp"), get("BetaPerp")};
    }
  }
}

void InputData::validateEnums() const {
  auto chk = [](const char *name, long v, long max) {
    if (v < 0 || v >= max)
      throw InputError(std::string("[Error] : ") + name + " Wrong value. Max acceptable value " +
                       std::to_string(max - 1) + ". Value found = " + std::to_string(v));
  };
  chk("FFT Windowing", windowingType, 2);
  chk("Scatter Approach", scatterApproach, 2);
  chk("Ewalds Interpolation", ewaldsInterpolation, 2);
  chk("Morphology Type", morphologyType, 2);
  chk("Case Type", caseType, 3);
  chk("Algorithm", algorithmType, 2);
}

cyr_sim_desc InputData::desc() const {
  cyr_sim_desc d{};
  d.dims[0] = voxelDims[0];
  d.dims[1] = voxelDims[1];
  d.dims[2] = voxelDims[2];
  d.num_materials = (uint32_t)NUM_MATERIAL;
  d.phys_size = physSize;
  d.num_energies = (uint32_t)energies.size();
  d.energies = energies.data();
  d.angle_start = startAngle;
  d.angle_increment = incrementAngle;
  d.angle_end = endAngle;
  kflat_.clear();
  for (const auto &k : kVectors) kflat_.insert(kflat_.end(), k.begin(), k.end());
  d.num_kvectors = (uint32_t)kVectors.size();
  d.kvectors = kflat_.empty() ? nullptr : kflat_.data();
  d.detector[0] = detectorCoordinates[0];
  d.detector[1] = detectorCoordinates[1];
  d.detector[2] = detectorCoordinates[2];
  d.case_type = (int32_t)caseType;
  d.morphology_type = (int32_t)morphologyType;
  d.reference_frame = referenceFrame ? 1 : 0;
  d.interpolation = ewaldsInterpolation;
  d.windowing = (int32_t)windowingType;
  d.rot_mask = rotMask ? 1 : 0;
  d.scatter_approach = (int32_t)scatterApproach;
  d.algorithm = algorithmType;
  d.precision = precision;
  d.storage = CYR_STORE_NATIVE;
  d.pipeline = pipeline;
  return d;
}

void InputData::print(std::ostream &os) const {
  os << "NumMaterial          : " << NUM_MATERIAL << "\n";
  if (morphologyOrder == XYZ)
    os << "Dimensions [X Y Z]   : [" << voxelDims[0] << " " << voxelDims[1] << " " << voxelDims[2] << "]\n";
  else
    os << "Dimensions [Z Y X]   : [" << voxelDims[2] << " " << voxelDims[1] << " " << voxelDims[0] << "]\n";
  os << "PhysSize             : " << physSize << " nm \n";
  os << "E Rotation Angle     : " << startAngle << " : " << incrementAngle << " : " << endAngle << "\n";
  os << "Morphology Type      : " << morph_name[morphologyType % 2] << "\n";
  if (morphologyOrder == ZYX || morphologyOrder == XYZ)
    os << "Morphology Order     : " << order_name[morphologyOrder] << "\n";
  os << "Energies simulated   : [";
  for (double e : energies) os << e << " ";
  os << "]\n";
  os << "Windowing Type       : " << window_name[windowingType % 2] << "\n";
  os << "Rotation Mask        : " << rotMask << "\n";
  os << "Interpolation Type   : " << interp_name[ewaldsInterpolation % 2] << "\n";
  os << "HDF Output Directory : " << HDF5DirName << "\n";
  os << "Scatter Approach     : " << scatter_name[scatterApproach % 2] << "\n";
  os << "Algorithm            : " << alg_name[algorithmType % 2] << "\n";
  os << "Reference Frame      : " << frame_name[referenceFrame ? 1 : 0] << "(" << referenceFrame << ")\n";
  os << "Precision            : " << (precision == CYR_FP64 ? "FP64" : "FP32") << "\n";
  os << "Pipeline             : " << (pipeline == CYR_PIPE_EBASIS ? "E-basis" : pipeline == CYR_PIPE_MBASIS ? "material basis" : pipeline == CYR_PIPE_AUTO ? "auto (E-basis or material basis)" : "plain") << "\n";
}

void InputData::printToFile(std::ostream &os) const {
  os << "CaseType             : " << case_name[caseType % 3] << "\n";
  os << "Morphology Type      : " << morph_name[morphologyType % 2] << "\n";
  if (morphologyOrder == ZYX || morphologyOrder == XYZ)
    os << "Morphology Order     : " << order_name[morphologyOrder] << "\n";
  os << "Reference Frame      : " << frame_name[referenceFrame ? 1 : 0] << "(" << referenceFrame << ")\n";
  if (morphologyOrder == XYZ)
    os << "Dimensions [X Y Z]   : [" << voxelDims[0] << " " << voxelDims[1] << " " << voxelDims[2] << "]\n";
  else
    os << "Dimensions [Z Y X]   : [" << voxelDims[2] << " " << voxelDims[1] << " " << voxelDims[0] << "]\n";
  os << "PhysSize             : " << physSize << "nm \n";
  os << "E Rotation Angle     : " << startAngle << " : " << incrementAngle << " : " << endAngle << "\n";
  os << "Energies simulated   : [";
  for (double e : energies) os << e << " ";
  os << "]\n";
  os << "Windowing Type       : " << window_name[windowingType % 2] << "\n";
  os << "Rotation Mask        : " << rotMask << "\n";
  os << "Interpolation Type   : " << interp_name[ewaldsInterpolation % 2] << "\n";
  os << "kVectors             : (";
  for (const auto &k : kVectors) os << "[" << k[0] << "," << k[1] << "," << k[2] << "]";
  os << ")\n";
  os << "HDF Output Directory : " << HDF5DirName << "\n";
  os << "Scatter Approach     : " << scatter_name[scatterApproach % 2] << "\n";
  os << "Algorithm            : " << alg_name[algorithmType % 2] << "\n";
  if (algorithmType == 1) os << "MaxStreams           : " << numMaxStreams << "\n";
}

std::vector<double> flatten_materials(const std::vector<Material> &m) {
  std::vector<double> out(4 * m.size());
  for (size_t i = 0; i < m.size(); i++) {
    out[4 * i + 0] = m[i].npara.real();
    out[4 * i + 1] = m[i].npara.imag();
    out[4 * i + 2] = m[i].nperp.real();
    out[4 * i + 3] = m[i].nperp.imag();
  }
  return out;
}

void printMetaData(const InputData &in, const std::string &path) {
  std::ofstream f(path);
  f << "CyRSoXS (MI355X HIP build): " << cyr_version() << "\n";
  f << "=========================================================================================\n";
  f << "Size of Real         : " << (in.precision == CYR_FP64 ? 8 : 4) << "\n";
  f << "Number of material   : " << in.NUM_MATERIAL << "\n";
  f << "\nScaling Information:\n";
  f << "=========================================================================================\n";
  f << "Number of pixel      :[" << in.voxelDims[0] << "," << in.voxelDims[1] << "]\n";
  f << "Q range              :[" << -M_PI / in.physSize << "," << M_PI / in.physSize << "]\n";
  f << "\n\nInputData : \n";
  f << "=========================================================================================\n";
  in.printToFile(f);
  f << "\n\nRotation Matrices : \n";
  f << "=========================================================================================\n";
  const cyr::M3 D = cyr::rotation_matrix_k(cyr::V3{in.detectorCoordinates[0], in.detectorCoordinates[1],
                                                   in.detectorCoordinates[2]});
  auto pm = [&](const cyr::M3 &M) {
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) f << M.m[3 * i + j] << " ";
      f << "\n";
    }
  };
  f << "Detector Rotation Matrix\n-------------------------------\n";
  pm(in.caseType == 2 ? D : cyr::rotation_matrix_k(cyr::V3{0, 0, 1}));
  f << "-------------------------------\nK Rotation Matrix And Base Configuration \n";
  f << "-----------------------------------------\n";
  for (const auto &k : in.kVectors) {
    const cyr::V3 kv{k[0], k[1], k[2]};
    const cyr::M3 RK = cyr::rotation_matrix_k(kv);
    cyr::M3 R;
    const double ang = cyr::base_configuration(kv, RK, &R);
    f << "-----------------------------------------\n";
    f << "K = " << k[0] << " " << k[1] << " " << k[2] << "\n";
    pm(RK);
    f << "RotAngle = " << ang << "\n";
    for (int c = 0; c < 3; c++) {
      const char *nm[] = {"Base X   = ", "Base Y   = ", "Base Z   = "};
      f << nm[c] << R.m[c] << " " << R.m[3 + c] << " " << R.m[6 + c] << "\n";
    }
    f << "-----------------------------------------\n";
  }
  f << "-------------------------------\n";
}

bool checkMorphology(const double *voxels, uint64_t n) {
  for (uint64_t i = 0; i < n; i++)
    if (std::isnan(voxels[i])) return false;
  return true;
}

}  // namespace cyrhost
