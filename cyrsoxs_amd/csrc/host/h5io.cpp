// HDF5 I/O over a dlopen'ed libhdf5 C API (1.10 ABI: hid_t is int64_t).
#include "h5io.hpp"

#include <dlfcn.h>
#include <sys/stat.h>

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <mutex>
#include <sstream>

namespace cyrhost {

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;

constexpr unsigned F_ACC_RDONLY = 0x0000u;
constexpr unsigned F_ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0;
constexpr hid_t S_ALL = 0;
constexpr int S_SCALAR = 0;
constexpr int T_INTEGER = 0, T_FLOAT = 1;

struct H5Api {
  bool ok = false;
  std::string why;
  herr_t (*open)(void);
  hid_t (*Fopen)(const char *, unsigned, hid_t);
  hid_t (*Fcreate)(const char *, unsigned, hid_t, hid_t);
  herr_t (*Fclose)(hid_t);
  htri_t (*Lexists)(hid_t, const char *, hid_t);
  hid_t (*Gopen2)(hid_t, const char *, hid_t);
  hid_t (*Gcreate2)(hid_t, const char *, hid_t, hid_t, hid_t);
  herr_t (*Gclose)(hid_t);
  hid_t (*Dopen2)(hid_t, const char *, hid_t);
  hid_t (*Dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*Dclose)(hid_t);
  hid_t (*Dget_space)(hid_t);
  hid_t (*Dget_type)(hid_t);
  int (*Sget_simple_extent_ndims)(hid_t);
  int (*Sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
  hid_t (*Screate_simple)(int, const hsize_t *, const hsize_t *);
  hid_t (*Screate)(int);
  herr_t (*Sclose)(hid_t);
  int (*Tget_class)(hid_t);
  herr_t (*Tclose)(hid_t);
  herr_t (*Eset_auto2)(hid_t, void *, void *);
  herr_t (*DSset_label)(hid_t, unsigned, const char *);
  long (*DSget_label)(hid_t, unsigned, char *, size_t);
  hid_t *NATIVE_DOUBLE, *NATIVE_FLOAT, *NATIVE_INT, *NATIVE_LLONG;
};

H5Api &api() {
  static H5Api a;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<std::string> dirs;
    if (const char *e = std::getenv("CYRSOXS_HDF5_DIR")) dirs.push_back(std::string(e) + "/");
    dirs.push_back("/opt/conda/lib/");
    dirs.push_back("");
    void *h = nullptr, *hl = nullptr;
    for (const auto &d : dirs) {
      h = dlopen((d + "libhdf5.so").c_str(), RTLD_NOW | RTLD_GLOBAL);
      if (h) {
        hl = dlopen((d + "libhdf5_hl.so").c_str(), RTLD_NOW | RTLD_GLOBAL);
        break;
      }
    }
    if (!h || !hl) {
      a.why = "libhdf5/libhdf5_hl not found (set CYRSOXS_HDF5_DIR)";
      return;
    }
    bool okall = true;
    auto sym = [&](void *lib, const char *n) {
      void *p = dlsym(lib, n);
      if (!p) okall = false;
      return p;
    };
#define CYR_SYM(field, lib, name) a.field = reinterpret_cast<decltype(a.field)>(sym(lib, name))
    CYR_SYM(open, h, "H5open");
    CYR_SYM(Fopen, h, "H5Fopen");
    CYR_SYM(Fcreate, h, "H5Fcreate");
    CYR_SYM(Fclose, h, "H5Fclose");
    CYR_SYM(Lexists, h, "H5Lexists");
    CYR_SYM(Gopen2, h, "H5Gopen2");
    CYR_SYM(Gcreate2, h, "H5Gcreate2");
    CYR_SYM(Gclose, h, "H5Gclose");
    CYR_SYM(Dopen2, h, "H5Dopen2");
    CYR_SYM(Dcreate2, h, "H5Dcreate2");
    CYR_SYM(Dread, h, "H5Dread");
    CYR_SYM(Dwrite, h, "H5Dwrite");
    CYR_SYM(Dclose, h, "H5Dclose");
    CYR_SYM(Dget_space, h, "H5Dget_space");
    CYR_SYM(Dget_type, h, "H5Dget_type");
    CYR_SYM(Sget_simple_extent_ndims, h, "H5Sget_simple_extent_ndims");
    CYR_SYM(Sget_simple_extent_dims, h, "H5Sget_simple_extent_dims");
    CYR_SYM(Screate_simple, h, "H5Screate_simple");
    CYR_SYM(Screate, h, "H5Screate");
    CYR_SYM(Sclose, h, "H5Sclose");
    CYR_SYM(Tget_class, h, "H5Tget_class");
    CYR_SYM(Tclose, h, "H5Tclose");
    CYR_SYM(Eset_auto2, h, "H5Eset_auto2");
    CYR_SYM(DSset_label, hl, "H5DSset_label");
    CYR_SYM(DSget_label, hl, "H5DSget_label");
    CYR_SYM(NATIVE_DOUBLE, h, "H5T_NATIVE_DOUBLE_g");
    CYR_SYM(NATIVE_FLOAT, h, "H5T_NATIVE_FLOAT_g");
    CYR_SYM(NATIVE_INT, h, "H5T_NATIVE_INT_g");
    CYR_SYM(NATIVE_LLONG, h, "H5T_NATIVE_LLONG_g");
#undef CYR_SYM
    if (!okall) {
      a.why = "libhdf5 is missing expected symbols";
      return;
    }
    if (a.open() < 0) {
      a.why = "H5open failed";
      return;
    }
    a.Eset_auto2(0 /* H5E_DEFAULT */, nullptr, nullptr);  // errors are reported by return codes
    a.ok = true;
  });
  if (!a.ok) throw H5Error("[HDF5 Error] " + a.why);
  return a;
}

// RAII handle closers
struct Hid {
  hid_t id = -1;
  herr_t (*closer)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
  ~Hid() {
    if (id >= 0 && closer) closer(id);
  }
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
};

hid_t check(hid_t id, const std::string &what) {
  if (id < 0) throw H5Error("[HDF5 Error] " + what);
  return id;
}

bool exists(hid_t loc, const std::string &name) { return api().Lexists(loc, name.c_str(), P_DEFAULT) > 0; }

std::string label(hid_t ds, unsigned idx) {
  char buf[2] = {0, 0};  // AXIS_LABEL_LEN = 2 (include/Input/readH5.h:40)
  if (api().DSget_label(ds, idx, buf, 2) < 0) return "";
  return std::string(buf);
}

// XYZ (Z fastest) -> ZYX (X fastest) for `nc` interleaved components
// (include/Input/readH5.h:164-184).
template <typename T>
void xyz_to_zyx(std::vector<T> &data, int nc, const uint32_t *dims) {
  std::vector<T> out(data.size());
  const uint64_t X = dims[0], Y = dims[1], Z = dims[2];
  for (uint64_t k = 0; k < Z; k++)
    for (uint64_t j = 0; j < Y; j++)
      for (uint64_t i = 0; i < X; i++) {
        const uint64_t zyx = k * (X * Y) + j * X + i, xyz = i * (Y * Z) + j * Z + k;
        for (int c = 0; c < nc; c++) out[zyx * nc + c] = data[xyz * nc + c];
      }
  data.swap(out);
}

// Reads a float/double dataset of `count` elements as double.
void read_doubles(hid_t ds, std::vector<double> &out, size_t count) {
  H5Api &a = api();
  out.resize(count);
  Hid t(a.Dget_type(ds), a.Tclose);
  if (a.Tget_class(t.id) != T_FLOAT && a.Tget_class(t.id) != T_INTEGER)
    throw H5Error("[HDF5 Error] This data format is not supported");
  if (a.Dread(ds, *a.NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, out.data()) < 0)
    throw H5Error("[HDF5 Error] read failed");
}

void check_labels_and_dims(hid_t ds, const std::string &name, const uint32_t *dims, int order, int rank,
                           int last) {
  H5Api &a = api();
  Hid sp(a.Dget_space(ds), a.Sclose);
  const int nd = a.Sget_simple_extent_ndims(sp.id);
  if (nd != rank)
    throw H5Error("[HDF5 Error] Expected " + std::to_string(rank) + "D array. Found Dim = " + std::to_string(nd) +
                  " for " + name);
  hsize_t d[4] = {0, 0, 0, 0};
  a.Sget_simple_extent_dims(sp.id, d, nullptr);
  const std::string l0 = label(ds, 0), l2 = label(ds, 2);
  bool ok;
  if (order == ZYX) {
    if (!(l0 == "Z" && l2 == "X")) throw H5Error("[HDF5 Error] Axis label mismatch for morphology for " + name);
    ok = d[0] == dims[2] && d[1] == dims[1] && d[2] == dims[0];
  } else {
    if (!(l0 == "X" && l2 == "Z")) throw H5Error("[HDF5 Error] Axis label mismatch for morphology for " + name);
    ok = d[0] == dims[0] && d[1] == dims[1] && d[2] == dims[2];
  }
  if (rank == 4) ok = ok && d[3] == (hsize_t)last;
  if (!ok) throw H5Error("Dimension mismatch for morphology (" + name + ")");
}

// Returns false if the dataset is absent and !required (data zero-filled).
bool get_scalar(hid_t file, const std::string &group, const std::string &suffix, const uint32_t *dims, int order,
                int mat, bool required, std::vector<double> &data) {
  H5Api &a = api();
  const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2];
  if (!exists(file, group)) throw H5Error("[HDF5 Error] Group " + group + " not found");
  Hid g(check(a.Gopen2(file, group.c_str(), P_DEFAULT), "open group " + group), a.Gclose);
  const std::string name = "Mat_" + std::to_string(mat) + suffix;
  if (!exists(g.id, name)) {
    if (required) throw H5Error("[HDF5 Error] Dataset = " + name + " does not exist");
    data.assign(V, 0.0);
    return false;
  }
  Hid ds(check(a.Dopen2(g.id, name.c_str(), P_DEFAULT), "open " + name), a.Dclose);
  check_labels_and_dims(ds.id, name, dims, order, 3, 0);
  read_doubles(ds.id, data, V);
  if (order == XYZ) xyz_to_zyx(data, 1, dims);
  return true;
}

void write_labels(hid_t ds, std::initializer_list<const char *> names) {
  unsigned i = 0;
  for (const char *n : names) api().DSset_label(ds, i++, n);
}

}  // namespace

bool h5_available() {
  try {
    api();
    return true;
  } catch (const H5Error &) {
    return false;
  }
}

int h5_number_of_materials(const std::string &file) {
  H5Api &a = api();
  Hid f(check(a.Fopen(file.c_str(), F_ACC_RDONLY, P_DEFAULT), "cannot open " + file), a.Fclose);
  if (!exists(f.id, "Morphology_Parameters")) throw H5Error("[HDF5 Error] Group Morphology_Parameters not found");
  Hid g(a.Gopen2(f.id, "Morphology_Parameters", P_DEFAULT), a.Gclose);
  if (!exists(g.id, "NumMaterial")) throw H5Error("[HDF5 Error] DataSet NumMaterial not found");
  Hid ds(a.Dopen2(g.id, "NumMaterial", P_DEFAULT), a.Dclose);
  int n = 0;
  if (a.Dread(ds.id, *a.NATIVE_INT, S_ALL, S_ALL, P_DEFAULT, &n) < 0) throw H5Error("[HDF5 Error] NumMaterial");
  return n;
}

void h5_dimension_and_order(const std::string &file, int morphology_type, uint32_t dims[3], double &phys_size,
                            int &order) {
  H5Api &a = api();
  Hid f(check(a.Fopen(file.c_str(), F_ACC_RDONLY, P_DEFAULT), "cannot open " + file), a.Fclose);
  {
    if (!exists(f.id, "Morphology_Parameters")) throw H5Error("[HDF5 Error] Group Morphology_Parameters not found");
    Hid g(a.Gopen2(f.id, "Morphology_Parameters", P_DEFAULT), a.Gclose);
    if (!exists(g.id, "PhysSize")) throw H5Error("[HDF5 Error] DataSet PhysSize not found");
    Hid ds(a.Dopen2(g.id, "PhysSize", P_DEFAULT), a.Dclose);
    std::vector<double> v;
    read_doubles(ds.id, v, 1);
    phys_size = v[0];
  }
  const std::string group = morphology_type == 0 ? "Euler_Angles" : "Vector_Morphology";
  const std::string name = morphology_type == 0 ? "Mat_1_Vfrac" : "Mat_1_unaligned";
  if (!exists(f.id, group)) throw H5Error("[HDF5 Error] Group " + group + " not found");
  Hid g(a.Gopen2(f.id, group.c_str(), P_DEFAULT), a.Gclose);
  if (!exists(g.id, name)) throw H5Error("[HDF5 Error] DataSet " + name + " not found");
  Hid ds(a.Dopen2(g.id, name.c_str(), P_DEFAULT), a.Dclose);
  Hid sp(a.Dget_space(ds.id), a.Sclose);
  if (a.Sget_simple_extent_ndims(sp.id) != 3) throw H5Error("[HDF5 Error] Expected 3D array for " + name);
  hsize_t d[3];
  a.Sget_simple_extent_dims(sp.id, d, nullptr);
  const std::string l0 = label(ds.id, 0), l2 = label(ds.id, 2);
  if (l0 == "Z" && l2 == "X") {
    order = ZYX;
    dims[0] = (uint32_t)d[2];
    dims[1] = (uint32_t)d[1];
    dims[2] = (uint32_t)d[0];
  } else if (l0 == "X" && l2 == "Z") {
    order = XYZ;
    dims[0] = (uint32_t)d[0];
    dims[1] = (uint32_t)d[1];
    dims[2] = (uint32_t)d[2];
  } else {
    throw H5Error("[HDF5 Error] Only XYZ/ZYX ordering supported");
  }
}

void h5_read_morphology(const std::string &file, const uint32_t dims[3], int morphology_type, int order,
                        int num_materials, std::vector<double> &voxels) {
  H5Api &a = api();
  const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2];
  voxels.assign(V * num_materials * 4, 0.0);
  Hid f(check(a.Fopen(file.c_str(), F_ACC_RDONLY, P_DEFAULT), "cannot open " + file), a.Fclose);
  std::vector<double> s;
  for (int m = 1; m <= num_materials; m++) {
    double *vm = &voxels[(uint64_t)(m - 1) * V * 4];
    if (morphology_type == 1) {  // vector morphology (include/Input/readH5.h:458-481)
      get_scalar(f.id, "Vector_Morphology", "_unaligned", dims, order, m, true, s);
      for (uint64_t i = 0; i < V; i++) vm[4 * i + 3] = s[i];
      Hid g(check(a.Gopen2(f.id, "Vector_Morphology", P_DEFAULT), "open Vector_Morphology"), a.Gclose);
      const std::string name = "Mat_" + std::to_string(m) + "_alignment";
      if (!exists(g.id, name)) throw H5Error("[HDF5 Error] Dataset = " + name + " does not exist");
      Hid ds(a.Dopen2(g.id, name.c_str(), P_DEFAULT), a.Dclose);
      check_labels_and_dims(ds.id, name, dims, order, 4, 3);
      std::vector<double> al;
      read_doubles(ds.id, al, V * 3);
      if (order == XYZ) xyz_to_zyx(al, 3, dims);
      for (uint64_t i = 0; i < V; i++) {
        vm[4 * i + 0] = al[3 * i + 0];
        vm[4 * i + 1] = al[3 * i + 1];
        vm[4 * i + 2] = al[3 * i + 2];
      }
    } else {  // Euler angles (include/Input/readH5.h:482-533)
      get_scalar(f.id, "Euler_Angles", "_Vfrac", dims, order, m, true, s);
      for (uint64_t i = 0; i < V; i++) vm[4 * i + 3] = s[i];
      const bool has_s = get_scalar(f.id, "Euler_Angles", "_S", dims, order, m, false, s);
      for (uint64_t i = 0; i < V; i++) vm[4 * i + 0] = s[i];
      if (has_s) {
        get_scalar(f.id, "Euler_Angles", "_Theta", dims, order, m, true, s);
        for (uint64_t i = 0; i < V; i++) vm[4 * i + 1] = vm[4 * i + 0] == 0 ? 0.0 : s[i];
        get_scalar(f.id, "Euler_Angles", "_Psi", dims, order, m, true, s);
        for (uint64_t i = 0; i < V; i++) vm[4 * i + 2] = vm[4 * i + 0] == 0 ? 0.0 : s[i];
      }
    }
  }
}

void create_directory(const std::string &dir) {
  std::string cur;
  std::stringstream ss(dir);
  std::string part;
  if (!dir.empty() && dir[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) throw H5Error("cannot create directory " + cur);
  }
}

void h5_write_patterns(const InputData &in, const void *patterns, const std::string &dir) {
  H5Api &a = api();
  create_directory(dir);
  const uint64_t P = (uint64_t)in.voxelDims[0] * in.voxelDims[1];
  const size_t nK = in.kVectors.size();
  const bool dbl = in.precision == CYR_FP64;
  const hid_t mt = dbl ? *a.NATIVE_DOUBLE : *a.NATIVE_FLOAT;
  const size_t es = dbl ? 8 : 4;
  for (size_t e = 0; e < in.energies.size(); e++) {
    std::stringstream st;
    st << std::fixed << std::setprecision(2) << in.energies[e];
    const std::string fname = dir + "/Energy_" + st.str() + ".h5";
    Hid f(check(a.Fcreate(fname.c_str(), F_ACC_TRUNC, P_DEFAULT, P_DEFAULT), "cannot create " + fname), a.Fclose);
    {
      Hid g(check(a.Gcreate2(f.id, "KIDList", P_DEFAULT, P_DEFAULT, P_DEFAULT), "KIDList"), a.Gclose);
      const hsize_t d2[2] = {nK, 3};
      Hid sp(a.Screate_simple(2, d2, nullptr), a.Sclose);
      Hid ds(check(a.Dcreate2(g.id, "KVec", *a.NATIVE_FLOAT, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), "KVec"),
             a.Dclose);
      std::vector<double> kf;
      for (const auto &k : in.kVectors) kf.insert(kf.end(), k.begin(), k.end());
      a.Dwrite(ds.id, *a.NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, kf.data());
    }
    for (size_t k = 0; k < nK; k++) {
      const std::string gname = "K" + std::to_string(k);
      Hid g(check(a.Gcreate2(f.id, gname.c_str(), P_DEFAULT, P_DEFAULT, P_DEFAULT), gname), a.Gclose);
      const hsize_t d2[2] = {in.voxelDims[1], in.voxelDims[0]};
      Hid sp(a.Screate_simple(2, d2, nullptr), a.Sclose);
      Hid ds(check(a.Dcreate2(g.id, "projection", mt, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), "projection"),
             a.Dclose);
      const char *src = reinterpret_cast<const char *>(patterns) + ((uint64_t)e * nK + k) * P * es;
      if (a.Dwrite(ds.id, mt, S_ALL, S_ALL, P_DEFAULT, src) < 0) throw H5Error("[HDF5 Error] write " + fname);
      write_labels(ds.id, {"Qy", "Qx"});
    }
  }
}

void h5_write_morphology(const std::string &file, const InputData &in, const double *voxels) {
  H5Api &a = api();
  const uint64_t V = in.numVoxels();
  Hid f(check(a.Fcreate(file.c_str(), F_ACC_TRUNC, P_DEFAULT, P_DEFAULT), "cannot create " + file), a.Fclose);
  {
    Hid g(a.Gcreate2(f.id, "Morphology_Parameters", P_DEFAULT, P_DEFAULT, P_DEFAULT), a.Gclose);
    Hid sp(a.Screate(S_SCALAR), a.Sclose);
    Hid d1(a.Dcreate2(g.id, "NumMaterial", *a.NATIVE_LLONG, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), a.Dclose);
    long long nm = in.NUM_MATERIAL;
    a.Dwrite(d1.id, *a.NATIVE_LLONG, S_ALL, S_ALL, P_DEFAULT, &nm);
    Hid d2(a.Dcreate2(g.id, "PhysSize", *a.NATIVE_DOUBLE, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), a.Dclose);
    a.Dwrite(d2.id, *a.NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, &in.physSize);
  }
  const bool vec = in.morphologyType == 1;
  Hid g(a.Gcreate2(f.id, vec ? "Vector_Morphology" : "Euler_Angles", P_DEFAULT, P_DEFAULT, P_DEFAULT), a.Gclose);
  const hsize_t d3[3] = {in.voxelDims[2], in.voxelDims[1], in.voxelDims[0]};
  const hsize_t d4[4] = {in.voxelDims[2], in.voxelDims[1], in.voxelDims[0], 3};
  std::vector<double> buf;
  for (int m = 0; m < in.NUM_MATERIAL; m++) {
    const double *vm = voxels + (uint64_t)m * V * 4;
    auto put = [&](const std::string &name, int comp) {
      buf.resize(V);
      for (uint64_t i = 0; i < V; i++) buf[i] = vm[4 * i + comp];
      Hid sp(a.Screate_simple(3, d3, nullptr), a.Sclose);
      Hid ds(a.Dcreate2(g.id, name.c_str(), *a.NATIVE_DOUBLE, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), a.Dclose);
      a.Dwrite(ds.id, *a.NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, buf.data());
      write_labels(ds.id, {"Z", "Y", "X"});
    };
    const std::string pre = "Mat_" + std::to_string(m + 1);
    if (vec) {
      put(pre + "_unaligned", 3);
      buf.resize(V * 3);
      for (uint64_t i = 0; i < V; i++)
        for (int c = 0; c < 3; c++) buf[3 * i + c] = vm[4 * i + c];
      Hid sp(a.Screate_simple(4, d4, nullptr), a.Sclose);
      Hid ds(a.Dcreate2(g.id, (pre + "_alignment").c_str(), *a.NATIVE_DOUBLE, sp.id, P_DEFAULT, P_DEFAULT,
                        P_DEFAULT),
             a.Dclose);
      a.Dwrite(ds.id, *a.NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, buf.data());
      write_labels(ds.id, {"Z", "Y", "X", "C"});
    } else {
      put(pre + "_S", 0);
      put(pre + "_Theta", 1);
      put(pre + "_Psi", 2);
      put(pre + "_Vfrac", 3);
    }
  }
}

void h5_write_polarization(const std::string &file, const std::string &name, const InputData &in,
                           const void *data) {
  H5Api &a = api();
  Hid f(check(a.Fcreate(file.c_str(), F_ACC_TRUNC, P_DEFAULT, P_DEFAULT), "cannot create " + file), a.Fclose);
  const hsize_t d4[4] = {in.voxelDims[2], in.voxelDims[1], in.voxelDims[0], 2};
  Hid sp(a.Screate_simple(4, d4, nullptr), a.Sclose);
  const hid_t mt = in.precision == CYR_FP64 ? *a.NATIVE_DOUBLE : *a.NATIVE_FLOAT;
  Hid ds(a.Dcreate2(f.id, name.c_str(), mt, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), a.Dclose);
  a.Dwrite(ds.id, mt, S_ALL, S_ALL, P_DEFAULT, data);
  write_labels(ds.id, {"Z", "Y", "X"});
}

}  // namespace cyrhost
