// openPMD HDF5 beam files of charged particles (bmad/hdf5/hdf5_read_beam.f90,
// hdf5_write_beam.f90, and the record helpers of hdf5_openpmd_mod.f90), host only.
//
// The HDF5 library is loaded when a file is first opened, not linked: BMX_HDF5_LIB if set, else
// libhdf5.so.103 or libhdf5.so from the loader's path, else /opt/conda/lib. Its header supplies
// the types and enumerations at build time; the functions and the predefined type identifiers
// are looked up in the loaded library. Without the header the entry points report
// BMX_ERR_UNSUPPORTED. HDF5 builds are not thread-safe in general, so every call holds one mutex.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <vector>

#include "bmx_internal.hpp"

#ifdef BMX_HAVE_HDF5
#include <dlfcn.h>
#include <hdf5.h>
#endif

using namespace bmx;

namespace {

std::mutex g_h5_mutex;

#ifdef BMX_HAVE_HDF5

struct H5Api {
    bool ok = false;
    std::string why;
    decltype(&H5open) open;
    decltype(&H5Eset_auto2) eset_auto;
    decltype(&H5Fopen) fopen;
    decltype(&H5Fcreate) fcreate;
    decltype(&H5Fclose) fclose;
    decltype(&H5Gopen2) gopen;
    decltype(&H5Gcreate2) gcreate;
    decltype(&H5Gclose) gclose;
    decltype(&H5Gget_info) gget_info;
    decltype(&H5Lget_name_by_idx) lget_name;
    decltype(&H5Lexists) lexists;
    decltype(&H5Oopen) oopen;
    decltype(&H5Oclose) oclose;
    decltype(&H5Iget_type) iget_type;
    decltype(&H5Aexists) aexists;
    decltype(&H5Aopen) aopen;
    decltype(&H5Aread) aread;
    decltype(&H5Awrite) awrite;
    decltype(&H5Acreate2) acreate;
    decltype(&H5Aclose) aclose;
    decltype(&H5Adelete) adelete;
    decltype(&H5Aget_type) aget_type;
    decltype(&H5Aget_space) aget_space;
    decltype(&H5Sget_simple_extent_npoints) s_npoints;
    decltype(&H5Sget_simple_extent_ndims) s_ndims;
    decltype(&H5Screate) screate;
    decltype(&H5Screate_simple) screate_simple;
    decltype(&H5Sclose) sclose;
    decltype(&H5Tget_class) tget_class;
    decltype(&H5Tget_size) tget_size;
    decltype(&H5Tis_variable_str) tis_vlen;
    decltype(&H5Tcopy) tcopy;
    decltype(&H5Tset_size) tset_size;
    decltype(&H5Tset_strpad) tset_strpad;
    decltype(&H5Tclose) tclose;
    decltype(&H5Dopen2) dopen;
    decltype(&H5Dcreate2) dcreate;
    decltype(&H5Dread) dread;
    decltype(&H5Dwrite) dwrite;
    decltype(&H5Dclose) dclose;
    decltype(&H5Dget_space) dget_space;
    decltype(&H5free_memory) free_memory;
    hid_t *native_double = nullptr, *native_int = nullptr, *c_s1 = nullptr;
};

H5Api g_api;
bool g_tried = false;

// File access flags (H5Fpublic.h), spelled out: the header's macros call H5open, which is
// not linked here.
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;

template <class T>
bool sym(void *lib, const char *name, T &f)
{
    f = reinterpret_cast<T>(dlsym(lib, name));
    return f != nullptr;
}

// Loads the library once (under g_h5_mutex).
const H5Api *h5()
{
    if (g_tried) return g_api.ok ? &g_api : nullptr;
    g_tried = true;
    std::vector<std::string> names;
    if (const char *e = std::getenv("BMX_HDF5_LIB")) names.push_back(e);
    names.push_back("libhdf5.so.103");
    names.push_back("libhdf5.so");
    names.push_back("/opt/conda/lib/libhdf5.so.103");
    names.push_back("/opt/conda/lib/libhdf5.so");
    void *lib = nullptr;
    for (const std::string &n : names)
        if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!lib) {
        g_api.why = "libhdf5 not found (set BMX_HDF5_LIB)";
        return nullptr;
    }
    H5Api &a = g_api;
    bool ok = sym(lib, "H5open", a.open) && sym(lib, "H5Eset_auto2", a.eset_auto) &&
              sym(lib, "H5Fopen", a.fopen) && sym(lib, "H5Fcreate", a.fcreate) &&
              sym(lib, "H5Fclose", a.fclose) && sym(lib, "H5Gopen2", a.gopen) &&
              sym(lib, "H5Gcreate2", a.gcreate) && sym(lib, "H5Gclose", a.gclose) &&
              sym(lib, "H5Gget_info", a.gget_info) && sym(lib, "H5Lget_name_by_idx", a.lget_name) &&
              sym(lib, "H5Lexists", a.lexists) && sym(lib, "H5Oopen", a.oopen) &&
              sym(lib, "H5Oclose", a.oclose) && sym(lib, "H5Iget_type", a.iget_type) &&
              sym(lib, "H5Aexists", a.aexists) && sym(lib, "H5Aopen", a.aopen) &&
              sym(lib, "H5Aread", a.aread) && sym(lib, "H5Awrite", a.awrite) &&
              sym(lib, "H5Acreate2", a.acreate) && sym(lib, "H5Aclose", a.aclose) && sym(lib, "H5Adelete", a.adelete) &&
              sym(lib, "H5Aget_type", a.aget_type) && sym(lib, "H5Aget_space", a.aget_space) &&
              sym(lib, "H5Sget_simple_extent_npoints", a.s_npoints) &&
              sym(lib, "H5Sget_simple_extent_ndims", a.s_ndims) && sym(lib, "H5Screate", a.screate) &&
              sym(lib, "H5Screate_simple", a.screate_simple) && sym(lib, "H5Sclose", a.sclose) &&
              sym(lib, "H5Tget_class", a.tget_class) && sym(lib, "H5Tget_size", a.tget_size) &&
              sym(lib, "H5Tis_variable_str", a.tis_vlen) && sym(lib, "H5Tcopy", a.tcopy) &&
              sym(lib, "H5Tset_size", a.tset_size) && sym(lib, "H5Tset_strpad", a.tset_strpad) &&
              sym(lib, "H5Tclose", a.tclose) && sym(lib, "H5Dopen2", a.dopen) &&
              sym(lib, "H5Dcreate2", a.dcreate) && sym(lib, "H5Dread", a.dread) &&
              sym(lib, "H5Dwrite", a.dwrite) && sym(lib, "H5Dclose", a.dclose) &&
              sym(lib, "H5Dget_space", a.dget_space) && sym(lib, "H5free_memory", a.free_memory);
    if (ok) {
        // the predefined types are globals that H5open initialises
        a.native_double = static_cast<hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
        a.native_int = static_cast<hid_t *>(dlsym(lib, "H5T_NATIVE_INT_g"));
        a.c_s1 = static_cast<hid_t *>(dlsym(lib, "H5T_C_S1_g"));
        ok = a.native_double && a.native_int && a.c_s1 && a.open() >= 0;
    }
    if (!ok) {
        g_api.why = "libhdf5 lacks an expected function";
        return nullptr;
    }
    a.eset_auto(H5E_DEFAULT, nullptr, nullptr);  // errors come back as return codes only
    a.ok = true;
    return &g_api;
}

// ---- attributes ---------------------------------------------------------------------------

bool has_attr(const H5Api &a, hid_t obj, const char *name) { return a.aexists(obj, name) > 0; }

bool read_attr_string(const H5Api &a, hid_t obj, const char *name, std::string &out)
{
    out.clear();
    if (!has_attr(a, obj, name)) return false;
    const hid_t at = a.aopen(obj, name, H5P_DEFAULT);
    if (at < 0) return false;
    const hid_t ty = a.aget_type(at);
    bool ok = false;
    if (a.tget_class(ty) == H5T_STRING) {
        const hid_t mem = a.tcopy(*a.c_s1);
        if (a.tis_vlen(ty) > 0) {
            a.tset_size(mem, H5T_VARIABLE);
            char *s = nullptr;
            ok = a.aread(at, mem, &s) >= 0;
            if (ok && s) out = s;
            if (s) a.free_memory(s);
        } else {
            const size_t n = a.tget_size(ty);
            a.tset_size(mem, n + 1);
            std::vector<char> buf(n + 1, '\0');
            ok = a.aread(at, mem, buf.data()) >= 0;
            if (ok) out.assign(buf.data());
        }
        a.tclose(mem);
    }
    a.tclose(ty);
    a.aclose(at);
    return ok;
}

bool read_attr_doubles(const H5Api &a, hid_t obj, const char *name, std::vector<double> &out)
{
    if (!has_attr(a, obj, name)) return false;
    const hid_t at = a.aopen(obj, name, H5P_DEFAULT);
    if (at < 0) return false;
    const hid_t sp = a.aget_space(at);
    const hssize_t n = a.s_npoints(sp);
    a.sclose(sp);
    out.assign(n > 0 ? (size_t)n : 0, 0.0);
    const bool ok = n > 0 && a.aread(at, *a.native_double, out.data()) >= 0;
    a.aclose(at);
    return ok;
}

bool read_attr_double(const H5Api &a, hid_t obj, const char *name, double &v)
{
    std::vector<double> d;
    if (!read_attr_doubles(a, obj, name, d)) return false;
    v = d[0];
    return true;
}

bool read_attr_int(const H5Api &a, hid_t obj, const char *name, int &v)
{
    if (!has_attr(a, obj, name)) return false;
    const hid_t at = a.aopen(obj, name, H5P_DEFAULT);
    if (at < 0) return false;
    const hid_t sp = a.aget_space(at);
    const hssize_t n = a.s_npoints(sp);
    a.sclose(sp);
    std::vector<int> buf(n > 0 ? (size_t)n : 1, 0);
    const bool ok = n > 0 && a.aread(at, *a.native_int, buf.data()) >= 0;
    a.aclose(at);
    if (ok) v = buf[0];
    return ok;
}

// H5LTset_attribute_string's form: a scalar, fixed-length, null-terminated string.
void write_attr_string(const H5Api &a, hid_t obj, const char *name, const std::string &s)
{
    const hid_t ty = a.tcopy(*a.c_s1);
    a.tset_size(ty, s.size() + 1);
    a.tset_strpad(ty, H5T_STR_NULLTERM);
    const hid_t sp = a.screate(H5S_SCALAR);
    const hid_t at = a.acreate(obj, name, ty, sp, H5P_DEFAULT, H5P_DEFAULT);
    a.awrite(at, ty, s.c_str());
    a.aclose(at);
    a.sclose(sp);
    a.tclose(ty);
}

void write_attr_doubles(const H5Api &a, hid_t obj, const char *name, const double *v, hsize_t n)
{
    const hid_t sp = a.screate_simple(1, &n, nullptr);
    const hid_t at = a.acreate(obj, name, *a.native_double, sp, H5P_DEFAULT, H5P_DEFAULT);
    a.awrite(at, *a.native_double, v);
    a.aclose(at);
    a.sclose(sp);
}

void write_attr_ints(const H5Api &a, hid_t obj, const char *name, const int *v, hsize_t n)
{
    const hid_t sp = a.screate_simple(1, &n, nullptr);
    const hid_t at = a.acreate(obj, name, *a.native_int, sp, H5P_DEFAULT, H5P_DEFAULT);
    a.awrite(at, *a.native_int, v);
    a.aclose(at);
    a.sclose(sp);
}

// ---- records (hdf5_openpmd_mod.f90) ---------------------------------------------------------

// A record component: a one-dimensional dataset, or a "constant record component" (a group
// holding a 'value' attribute). Its unitSI and the conversion factor of the quantity wanted:
// values are multiplied by conversion / unitSI when the two differ by more than 1e-15
// relative (pmd_read_real_dataset_rank1, :1227-1270).
bool read_record(const H5Api &a, hid_t parent, const std::string &name, double conversion, size_t n,
                 std::vector<double> &out, std::string &err)
{
    if (a.lexists(parent, name.c_str(), H5P_DEFAULT) <= 0) return false;
    const hid_t obj = a.oopen(parent, name.c_str(), H5P_DEFAULT);
    if (obj < 0) {
        err = "cannot open " + name;
        return false;
    }
    double unit_si = 0;
    if (!read_attr_double(a, obj, "unitSI", unit_si)) {
        err = "no unitSI on " + name;
        a.oclose(obj);
        return false;
    }
    out.assign(n, 0.0);
    bool ok = true;
    if (a.iget_type(obj) == H5I_DATASET) {
        const hid_t sp = a.dget_space(obj);
        const int nd = a.s_ndims(sp);
        const hssize_t np = a.s_npoints(sp);
        a.sclose(sp);
        if (nd != 1 || (size_t)np != n) {
            err = name + " is not a one-dimensional array of the bunch's length";
            ok = false;
        } else {
            ok = a.dread(obj, *a.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
            if (!ok) err = "cannot read " + name;
        }
    } else {
        double v = 0;
        ok = read_attr_double(a, obj, "value", v);
        if (ok) std::fill(out.begin(), out.end(), v);
        else err = "constant record " + name + " has no value";
    }
    a.oclose(obj);
    if (ok && std::fabs(unit_si - conversion) > 1e-15 * conversion)
        for (double &v : out) v = v * (conversion / unit_si);
    return ok;
}

struct Unit {
    const char *symbol;
    double unit_si;
    double dim[7];
};

// hdf5_openpmd_mod.f90:31-47 (unit_1, unit_m, unit_sec, unit_eV_per_c, unit_hbar)
const double kECharge = 1.602176634e-19, kCLight = 2.99792458e8;
const double kHbarEvS = 4.135667696e-15 / (2 * 3.141592653589793238462643383279);
const Unit kUnit1 = {"", 1.0, {0, 0, 0, 0, 0, 0, 0}};
const Unit kUnitM = {"m", 1.0, {1, 0, 0, 0, 0, 0, 0}};
const Unit kUnitSec = {"sec", 1.0, {0, 0, 1, 0, 0, 0, 0}};
const Unit kUnitEvPerC = {"eV/c", kECharge / kCLight, {1, 1, -1, 0, 0, 0, 0}};
const Unit kUnitHbar = {"hbar", kECharge * kHbarEvS, {2, 1, -1, 0, 0, 0, 0}};

void write_units(const H5Api &a, hid_t obj, const char *local, const Unit &u)
{
    if (local[0]) write_attr_string(a, obj, "localName", local);
    write_attr_doubles(a, obj, "unitSI", &u.unit_si, 1);
    write_attr_doubles(a, obj, "unitDimension", u.dim, 7);
    write_attr_string(a, obj, "unitSymbol", u.symbol);
}

// pmd_write_real_to_dataset_rank1 (:393-420): equal values become a constant record (a group
// with 'value' and 'shape'), others a dataset with minValue and maxValue; then the units.
void write_real_record(const H5Api &a, hid_t parent, const char *name, const char *local, const Unit &u,
                       const std::vector<double> &v)
{
    double lo = v.empty() ? 0 : v[0], hi = lo;
    for (double x : v) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
    }
    const int shape = (int)v.size();
    if (lo == hi) {
        const hid_t g = a.gcreate(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        write_attr_doubles(a, g, "value", &hi, 1);
        write_attr_ints(a, g, "shape", &shape, 1);
        write_units(a, g, local, u);
        a.gclose(g);
        return;
    }
    const hsize_t n = v.size();
    const hid_t sp = a.screate_simple(1, &n, nullptr);
    const hid_t d = a.dcreate(parent, name, *a.native_double, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    a.dwrite(d, *a.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
    write_attr_doubles(a, d, "minValue", &lo, 1);
    write_attr_doubles(a, d, "maxValue", &hi, 1);
    write_units(a, d, local, u);
    a.dclose(d);
    a.sclose(sp);
}

void write_int_record(const H5Api &a, hid_t parent, const char *name, const char *local,
                      const std::vector<int> &v)
{
    int lo = v.empty() ? 0 : v[0], hi = lo;
    for (int x : v) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
    }
    const int shape = (int)v.size();
    if (lo == hi) {
        const hid_t g = a.gcreate(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        write_attr_ints(a, g, "value", &hi, 1);
        write_attr_ints(a, g, "shape", &shape, 1);
        write_units(a, g, local, kUnit1);
        a.gclose(g);
        return;
    }
    const hsize_t n = v.size();
    const hid_t sp = a.screate_simple(1, &n, nullptr);
    const hid_t d = a.dcreate(parent, name, *a.native_int, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    a.dwrite(d, *a.native_int, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
    write_attr_ints(a, d, "minValue", &lo, 1);
    write_attr_ints(a, d, "maxValue", &hi, 1);
    write_units(a, d, local, kUnit1);
    a.dclose(d);
    a.sclose(sp);
}

// ---- bunch discovery (hdf5_read_beam.f90:43-120) ---------------------------------------------

struct FileHeader {
    std::string base_path, particles_path, software;
};

std::vector<std::string> child_names(const H5Api &a, hid_t g)
{
    std::vector<std::string> out;
    H5G_info_t info;
    if (a.gget_info(g, &info) < 0) return out;
    for (hsize_t i = 0; i < info.nlinks; i++) {
        char name[256] = {0};
        if (a.lget_name(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, name, sizeof name, H5P_DEFAULT) < 0) continue;
        out.emplace_back(name);
    }
    return out;
}

bool is_group(const H5Api &a, hid_t parent, const std::string &name)
{
    const hid_t o = a.oopen(parent, name.c_str(), H5P_DEFAULT);
    if (o < 0) return false;
    const bool g = a.iget_type(o) == H5I_GROUP;
    a.oclose(o);
    return g;
}

// The bunch groups of the file, in name order: basePath without %T is one bunch; with %T the
// groups of its directory whose names match it with an integer in place of %T.
bool bunch_paths(const H5Api &a, hid_t f, const FileHeader &h, std::vector<std::string> &paths,
                 std::string &err)
{
    paths.clear();
    const size_t it = h.base_path.find("%T");
    if (it == std::string::npos) {
        paths.push_back(h.base_path);
        return true;
    }
    const size_t is = h.base_path.rfind('/', it);
    const std::string dir = is == std::string::npos ? "/" : h.base_path.substr(0, is + 1);
    std::string t_match = h.base_path.substr(is == std::string::npos ? 0 : is + 1);
    if (!t_match.empty() && t_match.back() == '/') t_match.pop_back();
    std::string sub_dir;
    const size_t sl = t_match.find('/');
    if (sl != std::string::npos) {
        sub_dir = t_match.substr(sl);
        t_match = t_match.substr(0, sl);
    }
    const size_t tt = t_match.find("%T");
    const std::string pre = t_match.substr(0, tt), post = t_match.substr(tt + 2);
    const hid_t g = a.gopen(f, dir.c_str(), H5P_DEFAULT);
    if (g < 0) {
        err = "no group " + dir;
        return false;
    }
    for (const std::string &n : child_names(a, g)) {
        if (!is_group(a, g, n)) continue;
        if (n.size() < pre.size() + post.size() + 1) continue;
        if (n.compare(0, pre.size(), pre) != 0) continue;
        if (n.compare(n.size() - post.size(), post.size(), post) != 0) continue;
        const std::string mid = n.substr(pre.size(), n.size() - pre.size() - post.size());
        bool digits = !mid.empty();
        for (size_t k = 0; k < mid.size(); k++)
            digits &= (mid[k] >= '0' && mid[k] <= '9') || (k == 0 && (mid[k] == '-' || mid[k] == '+'));
        if (!digits) continue;
        paths.push_back(dir + n + sub_dir);
    }
    a.gclose(g);
    if (paths.empty()) {
        err = "no beam data in the file";
        return false;
    }
    return true;
}

bool read_header(const H5Api &a, hid_t f, FileHeader &h, std::string &err)
{
    std::string s;
    if (!read_attr_string(a, f, "openPMD", s) || !read_attr_string(a, f, "openPMDextension", s) ||
        !read_attr_string(a, f, "basePath", h.base_path) ||
        !read_attr_string(a, f, "particlesPath", h.particles_path)) {
        err = "not an openPMD beam file (openPMD, openPMDextension, basePath or particlesPath missing)";
        return false;
    }
    read_attr_string(a, f, "software", h.software);
    return true;
}

// The particle group of a bunch: the particles group itself when it carries numParticles (the
// older layout), else its first subgroup that does (one per species).
hid_t particle_group(const H5Api &a, hid_t f, const std::string &bunch, const FileHeader &h, int &n,
                     std::string &err)
{
    std::string path = bunch;
    if (!path.empty() && path.back() != '/') path += '/';
    path += h.particles_path;
    const hid_t g1 = a.gopen(f, path.c_str(), H5P_DEFAULT);
    if (g1 < 0) {
        err = "no particle group " + path;
        return -1;
    }
    if (read_attr_int(a, g1, "numParticles", n)) return g1;
    for (const std::string &c : child_names(a, g1)) {
        if (!is_group(a, g1, c)) continue;
        const hid_t g3 = a.gopen(g1, c.c_str(), H5P_DEFAULT);
        if (g3 >= 0 && read_attr_int(a, g3, "numParticles", n)) {
            a.gclose(g1);
            return g3;
        }
        if (g3 >= 0) a.gclose(g3);
    }
    a.gclose(g1);
    err = "no numParticles under " + path;
    return -1;
}

// openPMD species names of the charged particles this path tracks (particle_species_mod.f90's
// openPMD_subatomic_species_name), with mc^2 [eV] as particle_species_mod.f90 gives them.
struct SpeciesMass {
    const char *name;
    double mc2;
    int charge;
};
const SpeciesMass kSpecies[] = {{"electron", 0.51099895069e6, -1}, {"positron", 0.51099895069e6, 1},
                                {"proton", 938.27208816e6, 1},    {"anti-proton", 938.27208816e6, -1},
                                {"muon", 105.6583755e6, -1},      {"anti-muon", 105.6583755e6, 1}};

#endif  // BMX_HAVE_HDF5

}  // namespace

extern "C" {

int32_t bmx_beam_h5_info(const char *path, int32_t *n_bunch, int64_t *n_particle, char *species,
                         int32_t species_len)
{
#ifndef BMX_HAVE_HDF5
    (void)path; (void)n_bunch; (void)n_particle; (void)species; (void)species_len;
    return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_info: built without HDF5 headers");
#else
    if (!path || !n_bunch || !n_particle) return set_error(BMX_ERR_ARG, "bmx_beam_h5_info: null argument");
    std::lock_guard<std::mutex> lock(g_h5_mutex);
    const H5Api *a = h5();
    if (!a) return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_info: " + g_api.why);
    const hid_t f = a->fopen(path, kAccRdonly, H5P_DEFAULT);
    if (f < 0) return set_error(BMX_ERR_ARG, std::string("bmx_beam_h5_info: cannot open ") + path);
    FileHeader h;
    std::vector<std::string> paths;
    std::string err;
    int32_t rc = BMX_OK;
    if (!read_header(*a, f, h, err) || !bunch_paths(*a, f, h, paths, err)) {
        rc = set_error(BMX_ERR_ARG, "bmx_beam_h5_info: " + err);
    } else {
        int n = 0;
        const hid_t g = particle_group(*a, f, paths[0], h, n, err);
        if (g < 0) {
            rc = set_error(BMX_ERR_ARG, "bmx_beam_h5_info: " + err);
        } else {
            *n_bunch = (int32_t)paths.size();
            *n_particle = n;
            if (species && species_len > 0) {
                std::string s;
                read_attr_string(*a, g, "speciesType", s);
                std::snprintf(species, (size_t)species_len, "%s", s.c_str());
            }
            a->gclose(g);
        }
    }
    a->fclose(f);
    return rc;
#endif
}

int32_t bmx_beam_h5_read(const char *path, int32_t ix_bunch, const bmx_beam_view *host, int64_t cap,
                         double mc2, int64_t *n_out, double *p0c_out, double *charge_tot)
{
#ifndef BMX_HAVE_HDF5
    (void)path; (void)ix_bunch; (void)host; (void)cap; (void)mc2; (void)n_out; (void)p0c_out; (void)charge_tot;
    return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_read: built without HDF5 headers");
#else
    if (!path || !host || cap < 0 || ix_bunch < 1 || !(mc2 > 0))
        return set_error(BMX_ERR_ARG, "bmx_beam_h5_read: bad argument");
    std::lock_guard<std::mutex> lock(g_h5_mutex);
    const H5Api *ap = h5();
    if (!ap) return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_read: " + g_api.why);
    const H5Api &a = *ap;
    const hid_t f = a.fopen(path, kAccRdonly, H5P_DEFAULT);
    if (f < 0) return set_error(BMX_ERR_ARG, std::string("bmx_beam_h5_read: cannot open ") + path);
    FileHeader h;
    std::vector<std::string> paths;
    std::string err;
    int n = 0;
    hid_t g = -1;
    if (read_header(a, f, h, err) && bunch_paths(a, f, h, paths, err)) {
        if (ix_bunch > (int32_t)paths.size()) err = "the file has " + std::to_string(paths.size()) + " bunches";
        else g = particle_group(a, f, paths[ix_bunch - 1], h, n, err);
    }
    if (g < 0) {
        a.fclose(f);
        return set_error(BMX_ERR_ARG, "bmx_beam_h5_read: " + err);
    }
    if (n > cap) {
        a.gclose(g);
        a.fclose(f);
        return set_error(BMX_ERR_ARG, "bmx_beam_h5_read: bunch larger than the arrays");
    }
    // bunch attributes (hdf5_read_beam.f90:200-223)
    std::string species;
    double q_tot = 0, q_unit = 0, q_live = 0;
    read_attr_string(a, g, "speciesType", species);
    read_attr_double(a, g, "totalCharge", q_tot);
    read_attr_double(a, g, "chargeLive", q_live);
    read_attr_double(a, g, "chargeUnitSI", q_unit);
    const size_t N = (size_t)n;
    bool charged = false;
    for (const SpeciesMass &s : kSpecies) charged |= species == s.name;
    if (!charged) {
        a.gclose(g);
        a.fclose(f);
        return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_read: species '" + species +
                                                  "' (this path tracks charged leptons and protons)");
    }
    // the records this path carries (:227-299)
    const double f_ev = kUnitEvPerC.unit_si;
    std::vector<double> x(N, 0), y(N, 0), z(N, 0), xo(N, 0), yo(N, 0), zo(N, 0), px(N, 0), py(N, 0),
        pz(N, 0), pxo(N, 0), pyo(N, 0), pzo(N, 0), dt(N, 0), t0(N, 0), p0c(N, 0), w(N, 1), st(N, 1),
        ixe(N, 0), tmp;
    std::vector<double> tot(N, NAN);  // garbage until read
    bool has_tot = false;
    struct Rec {
        const char *name;
        double conv;
        std::vector<double> *dst;
        bool *flag;
    };
    const Rec recs[] = {{"position/x", 1.0, &x, nullptr},
                        {"position/y", 1.0, &y, nullptr},
                        {"position/z", 1.0, &z, nullptr},
                        {"positionOffset/x", 1.0, &xo, nullptr},
                        {"positionOffset/y", 1.0, &yo, nullptr},
                        {"positionOffset/z", 1.0, &zo, nullptr},
                        {"momentum/x", f_ev, &px, nullptr},
                        {"momentum/y", f_ev, &py, nullptr},
                        {"momentum/z", f_ev, &pz, nullptr},
                        {"momentumOffset/x", f_ev, &pxo, nullptr},
                        {"momentumOffset/y", f_ev, &pyo, nullptr},
                        {"momentumOffset/z", f_ev, &pzo, nullptr},
                        {"time", 1.0, &dt, nullptr},
                        {"timeOffset", 1.0, &t0, nullptr},
                        {"totalMomentum", f_ev, &tot, &has_tot},
                        {"totalMomentumOffset", f_ev, &p0c, nullptr},
                        {"weight", 1.0, &w, nullptr},
                        {"particleStatus", 1.0, &st, nullptr},
                        {"elementIndex", 1.0, &ixe, nullptr}};
    for (const Rec &r : recs) {
        err.clear();
        if (read_record(a, g, r.name, r.conv, N, tmp, err)) {
            *r.dst = tmp;
            if (r.flag) *r.flag = true;
        } else if (!err.empty()) {
            a.gclose(g);
            a.fclose(f);
            return set_error(BMX_ERR_ARG, "bmx_beam_h5_read: " + err);
        }
    }
    a.gclose(g);
    a.fclose(f);
    // coordinates (:327-399), no lattice element given: the particles keep the file's reference
    // momentum (p0c_initial = p0c_final = totalMomentumOffset)
    const bool bmad = h.software == "Bmad" || h.software == "bmadx-mi355x";
    double sum_w = 0;
    for (size_t i = 0; i < N; i++) {
        const double p0 = p0c[i];
        if (p0 == 0) return set_error(BMX_ERR_ARG, "bmx_beam_h5_read: reference momentum not present");
        if (std::fabs(p0 - p0c[0]) > 1e-12 * std::fabs(p0c[0]))
            return set_error(BMX_ERR_ARG, "bmx_beam_h5_read: particles with different reference momenta");
        const double vx = x[i] + xo[i], vy = y[i] + yo[i];
        const double mx = px[i] + pxo[i], my = py[i] + pyo[i], mz = pz[i] + pzo[i];
        int state = (int)st[i];
        if (state != BMX_STATE_ALIVE && state != 0 && !bmad) state = BMX_STATE_LOST;
        double rp;  // pz
        if (!has_tot) rp = (std::sqrt(mx * mx + my * my + mz * mz) - p0) / p0;
        else rp = tot[i] / p0;
        // convert_pc_to((1 + pz) p0c, species, beta)
        const double pc = (1 + rp) * p0;
        const double beta = pc / std::sqrt(pc * pc + mc2 * mc2);
        host->x[i] = vx;
        host->px[i] = mx / p0;
        host->y[i] = vy;
        host->py[i] = my / p0;
        host->z[i] = -beta * kCLight * dt[i];
        host->pz[i] = rp;
        host->beta[i] = beta;
        host->state[i] = state;
        host->loss_ele[i] = state == BMX_STATE_ALIVE ? 0 : (int32_t)ixe[i];
        host->loss_turn[i] = 0;
        sum_w += w[i];
    }
    (void)q_live;
    if (n_out) *n_out = n;
    if (p0c_out) *p0c_out = N ? p0c[0] : 0;
    if (charge_tot) *charge_tot = q_unit != 0 ? q_tot * q_unit : sum_w;
    return BMX_OK;
#endif
}

int32_t bmx_beam_h5_write(const char *path, const bmx_beam_view *host, int64_t n, const char *species,
                          double charge, double p0c, int32_t ix_ele, int32_t append)
{
#ifndef BMX_HAVE_HDF5
    (void)path; (void)host; (void)n; (void)species; (void)charge; (void)p0c; (void)ix_ele; (void)append;
    return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_write: built without HDF5 headers");
#else
    if (!path || !host || n < 1 || !species || !(p0c > 0))
        return set_error(BMX_ERR_ARG, "bmx_beam_h5_write: bad argument");
    std::string sp(species);
    for (char &c : sp) c = (char)std::tolower((unsigned char)c);
    bool known = false;
    for (const SpeciesMass &s : kSpecies) known |= sp == s.name;
    if (!known) return set_error(BMX_ERR_ARG, "bmx_beam_h5_write: species '" + sp + "' is not a charged species of this path");
    std::lock_guard<std::mutex> lock(g_h5_mutex);
    const H5Api *ap = h5();
    if (!ap) return set_error(BMX_ERR_UNSUPPORTED, "bmx_beam_h5_write: " + g_api.why);
    const H5Api &a = *ap;
    // the file and its header (hdf5_write_beam.f90:40-74)
    char date[32];
    {
        const std::time_t now = std::time(nullptr);
        std::tm tm{};
        localtime_r(&now, &tm);
        std::strftime(date, sizeof date, "%Y-%m-%d %H:%M:%S", &tm);
    }
    hid_t f;
    if (append) {
        f = a.fopen(path, kAccRdwr, H5P_DEFAULT);
        if (f < 0) return set_error(BMX_ERR_ARG, std::string("bmx_beam_h5_write: cannot open ") + path);
        if (has_attr(a, f, "lastUpdated")) a.adelete(f, "lastUpdated");
        write_attr_string(a, f, "lastUpdated", date);
    } else {
        f = a.fcreate(path, kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
        if (f < 0) return set_error(BMX_ERR_ARG, std::string("bmx_beam_h5_write: cannot create ") + path);
        write_attr_string(a, f, "dataType", "openPMD");
        write_attr_string(a, f, "openPMD", "2.0.0");
        write_attr_string(a, f, "openPMDextension", "BeamPhysics;SpeciesType");
        write_attr_string(a, f, "basePath", "/data/%T/");
        write_attr_string(a, f, "particlesPath", "particles/");
        write_attr_string(a, f, "software", "bmadx-mi355x");
        write_attr_string(a, f, "softwareVersion", "1.0");
        write_attr_string(a, f, "date", date);
        const hid_t r = a.gcreate(f, "/data/", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        a.gclose(r);
    }
    const hid_t r = a.gopen(f, "/data/", H5P_DEFAULT);
    if (r < 0) {
        a.fclose(f);
        return set_error(BMX_ERR_ARG, "bmx_beam_h5_write: no /data group in the file");
    }
    // the first free bunch index (:82-88)
    int ib = 0;
    char bname[16];
    do {
        ib++;
        std::snprintf(bname, sizeof bname, "%05d", ib);
    } while (a.lexists(r, bname, H5P_DEFAULT) > 0);
    const hid_t b = a.gcreate(r, bname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const hid_t b1 = a.gcreate(b, "particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const hid_t b2 = a.gcreate(b1, sp.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const size_t N = (size_t)n;
    double q_live = 0;
    for (size_t i = 0; i < N; i++)
        if (host->state[i] == BMX_STATE_ALIVE) q_live += charge;
    const double q_tot = charge * (double)N, one = 1.0;
    const int n32 = (int)n;
    write_attr_string(a, b2, "speciesType", sp);
    write_attr_doubles(a, b2, "totalCharge", &q_tot, 1);
    write_attr_doubles(a, b2, "chargeLive", &q_live, 1);
    write_attr_doubles(a, b2, "chargeUnitSI", &one, 1);
    write_attr_ints(a, b2, "numParticles", &n32, 1);
    // records (:125-219): momenta in eV/c, time t - t_ref = -z / (beta c)
    std::vector<double> v(N);
    hid_t z;
    z = a.gcreate(b2, "momentum", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    for (size_t i = 0; i < N; i++) v[i] = host->px[i] * p0c;
    write_real_record(a, z, "x", "px * p0c", kUnitEvPerC, v);
    for (size_t i = 0; i < N; i++) v[i] = host->py[i] * p0c;
    write_real_record(a, z, "y", "py * p0c", kUnitEvPerC, v);
    for (size_t i = 0; i < N; i++) {
        const double rp = 1 + host->pz[i];
        v[i] = std::sqrt(rp * rp - host->px[i] * host->px[i] - host->py[i] * host->py[i]) * p0c;
    }
    write_real_record(a, z, "z", "ps * p0c", kUnitEvPerC, v);
    a.gclose(z);
    for (size_t i = 0; i < N; i++) v[i] = host->pz[i] * p0c;
    write_real_record(a, b2, "totalMomentum", "pz * p0c", kUnitEvPerC, v);
    z = a.gcreate(b2, "spin", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    std::fill(v.begin(), v.end(), 0.0);
    write_real_record(a, z, "x", "Sx", kUnitHbar, v);
    write_real_record(a, z, "y", "Sy", kUnitHbar, v);
    write_real_record(a, z, "z", "Sz", kUnitHbar, v);
    a.gclose(z);
    std::fill(v.begin(), v.end(), p0c);
    write_real_record(a, b2, "totalMomentumOffset", "p0c", kUnitEvPerC, v);
    for (size_t i = 0; i < N; i++) v[i] = -host->z[i] / (host->beta[i] * kCLight);
    write_real_record(a, b2, "time", "t - t_ref", kUnitSec, v);
    // the absolute time is not carried (relative time tracking): t_ref = -(t - t_ref)
    for (size_t i = 0; i < N; i++) v[i] = 0.0 - v[i];
    write_real_record(a, b2, "timeOffset", "t_ref", kUnitSec, v);
    z = a.gcreate(b2, "position", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    for (size_t i = 0; i < N; i++) v[i] = host->x[i];
    write_real_record(a, z, "x", "x", kUnitM, v);
    for (size_t i = 0; i < N; i++) v[i] = host->y[i];
    write_real_record(a, z, "y", "y", kUnitM, v);
    std::fill(v.begin(), v.end(), 0.0);
    write_real_record(a, z, "z", "z", kUnitM, v);
    a.gclose(z);
    write_real_record(a, b2, "sPosition", "s", kUnitM, v);
    std::fill(v.begin(), v.end(), charge);
    write_real_record(a, b2, "weight", "macro-charge", kUnit1, v);
    std::vector<int> iv(N);
    for (size_t i = 0; i < N; i++) iv[i] = host->state[i];
    write_int_record(a, b2, "particleStatus", "state", iv);
    std::fill(iv.begin(), iv.end(), 0);
    write_int_record(a, b2, "branchIndex", "ix_branch", iv);
    for (size_t i = 0; i < N; i++) iv[i] = host->state[i] == BMX_STATE_ALIVE ? ix_ele : host->loss_ele[i];
    write_int_record(a, b2, "elementIndex", "ix_ele", iv);
    std::fill(iv.begin(), iv.end(), 1);  // downstream_end
    write_int_record(a, b2, "locationInElement", "location", iv);
    a.gclose(b2);
    a.gclose(b1);
    a.gclose(b);
    a.gclose(r);
    a.fclose(f);
    return BMX_OK;
#endif
}

}  // extern "C"
