// State files in the reference's format: the EpetraExt-HDF5 layout that
// Model::loadStateFromFile / Model::saveStateToFile read and write
// (src/utils/Model.C:15-160).
//   /State/Values       double [NumVectors][GlobalLength]   (the state, one vector)
//   /State/GlobalLength int scalar
//   /State/NumVectors   int scalar
//   /State/__type__     string[1] "Epetra_MultiVector"
//   /Parameters/<label> double scalar, one per THCM parameter, under the
//                       labels of THCM::int2par (THCM.C:1895-1937)
// The HDF5 C library is loaded at the first call (dlopen of libhdf5.so, else
// /opt/conda/lib/libhdf5.so, or the path in IEMIC_HDF5_LIB), so the library
// itself has no link-time dependency on it.  Host-only code.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "ctx.hpp"

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;

// HDF5 1.10 constants (H5Fpublic.h, H5Ppublic.h, H5Spublic.h, H5Epublic.h)
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0, kAll = 0, kErrDefault = 0;
constexpr int kScalar = 0;   // H5S_SCALAR

// par.F90 index (1..30) -> label of THCM::int2par
const char *const kLabels[31] = {nullptr,
                                 "AL_T", "Rayleigh-Number", "Vertical Ekman-Number", "Horizontal Ekman-Number",
                                 "Rossby-Number", "MIXP", "RESC", "SPL1", "Salinity Homotopy", "Solar Forcing",
                                 "Horizontal Peclet-Number", "Vertical Peclet-Number", "P_VC", "LAMB",
                                 "Salinity Forcing", "Wind Forcing", "Temperature Forcing", "Nonlinear Factor",
                                 "Combined Forcing", "ARCL", "NLES", "IFRICB", "CONT", "Energy", "ALPC", "CMPR",
                                 "Flux Perturbation", "Salinity Perturbation", "MKAP", "SPL2"};

struct H5 {
    void *so = nullptr;
    herr_t (*open)();
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    htri_t (*lexists)(hid_t, const char *, hid_t);
    hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    hid_t (*dget_space)(hid_t);
    herr_t (*dclose)(hid_t);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tclose)(hid_t);
    hid_t *native_double, *native_int, *c_s1;
};

H5 g_h5;
std::mutex g_h5_mu;   // load_h5 may be reached from several threads at once

template <typename F>
bool sym(void *so, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(so, name));
    return f != nullptr;
}

int load_h5()
{
    std::lock_guard<std::mutex> lock(g_h5_mu);
    if (g_h5.so) return IEMIC_OK;
    const char *cands[3] = {std::getenv("IEMIC_HDF5_LIB"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"};
    void *so = nullptr;
    for (const char *c : cands)
        if (c && (so = dlopen(c, RTLD_NOW | RTLD_LOCAL))) break;
    if (!so)
        return iemic::set_error(IEMIC_EUNSUPPORTED,
                                "state files: no HDF5 library found (libhdf5.so, /opt/conda/lib/libhdf5.so, "
                                "or set IEMIC_HDF5_LIB)");
    H5 h;
    h.so = so;
    bool ok = sym(so, "H5open", h.open) && sym(so, "H5Eset_auto2", h.eset_auto) && sym(so, "H5Fopen", h.fopen) &&
              sym(so, "H5Fcreate", h.fcreate) && sym(so, "H5Fclose", h.fclose) && sym(so, "H5Lexists", h.lexists) &&
              sym(so, "H5Gcreate2", h.gcreate) && sym(so, "H5Gclose", h.gclose) && sym(so, "H5Dopen2", h.dopen) &&
              sym(so, "H5Dcreate2", h.dcreate) && sym(so, "H5Dread", h.dread) && sym(so, "H5Dwrite", h.dwrite) &&
              sym(so, "H5Dget_space", h.dget_space) && sym(so, "H5Dclose", h.dclose) &&
              sym(so, "H5Sget_simple_extent_ndims", h.sget_ndims) &&
              sym(so, "H5Sget_simple_extent_dims", h.sget_dims) && sym(so, "H5Screate", h.screate) &&
              sym(so, "H5Screate_simple", h.screate_simple) && sym(so, "H5Sclose", h.sclose) &&
              sym(so, "H5Tcopy", h.tcopy) && sym(so, "H5Tset_size", h.tset_size) && sym(so, "H5Tclose", h.tclose) &&
              sym(so, "H5T_NATIVE_DOUBLE_g", h.native_double) && sym(so, "H5T_NATIVE_INT_g", h.native_int) &&
              sym(so, "H5T_C_S1_g", h.c_s1);
    if (!ok || h.open() < 0) {
        dlclose(so);
        return iemic::set_error(IEMIC_EUNSUPPORTED, "state files: the HDF5 library lacks a needed symbol");
    }
    h.eset_auto(kErrDefault, nullptr, nullptr);   // errors come back as codes, not printed stacks
    g_h5 = h;
    return IEMIC_OK;
}

// Reads the whole dataset `name` (below loc) of `count` elements of type t.
bool read_all(hid_t loc, const char *name, hid_t t, void *out, hsize_t count, hsize_t *found_count)
{
    H5 &h = g_h5;
    if (h.lexists(loc, name, kDefault) <= 0) return false;
    hid_t d = h.dopen(loc, name, kDefault);
    if (d < 0) return false;
    hid_t sp = h.dget_space(d);
    hsize_t dims[4] = {1, 1, 1, 1};
    int nd = sp >= 0 ? h.sget_ndims(sp) : -1;
    bool ok = nd >= 0 && nd <= 4 && h.sget_dims(sp, dims, nullptr) >= 0;
    hsize_t tot = 1;
    for (int i = 0; i < nd; i++) tot *= dims[i];
    if (found_count) *found_count = tot;
    ok = ok && tot == count && h.dread(d, t, kAll, kAll, kDefault, out) >= 0;
    if (sp >= 0) h.sclose(sp);
    h.dclose(d);
    return ok;
}

bool write_all(hid_t loc, const char *name, hid_t t, const void *data, int rank, const hsize_t *dims)
{
    H5 &h = g_h5;
    hid_t sp = rank == 0 ? h.screate(kScalar) : h.screate_simple(rank, dims, nullptr);
    if (sp < 0) return false;
    hid_t d = h.dcreate(loc, name, t, sp, kDefault, kDefault, kDefault);
    bool ok = d >= 0 && h.dwrite(d, t, kAll, kAll, kDefault, data) >= 0;
    if (d >= 0) h.dclose(d);
    h.sclose(sp);
    return ok;
}

}  // namespace

using iemic::set_error;

// Model::loadStateFromFile (Model.C:15-110): the state and the parameters
// that the file holds (a parameter missing from the file is skipped, as the
// reference does).  par[1..30] receive the values, found[1..30] 1 where the
// file holds the parameter (both optional, 31 entries each).
extern "C" int iemic_read_state(const char *path, double *x, int64_t n, double *par, int *found)
{
    if (!path || !x || n <= 0) return set_error(IEMIC_EINVAL, "iemic_read_state: null path or state, or n <= 0");
    if (int rc = load_h5()) return rc;
    H5 &h = g_h5;
    hid_t f = h.fopen(path, kAccRdonly, kDefault);
    if (f < 0) return set_error(IEMIC_EIO, std::string("iemic_read_state: cannot open ") + path);
    int rc = IEMIC_OK;
    int glen = 0, nvec = 0;
    hsize_t cnt = 0;
    if (h.lexists(f, "State", kDefault) <= 0) {
        rc = set_error(IEMIC_EIO, std::string("iemic_read_state: no group <State> in ") + path);
    } else if (!read_all(f, "State/GlobalLength", *h.native_int, &glen, 1, nullptr) ||
               !read_all(f, "State/NumVectors", *h.native_int, &nvec, 1, nullptr)) {
        rc = set_error(IEMIC_EIO, "iemic_read_state: <State> lacks GlobalLength or NumVectors");
    } else if ((int64_t)glen != n || nvec != 1) {
        rc = set_error(IEMIC_EINVAL, "iemic_read_state: the file's state has " + std::to_string(glen) + " x " +
                                         std::to_string(nvec) + " values, expected " + std::to_string(n) + " x 1");
    } else if (!read_all(f, "State/Values", *h.native_double, x, (hsize_t)n, &cnt)) {
        rc = set_error(IEMIC_EIO, "iemic_read_state: cannot read State/Values");
    }
    if (rc == IEMIC_OK && (par || found)) {
        // Model::loadStateFromFile stops with an error when the group is
        // absent (Model.C:87-90); a missing single parameter is skipped
        const bool has = h.lexists(f, "Parameters", kDefault) > 0;
        if (!has) rc = set_error(IEMIC_EIO, std::string("iemic_read_state: no group <Parameters> in ") + path);
        for (int i = 1; has && i <= 30; i++) {
            double v = 0.0;
            const std::string nm = std::string("Parameters/") + kLabels[i];
            const bool ok = has && read_all(f, nm.c_str(), *h.native_double, &v, 1, nullptr);
            if (par && ok) par[i] = v;
            if (found) found[i] = ok ? 1 : 0;
        }
    }
    h.fclose(f);
    return rc;
}

// Model::saveStateToFile (Model.C:112-160): the state and the 30
// parameters (par[1..30]).  The file is created or truncated.  The grid
// group the reference adds for its plotting tools is not written.
extern "C" int iemic_write_state(const char *path, const double *x, int64_t n, const double *par)
{
    if (!path || !x || n <= 0 || n > INT32_MAX)
        return set_error(IEMIC_EINVAL, "iemic_write_state: null path or state, or n out of range");
    if (int rc = load_h5()) return rc;
    H5 &h = g_h5;
    hid_t f = h.fcreate(path, kAccTrunc, kDefault, kDefault);
    if (f < 0) return set_error(IEMIC_EIO, std::string("iemic_write_state: cannot create ") + path);
    bool ok = true;
    hid_t g = h.gcreate(f, "State", kDefault, kDefault, kDefault);
    ok = g >= 0;
    if (ok) {
        const hsize_t dv[2] = {1, (hsize_t)n};
        const int glen = (int)n, nvec = 1;
        ok = write_all(g, "Values", *h.native_double, x, 2, dv) &&
             write_all(g, "GlobalLength", *h.native_int, &glen, 0, nullptr) &&
             write_all(g, "NumVectors", *h.native_int, &nvec, 0, nullptr);
        const char type_name[19] = "Epetra_MultiVector";
        hid_t st = h.tcopy(*h.c_s1);
        const hsize_t d1[1] = {1};
        ok = ok && st >= 0 && h.tset_size(st, sizeof(type_name)) >= 0 && write_all(g, "__type__", st, type_name, 1, d1);
        if (st >= 0) h.tclose(st);
        h.gclose(g);
    }
    if (ok && par) {
        hid_t gp = h.gcreate(f, "Parameters", kDefault, kDefault, kDefault);
        ok = gp >= 0;
        for (int i = 1; ok && i <= 30; i++) ok = write_all(gp, kLabels[i], *h.native_double, &par[i], 0, nullptr);
        if (gp >= 0) h.gclose(gp);
    }
    h.fclose(f);
    return ok ? IEMIC_OK : set_error(IEMIC_EIO, std::string("iemic_write_state: writing ") + path + " failed");
}
