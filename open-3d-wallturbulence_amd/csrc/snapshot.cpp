// Snapshot files in the reference's format (save_instant_field and WRITE_3D,
// src/IO.cpp:101-157, 198-283): one HDF5 file per snapshot holding 3-D double
// datasets. The serial HDF5 C library (1.10 or later, where hid_t is 64-bit) is
// loaded at run time, so libowt has no link-time HDF5 dependency; when it cannot
// be loaded, saving is an I/O error and nothing else is affected.
//
// Library lookup: $OWT_HDF5_LIB, then libhdf5.so / libhdf5.so.103 on the loader
// path, then /opt/conda/lib/libhdf5.so.
#include "snapshot.hpp"

#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <mutex>

#include "owt_internal.hpp"

namespace owt {
namespace {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;       // H5P_DEFAULT and H5S_ALL
constexpr unsigned kAccTrunc = 2u;  // H5F_ACC_TRUNC

struct Hdf5 {
    herr_t (*open)() = nullptr;
    herr_t (*libversion)(unsigned *, unsigned *, unsigned *) = nullptr;
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t) = nullptr;
    herr_t (*fclose)(hid_t) = nullptr;
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *) = nullptr;
    herr_t (*sclose)(hid_t) = nullptr;
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *) = nullptr;
    herr_t (*dclose)(hid_t) = nullptr;
    const hid_t *native_double = nullptr;  // H5T_NATIVE_DOUBLE, valid after H5open
    std::string err;
};

template <typename T>
bool bind(void *lib, const char *name, T &fn) {
    fn = reinterpret_cast<T>(dlsym(lib, name));
    return fn != nullptr;
}

Hdf5 load_hdf5() {
    Hdf5 h;
    const char *env = std::getenv("OWT_HDF5_LIB");
    const char *names[] = {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    void *lib = nullptr;
    for (const char *n : names)
        if (n && *n && (lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!lib) {
        h.err = "HDF5 C library not found (set OWT_HDF5_LIB to libhdf5.so)";
        return h;
    }
    // the library stays loaded for the life of the process (it registers exit handlers)
    bool ok = bind(lib, "H5open", h.open) && bind(lib, "H5get_libversion", h.libversion) &&
              bind(lib, "H5Fcreate", h.fcreate) && bind(lib, "H5Fclose", h.fclose) &&
              bind(lib, "H5Screate_simple", h.screate_simple) && bind(lib, "H5Sclose", h.sclose) &&
              bind(lib, "H5Dcreate2", h.dcreate2) && bind(lib, "H5Dwrite", h.dwrite) && bind(lib, "H5Dclose", h.dclose);
    h.native_double = reinterpret_cast<const hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
    if (!ok || !h.native_double) {
        h.err = "HDF5 C library lacks a required symbol";
        return h;
    }
    unsigned maj = 0, min = 0, rel = 0;
    if (h.open() < 0 || h.libversion(&maj, &min, &rel) < 0 || maj != 1 || min < 10) {
        h.err = "HDF5 C library 1.10 or later required (found " + std::to_string(maj) + "." + std::to_string(min) + ")";
        return h;
    }
    return h;
}

const Hdf5 &hdf5() {
    static std::once_flag once;
    static Hdf5 h;
    std::call_once(once, [] { h = load_hdf5(); });
    if (!h.err.empty()) throw Error(5, h.err);
    return h;
}

}  // namespace

// FILE_CREATE (src/IO.cpp:101-116): create or truncate
H5Writer::H5Writer(const std::string &path) : path_(path) {
    const Hdf5 &h = hdf5();
    file_ = h.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
    if (file_ < 0) throw Error(5, "cannot create " + path);
}

H5Writer::~H5Writer() {
    if (file_ >= 0) (void)hdf5().fclose(file_);
}

// WRITE_3D (src/IO.cpp:118-157): a [d0][d1][d2] double dataset, written whole
void H5Writer::write(const std::string &name, const long dims[3], const double *data) {
    const Hdf5 &h = hdf5();
    const hsize_t d[3] = {(hsize_t)dims[0], (hsize_t)dims[1], (hsize_t)dims[2]};
    const hid_t space = h.screate_simple(3, d, nullptr);
    if (space < 0) throw Error(5, "HDF5 dataspace for " + name);
    const hid_t dset = h.dcreate2(file_, name.c_str(), *h.native_double, space, kDefault, kDefault, kDefault);
    (void)h.sclose(space);
    if (dset < 0) throw Error(5, "cannot create dataset " + name + " in " + path_);
    const herr_t rc = h.dwrite(dset, *h.native_double, kDefault, kDefault, kDefault, data);
    (void)h.dclose(dset);
    if (rc < 0) throw Error(5, "cannot write dataset " + name + " in " + path_);
}

void H5Writer::close() {
    if (file_ < 0) return;
    const herr_t rc = hdf5().fclose(file_);
    file_ = -1;
    if (rc < 0) throw Error(5, "cannot close " + path_);
}

}  // namespace owt
