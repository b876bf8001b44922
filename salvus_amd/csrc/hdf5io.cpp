// The reference's HDF5 file formats on either side of the time loop:
//   receivers  ReceiverHdf5::write (ReceiverHdf5.cpp:31-87): one group per
//              receiver name, one 1-D FP32 dataset per field
//              (/<receiver>/<field>, n_samples long); fields are the
//              physics' PullElementalFields names (Scalar.cpp:28 "u",
//              Elastic2D.cpp:40 "ux","uy", Elastic3D.cpp:57 "ux","uy","uz")
//   sources    SourceHdf5::loadData (SourceHdf5.cpp:40-50): dataset
//              "<source>/data" read as nsteps x ncomp doubles into a
//              column-major matrix (component k's series is the k-th run of
//              nsteps values); fire() returns row `time_idx`
//              (SourceHdf5.cpp:31-36).
// libhdf5 is opened at run time (SEM_HDF5_LIB, else the usual names), so
// the library needs it only when these entry points are called.
#include <dlfcn.h>
#include <unistd.h>

#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "sem_internal.h"

namespace {

using hid_t = int64_t;  // HDF5 >= 1.10
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
using hssize_t = long long;
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0;  // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT

struct H5 {
  void *lib = nullptr;
  herr_t (*open)() = nullptr;
  herr_t (*eset_auto)(hid_t, void *, void *) = nullptr;
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t) = nullptr;
  hid_t (*fopen)(const char *, unsigned, hid_t) = nullptr;
  herr_t (*fclose)(hid_t) = nullptr;
  hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*gclose)(hid_t) = nullptr;
  htri_t (*lexists)(hid_t, const char *, hid_t) = nullptr;
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *) = nullptr;
  herr_t (*sclose)(hid_t) = nullptr;
  hssize_t (*snpoints)(hid_t) = nullptr;
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
  hid_t (*dopen)(hid_t, const char *, hid_t) = nullptr;
  hid_t (*dspace)(hid_t) = nullptr;
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *) = nullptr;
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *) = nullptr;
  herr_t (*dclose)(hid_t) = nullptr;
  const hid_t *native_float = nullptr, *native_double = nullptr;
};

template <class F>
bool sym(void *lib, const char *name, F &f) {
  f = reinterpret_cast<F>(dlsym(lib, name));
  return f != nullptr;
}

// Loaded once per process; returns nullptr (with the error set) when no
// usable libhdf5 is found.
const H5 *hdf5() {
  static H5 h;
  static bool tried = false;
  if (tried) {
    if (!h.lib) sem::set_error("libhdf5 not available (set SEM_HDF5_LIB)");
    return h.lib ? &h : nullptr;
  }
  tried = true;
  std::vector<std::string> names;
  if (const char *env = getenv("SEM_HDF5_LIB")) names.push_back(env);
  for (const char *n : {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
                        "/opt/conda/lib/libhdf5.so.103"})
    names.push_back(n);
  void *lib = nullptr;
  for (auto &n : names)
    if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!lib) {
    sem::set_error("libhdf5 not available (set SEM_HDF5_LIB)");
    return nullptr;
  }
  bool ok = sym(lib, "H5open", h.open) && sym(lib, "H5Eset_auto2", h.eset_auto) &&
            sym(lib, "H5Fcreate", h.fcreate) && sym(lib, "H5Fopen", h.fopen) &&
            sym(lib, "H5Fclose", h.fclose) && sym(lib, "H5Gcreate2", h.gcreate) &&
            sym(lib, "H5Gclose", h.gclose) && sym(lib, "H5Lexists", h.lexists) &&
            sym(lib, "H5Screate_simple", h.screate_simple) && sym(lib, "H5Sclose", h.sclose) &&
            sym(lib, "H5Sget_simple_extent_npoints", h.snpoints) &&
            sym(lib, "H5Dcreate2", h.dcreate) && sym(lib, "H5Dopen2", h.dopen) &&
            sym(lib, "H5Dget_space", h.dspace) && sym(lib, "H5Dwrite", h.dwrite) &&
            sym(lib, "H5Dread", h.dread) && sym(lib, "H5Dclose", h.dclose) &&
            sym(lib, "H5T_NATIVE_FLOAT_g", h.native_float) &&
            sym(lib, "H5T_NATIVE_DOUBLE_g", h.native_double);
  if (!ok || h.open() < 0) {
    dlclose(lib);
    sem::set_error("libhdf5 found but unusable (needs the HDF5 >= 1.10 API)");
    return nullptr;
  }
  h.eset_auto(kDefault, nullptr, nullptr);  // errors come back as return codes
  h.lib = lib;
  return &h;
}

bool plain_name(const char *s) {
  if (!s || !*s) return false;
  for (; *s; ++s)
    if (*s == '/') return false;
  return true;
}

}  // namespace

extern "C" {

int sem_receivers_write_h5(const char *path, int append, int nrec, const char *const *names,
                           int ncomp, const char *const *fields, int64_t nsamples,
                           const float *data) {
  if (!path || nrec < 0 || ncomp < 1 || nsamples < 0 || (nrec > 0 && (!names || !fields || !data))) {
    sem::set_error("receivers_write_h5: bad arguments");
    return -1;
  }
  for (int r = 0; r < nrec; ++r)
    if (!plain_name(names[r])) { sem::set_error("receiver names must be non-empty and contain no '/'"); return -1; }
  for (int k = 0; k < ncomp; ++k)
    if (!plain_name(fields[k])) { sem::set_error("field names must be non-empty and contain no '/'"); return -1; }
  const H5 *h = hdf5();
  if (!h) return -2;
  hid_t f = (append && access(path, F_OK) == 0) ? h->fopen(path, kAccRdwr, kDefault)
                                                : h->fcreate(path, kAccTrunc, kDefault, kDefault);
  if (f < 0) { sem::set_error(std::string("cannot open ") + path); return -2; }
  int rc = 0;
  std::vector<float> series((size_t)nsamples);
  const hsize_t dims = (hsize_t)nsamples;
  for (int r = 0; r < nrec && !rc; ++r) {
    if (h->lexists(f, names[r], kDefault) > 0) {
      sem::set_error(std::string("receiver '") + names[r] + "' already in the file");
      rc = -1;
      break;
    }
    const hid_t g = h->gcreate(f, names[r], kDefault, kDefault, kDefault);
    if (g < 0) { sem::set_error("H5Gcreate failed"); rc = -2; break; }
    h->gclose(g);
    for (int k = 0; k < ncomp && !rc; ++k) {
      for (int64_t t = 0; t < nsamples; ++t) series[t] = data[((size_t)t * nrec + r) * ncomp + k];
      const std::string full = std::string("/") + names[r] + "/" + fields[k];
      const hid_t sp = h->screate_simple(1, &dims, nullptr);
      const hid_t ds = sp < 0 ? -1 : h->dcreate(f, full.c_str(), *h->native_float, sp, kDefault, kDefault, kDefault);
      if (ds < 0 || h->dwrite(ds, *h->native_float, kDefault, kDefault, kDefault, series.data()) < 0) {
        sem::set_error("writing " + full + " failed");
        rc = -2;
      }
      if (ds >= 0) h->dclose(ds);
      if (sp >= 0) h->sclose(sp);
    }
  }
  if (h->fclose(f) < 0 && !rc) { sem::set_error("H5Fclose failed"); rc = -2; }
  return rc;
}

int sem_source_read_h5(const char *path, const char *name, int64_t nsteps, int ncomp, double *out) {
  if (!path || !name || !*name || nsteps < 1 || ncomp < 1 || !out) {
    sem::set_error("source_read_h5: bad arguments");
    return -1;
  }
  const H5 *h = hdf5();
  if (!h) return -2;
  const hid_t f = h->fopen(path, kAccRdonly, kDefault);
  if (f < 0) { sem::set_error(std::string("cannot open ") + path); return -2; }
  const std::string ds_name = std::string(name) + "/data";
  int rc = 0;
  const hid_t ds = h->dopen(f, ds_name.c_str(), kDefault);
  if (ds < 0) {
    sem::set_error("Can't read data for source '" + std::string(name) + "' from file '" + path + "'.");
    rc = -1;
  } else {
    const hid_t sp = h->dspace(ds);
    const hssize_t npts = sp < 0 ? -1 : h->snpoints(sp);
    if (sp >= 0) h->sclose(sp);
    if (npts != (hssize_t)(nsteps * ncomp)) {
      sem::set_error("source '" + std::string(name) + "' holds " + std::to_string(npts) +
                     " values, expected nsteps x ncomp = " + std::to_string(nsteps * ncomp));
      rc = -1;
    } else if (h->dread(ds, *h->native_double, kDefault, kDefault, kDefault, out) < 0) {
      sem::set_error("Can't read data for source '" + std::string(name) + "' from file '" + path + "'.");
      rc = -2;
    }
    h->dclose(ds);
  }
  h->fclose(f);
  return rc;
}

}  // extern "C"
