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
//   movies     the mesh first (Problem.cpp:133-137 opens the viewer and
//              DMViews the mesh before any frame): /geometry/vertices
//              [V][dim] FP64, /topology/cells [E][2^dim] int32 with the
//              attribute cell_dim, the names of PETSc's DMPlex HDF5 view;
//              and /geometry/nodes [G][dim] FP64, the coordinates of the
//              global DOFs the frames hold (the GLL nodes in DMPlex DOF
//              order), so that a reader can place each value.  Then
//              Problem::saveSolution (Problem.cpp:217-237), called after
//              every step by the driver (Main.cpp:55-62): each named field
//              of the frame goes to an extendible FP64 dataset
//              /fields/<name> [frames][n], and the frame's time (the time
//              after the step, the value DMSetOutputSequenceNumber records)
//              to /time [frames].  This follows PETSc's time-stepped
//              VecView layout for DMPlex vectors in outline; PETSc itself is
//              not available to check it against.
// libhdf5 is opened at run time (SEM_HDF5_LIB, else the usual names), so
// the library needs it only when these entry points are called.
#include <dlfcn.h>
#include <unistd.h>

#include <algorithm>
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
constexpr hsize_t kUnlimited = ~0ull;  // H5S_UNLIMITED
constexpr int kSelectSet = 0;          // H5S_SELECT_SET

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
  hid_t (*pcreate)(hid_t) = nullptr;
  herr_t (*pset_chunk)(hid_t, int, const hsize_t *) = nullptr;
  herr_t (*pclose)(hid_t) = nullptr;
  herr_t (*dset_extent)(hid_t, const hsize_t *) = nullptr;
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *,
                              const hsize_t *) = nullptr;
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *) = nullptr;
  hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*awrite)(hid_t, hid_t, const void *) = nullptr;
  herr_t (*aclose)(hid_t) = nullptr;
  const hid_t *native_float = nullptr, *native_double = nullptr, *native_int = nullptr,
              *dcpl_class = nullptr;
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
            sym(lib, "H5Pcreate", h.pcreate) && sym(lib, "H5Pset_chunk", h.pset_chunk) &&
            sym(lib, "H5Pclose", h.pclose) && sym(lib, "H5Dset_extent", h.dset_extent) &&
            sym(lib, "H5Sselect_hyperslab", h.sselect_hyperslab) &&
            sym(lib, "H5Sget_simple_extent_dims", h.sget_dims) &&
            sym(lib, "H5P_CLS_DATASET_CREATE_ID_g", h.dcpl_class) &&
            sym(lib, "H5T_NATIVE_FLOAT_g", h.native_float) &&
            sym(lib, "H5T_NATIVE_DOUBLE_g", h.native_double) &&
            sym(lib, "H5T_NATIVE_INT_g", h.native_int) && sym(lib, "H5Acreate2", h.acreate) &&
            sym(lib, "H5Awrite", h.awrite) && sym(lib, "H5Aclose", h.aclose);
  // hid_t is 64-bit from HDF5 1.10 on: check the version before any
  // hid_t global (H5T_NATIVE_*_g) is read
  int (*libversion)(unsigned *, unsigned *, unsigned *) = nullptr;
  unsigned vmaj = 0, vmin = 0, vrel = 0;
  if (ok && (!sym(lib, "H5get_libversion", libversion) || libversion(&vmaj, &vmin, &vrel) < 0 ||
             vmaj < 1 || (vmaj == 1 && vmin < 10))) {
    dlclose(lib);
    sem::set_error("libhdf5 " + std::to_string(vmaj) + "." + std::to_string(vmin) + "." +
                   std::to_string(vrel) + " is too old (the HDF5 >= 1.10 API is required)");
    return nullptr;
  }
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

// Rows already in the extendible dataset `name` ([rows][width], or [rows]
// when width == 0), -1 if it does not exist, -2 on a shape mismatch.
int64_t rows_of(const H5 *h, hid_t f, const std::string &name, int64_t width) {
  if (h->lexists(f, name.c_str(), kDefault) <= 0) return -1;
  const hid_t ds = h->dopen(f, name.c_str(), kDefault);
  if (ds < 0) return -2;
  const hid_t sp = h->dspace(ds);
  hsize_t dims[2] = {0, 0};
  const int rank = sp < 0 ? -1 : h->sget_dims(sp, dims, nullptr);
  if (sp >= 0) h->sclose(sp);
  h->dclose(ds);
  if (rank != (width > 0 ? 2 : 1) || (width > 0 && dims[1] != (hsize_t)width)) return -2;
  return (int64_t)dims[0];
}

// Appends one row (width doubles, or one double when width == 0) to the
// extendible dataset `name`, creating it ([0][width], chunks of one row)
// when absent.
bool append_row(const H5 *h, hid_t f, const std::string &name, int64_t width, const double *row) {
  const int rank = width > 0 ? 2 : 1;
  const hsize_t w = width > 0 ? (hsize_t)width : 1;
  int64_t rows = rows_of(h, f, name, width);
  if (rows == -2) return false;
  hid_t ds;
  if (rows < 0) {
    const hsize_t dims[2] = {0, w}, maxd[2] = {kUnlimited, w};
    const hsize_t chunk[2] = {width > 0 ? 1ull : 256ull, w};
    const hid_t sp = h->screate_simple(rank, dims, maxd);
    const hid_t pl = h->pcreate(*h->dcpl_class);
    const bool ok = sp >= 0 && pl >= 0 && h->pset_chunk(pl, rank, chunk) >= 0;
    ds = ok ? h->dcreate(f, name.c_str(), *h->native_double, sp, kDefault, pl, kDefault) : -1;
    if (pl >= 0) h->pclose(pl);
    if (sp >= 0) h->sclose(sp);
    rows = 0;
  } else {
    ds = h->dopen(f, name.c_str(), kDefault);
  }
  if (ds < 0) return false;
  const hsize_t ext[2] = {(hsize_t)rows + 1, w}, start[2] = {(hsize_t)rows, 0}, count[2] = {1, w};
  bool ok = h->dset_extent(ds, ext) >= 0;
  const hid_t fs = ok ? h->dspace(ds) : -1;
  const hid_t ms = fs >= 0 ? h->screate_simple(1, &w, nullptr) : -1;
  ok = ok && ms >= 0 && h->sselect_hyperslab(fs, kSelectSet, start, nullptr, count, nullptr) >= 0 &&
       h->dwrite(ds, *h->native_double, ms, fs, kDefault, row) >= 0;
  if (ms >= 0) h->sclose(ms);
  if (fs >= 0) h->sclose(fs);
  h->dclose(ds);
  return ok;
}

// Writes a fixed 2-D dataset [rows][cols] of `type` at `name`.
bool write_2d(const H5 *h, hid_t f, const char *name, hid_t type, int64_t rows, int64_t cols,
              const void *data) {
  const hsize_t dims[2] = {(hsize_t)rows, (hsize_t)cols};
  const hid_t sp = h->screate_simple(2, dims, nullptr);
  const hid_t ds = sp < 0 ? -1 : h->dcreate(f, name, type, sp, kDefault, kDefault, kDefault);
  const bool ok = ds >= 0 && h->dwrite(ds, type, kDefault, kDefault, kDefault, data) >= 0;
  if (ds >= 0) h->dclose(ds);
  if (sp >= 0) h->sclose(sp);
  return ok;
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

int sem_movie_write_mesh_h5(const char *path, int dim, int64_t num_verts, const double *coords,
                            int64_t num_cells, const int32_t *cells, int64_t num_nodes,
                            const double *node_xyz) {
  if (!path || (dim != 2 && dim != 3) || num_verts < 1 || !coords || num_cells < 1 || !cells ||
      num_nodes < 1 || !node_xyz) {
    sem::set_error("movie_write_mesh_h5: bad arguments");
    return -1;
  }
  const int nv = dim == 2 ? 4 : 8;
  for (int64_t k = 0; k < num_cells * nv; ++k)
    if (cells[k] < 0 || cells[k] >= num_verts) {
      sem::set_error("movie_write_mesh_h5: cell vertex out of range");
      return -1;
    }
  const H5 *h = hdf5();
  if (!h) return -2;
  const hid_t f = h->fcreate(path, kAccTrunc, kDefault, kDefault);
  if (f < 0) { sem::set_error(std::string("cannot create ") + path); return -2; }
  int rc = 0;
  for (const char *g : {"geometry", "topology"}) {
    const hid_t gid = h->gcreate(f, g, kDefault, kDefault, kDefault);
    if (gid < 0) rc = -2;
    else h->gclose(gid);
  }
  if (!rc && !(write_2d(h, f, "geometry/vertices", *h->native_double, num_verts, dim, coords) &&
               write_2d(h, f, "topology/cells", *h->native_int, num_cells, nv, cells) &&
               write_2d(h, f, "geometry/nodes", *h->native_double, num_nodes, dim, node_xyz)))
    rc = -2;
  if (!rc) {  // cell_dim on /topology/cells
    const hid_t ds = h->dopen(f, "topology/cells", kDefault);
    const hsize_t one = 1;
    const hid_t sp = h->screate_simple(1, &one, nullptr);
    const hid_t at = (ds < 0 || sp < 0) ? -1 : h->acreate(ds, "cell_dim", *h->native_int, sp, kDefault, kDefault);
    const int32_t cd = dim;
    if (at < 0 || h->awrite(at, *h->native_int, &cd) < 0) rc = -2;
    if (at >= 0) h->aclose(at);
    if (sp >= 0) h->sclose(sp);
    if (ds >= 0) h->dclose(ds);
  }
  if (rc) sem::set_error(std::string("writing the mesh to ") + path + " failed");
  if (h->fclose(f) < 0 && !rc) { sem::set_error("H5Fclose failed"); rc = -2; }
  return rc;
}

int sem_movie_write_h5(const char *path, int append, int nfields, const char *const *names,
                       int64_t n, const double *values, double time) {
  if (!path || nfields < 1 || !names || n < 1 || !values) {
    sem::set_error("movie_write_h5: bad arguments");
    return -1;
  }
  for (int k = 0; k < nfields; ++k)
    if (!plain_name(names[k])) { sem::set_error("field names must be non-empty and contain no '/'"); return -1; }
  const H5 *h = hdf5();
  if (!h) return -2;
  hid_t f = (append && access(path, F_OK) == 0) ? h->fopen(path, kAccRdwr, kDefault)
                                                : h->fcreate(path, kAccTrunc, kDefault, kDefault);
  if (f < 0) { sem::set_error(std::string("cannot open ") + path); return -2; }
  int rc = 0;
  // every dataset of the frame must hold the same number of frames
  const int64_t frames = std::max<int64_t>(rows_of(h, f, "time", 0), 0);
  if (rows_of(h, f, "time", 0) == -2) rc = -1;
  for (int k = 0; k < nfields && !rc; ++k) {
    const int64_t r = rows_of(h, f, std::string("fields/") + names[k], n);
    if (r == -2 || (r >= 0 && r != frames) || (r < 0 && frames > 0)) rc = -1;
  }
  if (rc) sem::set_error("movie file datasets do not match this frame (field set, length or frame count)");
  if (!rc && h->lexists(f, "fields", kDefault) <= 0) {
    const hid_t g = h->gcreate(f, "fields", kDefault, kDefault, kDefault);
    if (g < 0) { sem::set_error("H5Gcreate failed"); rc = -2; }
    else h->gclose(g);
  }
  for (int k = 0; k < nfields && !rc; ++k)
    if (!append_row(h, f, std::string("fields/") + names[k], n, values + (size_t)k * n)) {
      sem::set_error(std::string("writing /fields/") + names[k] + " failed");
      rc = -2;
    }
  if (!rc && !append_row(h, f, "time", 0, &time)) { sem::set_error("writing /time failed"); rc = -2; }
  if (h->fclose(f) < 0 && !rc) { sem::set_error("H5Fclose failed"); rc = -2; }
  return rc;
}

}  // extern "C"
