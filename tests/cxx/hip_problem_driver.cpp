// C++ driver of the C ABI, written the way the reference-side binding
// (Order2NewmarkHip, a Problem subclass selected in Problem::Factory,
// Problem.cpp:11-32, and driven from one branch in Main.cpp:45-66) calls it:
// every status code becomes a std::runtime_error carrying sem_last_error(),
// as the reference's own errors are exceptions caught in main
// (Main.cpp:70-74).
//
// The problem arrives as a binary file holding the arrays the binding
// extracts from the reference's DMPlex mesh and ExodusModel (cells in closure
// vertex order, vertex coordinates, per-element per-vertex material), plus
// the time step, step count, one Ricker source and receivers:
//   int32  dim, order, physics, ncells, nverts, nrec
//   int64  nsteps
//   f64    dt, src[3], amp, f0, t0, dir[3]
//   int32  cells[ncells * 2^dim]
//   f64    coords[nverts * dim], vertex_par[ncells * 2^dim * npar], rec[nrec * 3]
// It runs the whole loop (one sem_batch_run of every step, as the Main.cpp
// branch does), writes the seismograms (float32, nsteps x nrec x c) and the
// final displacement (float64, c x G) to <out>, and checks the error path:
// a source outside the mesh must come back as std::runtime_error.
//
// usage: hip_problem_driver <problem.bin> <out.bin>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
#include "salvus_sem.h"
}

namespace {

void check(int rc) {
  if (rc) throw std::runtime_error(sem_last_error());
}

template <class T>
void get(std::ifstream &f, T *p, size_t n) {
  f.read(reinterpret_cast<char *>(p), (std::streamsize)(n * sizeof(T)));
  if (!f) throw std::runtime_error("problem file truncated");
}

// The binding's state and calls (Order2NewmarkHip): mesh and one shot.
class HipProblem {
  sem_mesh *mMesh = nullptr;
  sem_batch *mBatch = nullptr;
  int mShot = 0;
  double mDt = 0;

 public:
  explicit HipProblem(double dt) : mDt(dt) {}
  ~HipProblem() {
    sem_batch_destroy(mBatch);
    sem_mesh_destroy(mMesh);
  }
  void setup(int dim, int order, int physics, int64_t nc, int64_t nv, const int32_t *cells,
             const double *coords, const double *vertex_par, int64_t nsteps) {
    check(sem_mesh_create(&mMesh, 0, dim, order, physics, nc, nv, cells, coords, vertex_par,
                          SEM_DEFAULT_QUIRKS));
    check(sem_batch_create(&mBatch, 0));
    check(sem_batch_add_shot(mBatch, mMesh, nsteps, mDt, &mShot));
  }
  void addSource(const double *loc, double amp, double f0, double t0, const double *dir) {
    check(sem_shot_add_source(mBatch, mShot, loc, amp, f0, t0, dir));
  }
  void addReceiver(const double *loc) { check(sem_shot_add_receiver(mBatch, mShot, loc)); }
  void initializeGlobalDofs() { check(sem_batch_prepare(mBatch, SEM_PATH_FUSED, 0, 0)); }
  void run(int64_t nsteps) { check(sem_batch_run(mBatch, nsteps)); }
  int64_t info(int k) const {
    int64_t v[16];
    check(sem_mesh_info(mMesh, v));
    return v[k];
  }
  std::vector<float> seismograms(size_t n) {
    std::vector<float> s(n);
    check(sem_shot_get_seismograms(mBatch, mShot, s.data()));
    return s;
  }
  std::vector<double> field(int which, size_t n) {
    std::vector<double> u(n);
    check(sem_shot_get_field(mBatch, mShot, which, u.data()));
    return u;
  }
};

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::cerr << "usage: hip_problem_driver <problem.bin> <out.bin>\n";
    return 2;
  }
  try {
    std::ifstream f(argv[1], std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + argv[1]);
    int32_t hdr[6];
    get(f, hdr, 6);
    const int dim = hdr[0], order = hdr[1], physics = hdr[2], ncells = hdr[3], nverts = hdr[4],
              nrec = hdr[5];
    int64_t nsteps = 0;
    get(f, &nsteps, 1);
    double par[10];
    get(f, par, 10);
    const double dt = par[0];
    const int nv = dim == 2 ? 4 : 8;
    const int npar = physics == SEM_ACOUSTIC ? 1 : (dim == 2 ? 5 : 6);
    const int c = physics == SEM_ACOUSTIC ? 1 : dim;
    std::vector<int32_t> cells((size_t)ncells * nv);
    std::vector<double> coords((size_t)nverts * dim), vpar((size_t)ncells * nv * npar),
        rec((size_t)nrec * 3);
    get(f, cells.data(), cells.size());
    get(f, coords.data(), coords.size());
    get(f, vpar.data(), vpar.size());
    get(f, rec.data(), rec.size());

    HipProblem p(dt);
    p.setup(dim, order, physics, ncells, nverts, cells.data(), coords.data(), vpar.data(), nsteps);
    p.addSource(par + 1, par[4], par[5], par[6], par + 7);
    for (int r = 0; r < nrec; ++r) p.addReceiver(&rec[3 * r]);

    // the error path: a point outside the mesh is an exception
    bool thrown = false;
    try {
      const double far[3] = {-1e30, -1e30, -1e30};
      p.addReceiver(far);
    } catch (const std::runtime_error &e) {
      thrown = true;
      std::cout << "error path: " << e.what() << "\n";
    }
    if (!thrown) throw std::runtime_error("a receiver outside the mesh was accepted");

    p.initializeGlobalDofs();
    p.run(nsteps);
    const int64_t G = p.info(1);
    const std::vector<float> seis = p.seismograms((size_t)nsteps * nrec * c);
    const std::vector<double> u = p.field(SEM_FIELD_U, (size_t)c * G);
    std::ofstream o(argv[2], std::ios::binary);
    o.write(reinterpret_cast<const char *>(seis.data()), (std::streamsize)(seis.size() * 4));
    o.write(reinterpret_cast<const char *>(u.data()), (std::streamsize)(u.size() * 8));
    if (!o) throw std::runtime_error(std::string("cannot write ") + argv[2]);
    std::cout << "ran " << nsteps << " steps on " << ncells << " cells, G = " << G << "\n";
  } catch (const std::runtime_error &e) {
    std::cerr << "runtime error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
