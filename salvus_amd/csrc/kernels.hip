// HIP kernels of the MI355X SEM path (gfx950).
//
// Element operator (per element, order N, n = N+1, nd = n^dim):
//   g_a(q)   = sum_m D[q_a][m] u(q with q_a <- m)               reference gradient
//   F_a(q)   = w_q detJ sum_c Jinv[c][a] sigma_c(q)              J^-T flux
//   (K u)(p) = sum_a sum_m D[m][p_a] F_a(p with p_a <- m)        test-and-integrate
// which is TensorQuad.cpp:176-210/:350-397 and Hexahedra.cpp:579-754 with
// the reference's quadrature; for acoustics F = w vp^2 Gsym g with
// Gsym = detJ Jinv^T Jinv (Scalar.cpp:42-71).  Stresses: Elastic2D.cpp:53-82
// (C11,C13,C33,C55; C12 = C23 = 0), Elastic3D.cpp:33-54 / :70-118 (TI).
//
// One thread per element node; an element's field and fluxes are staged in
// LDS; the 1-D derivative rows/columns live in registers.  Stiffness
// arithmetic is FP32 with FP32 accumulation.  Geometry and mass are
// computed in FP64 on the device and stored FP32.
#include <hip/hip_runtime.h>

#include <cstring>

#include "sem_internal.h"

namespace sem {

// ------------------------------------------------------------------ tables
struct DevTables {
  float D[kMaxN * kMaxN];  // D[i*n+m] = l'_m(xi_i)
  float w[kMaxN];
  float xi[kMaxN];
  double wd[kMaxN];
  double xid[kMaxN];
};
__constant__ DevTables c_tab[kMaxOrder2D + 1];

template <int DIM, int N>
struct Sh {
  static constexpr int n = N + 1;
  static constexpr int ND = DIM == 2 ? n * n : n * n * n;
  static constexpr int NV = DIM == 2 ? 4 : 8;
};

__device__ __forceinline__ float p1_weight(int dim, int v, const float *xi3) {
  // QuadP1.h:98-108 / HexP1.cpp:242-298 vertex signs
  const int qs[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  const int hs[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1},
                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  if (dim == 2) return 0.25f * (1.f + qs[v][0] * xi3[0]) * (1.f + qs[v][1] * xi3[1]);
  return 0.125f * (1.f + hs[v][0] * xi3[0]) * (1.f + hs[v][1] * xi3[1]) * (1.f + hs[v][2] * xi3[2]);
}

// ------------------------------------------------------------------ FP64 geometry
// Affine per-element factors from the Jacobian at the element centre.
// Acoustic: Gsym = detJ Jinv^T Jinv (upper triangle, row-major: xx xy [xz] yy [yz] [zz]).
// Elastic : Jinv (dim^2, [a][b] = d r_b / d x_a ... stored as Jinv[a*dim+b]) and detJ.
// quad_ref: the rows of QuadP1.h:20-27, where dN/dr is evaluated at r and
// dN/ds at s (each derivative of the bilinear map taken at the other
// coordinate; equal to the consistent Jacobian on parallelograms).
__device__ void dev_jacobian(int dim, const double *vtx, const double *ref, double *J,
                             bool quad_ref = false) {
  const int qs[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  const int hs[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1},
                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int i = 0; i < dim * dim; ++i) J[i] = 0.0;
  const int nv = dim == 2 ? 4 : 8;
  for (int v = 0; v < nv; ++v) {
    double d[3];
    if (dim == 2) {
      d[0] = 0.25 * qs[v][0] * (1 + ref[quad_ref ? 0 : 1] * qs[v][1]);
      d[1] = 0.25 * qs[v][1] * (1 + ref[quad_ref ? 1 : 0] * qs[v][0]);
    } else {
      d[0] = 0.125 * hs[v][0] * (1 + ref[1] * hs[v][1]) * (1 + ref[2] * hs[v][2]);
      d[1] = 0.125 * hs[v][1] * (1 + ref[0] * hs[v][0]) * (1 + ref[2] * hs[v][2]);
      d[2] = 0.125 * hs[v][2] * (1 + ref[0] * hs[v][0]) * (1 + ref[1] * hs[v][1]);
    }
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) J[a * dim + b] += d[a] * vtx[v * dim + b];
  }
}

__device__ double dev_invert(int dim, const double *J, double *inv) {
  if (dim == 2) {
    double d = J[0] * J[3] - J[1] * J[2];
    inv[0] = J[3] / d; inv[1] = -J[1] / d; inv[2] = -J[2] / d; inv[3] = J[0] / d;
    return d;
  }
  double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8],
         c02 = J[3] * J[7] - J[4] * J[6];
  double d = J[0] * c00 + J[1] * c01 + J[2] * c02;
  inv[0] = c00 / d; inv[1] = (J[2] * J[7] - J[1] * J[8]) / d; inv[2] = (J[1] * J[5] - J[2] * J[4]) / d;
  inv[3] = c01 / d; inv[4] = (J[0] * J[8] - J[2] * J[6]) / d; inv[5] = (J[2] * J[3] - J[0] * J[5]) / d;
  inv[6] = c02 / d; inv[7] = (J[1] * J[6] - J[0] * J[7]) / d; inv[8] = (J[0] * J[4] - J[1] * J[3]) / d;
  return d;
}

// Interpolated vertex parameter q of one element at reference point ref (FP64).
__device__ double dev_p1(int dim, const double *vpar, int npar, int q, const double *ref) {
  const int qs[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  const int hs[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {1, -1, -1},
                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  double s = 0;
  if (dim == 2) {
    for (int v = 0; v < 4; ++v)
      s += 0.25 * (1 + ref[0] * qs[v][0]) * (1 + ref[1] * qs[v][1]) * vpar[v * npar + q];
  } else {
    for (int v = 0; v < 8; ++v)
      s += 0.125 * (1 + ref[0] * hs[v][0]) * (1 + ref[1] * hs[v][1]) * (1 + ref[2] * hs[v][2]) *
           vpar[v * npar + q];
  }
  return s;
}

__global__ void k_affine_geometry(int dim, int physics, int64_t E, const double *__restrict__ vtx,
                                  float *__restrict__ geo) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int nv = dim == 2 ? 4 : 8;
  double J[9], inv[9], ref[3] = {0, 0, 0};
  dev_jacobian(dim, vtx + e * nv * dim, ref, J);
  double det = dev_invert(dim, J, inv);
  if (physics == SEM_ACOUSTIC) {
    const int ng = dim * (dim + 1) / 2;
    int k = 0;
    for (int a = 0; a < dim; ++a)
      for (int b = a; b < dim; ++b) {
        double s = 0;
        for (int c = 0; c < dim; ++c) s += inv[c * dim + a] * inv[c * dim + b];
        geo[e * ng + k++] = (float)(det * s);
      }
  } else {
    const int ng = dim * dim + 1;
    for (int i = 0; i < dim * dim; ++i) geo[e * ng + i] = (float)inv[i];
    geo[e * ng + dim * dim] = (float)det;
  }
}

// Pointwise factors (general geometry) and FP64 element mass, one thread
// per (element, node).  geo_p layout: [E][ngp][nd].  me: [E][nd].
template <int DIM, int N>
__global__ void k_point_geometry(int physics, int gmode, int64_t E, const double *__restrict__ vtx,
                                 const double *__restrict__ vpar, int npar,
                                 float *__restrict__ geo_p, double *__restrict__ me) {
  using S = Sh<DIM, N>;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * S::ND) return;
  int64_t e = t / S::ND;
  int p = (int)(t % S::ND);
  const DevTables &T = c_tab[N];
  int P[3] = {p % S::n, (p / S::n) % S::n, DIM == 3 ? p / (S::n * S::n) : 0};
  double ref[3] = {T.xid[P[0]], T.xid[P[1]], DIM == 3 ? T.xid[P[2]] : 0.0};
  double wq = T.wd[P[0]] * T.wd[P[1]] * (DIM == 3 ? T.wd[P[2]] : 1.0);
  double J[9], inv[9];
  const double *ev = vtx + e * S::NV * DIM;
  const double *ep = vpar + e * S::NV * npar;
  const bool qref = DIM == 2 && gmode == kGeoQuadRef;
  dev_jacobian(DIM, ev, ref, J, qref);
  double det = dev_invert(DIM, J, inv);
  // mass: Scalar.cpp:34-39 (no rho), Elastic*.cpp: rho at the node
  double rho = physics == SEM_ACOUSTIC ? 1.0 : dev_p1(DIM, ep, npar, 0, ref);
  if (me) me[e * S::ND + p] = wq * det * rho;
  if (!geo_p) return;
  const int ngp = point_geo_count(DIM, physics, qref ? kGeoQuadRef : kGeoPoint);
  float *g = geo_p + e * ngp * S::ND + p;
  if (physics == SEM_ACOUSTIC && qref) {
    // reference quad operator: J^-1 at the node and w det J vp^2
    const double vp = dev_p1(DIM, ep, npar, 0, ref);
    for (int i = 0; i < 4; ++i) g[i * S::ND] = (float)inv[i];
    g[4 * S::ND] = (float)(wq * det * vp * vp);
  } else if (physics == SEM_ACOUSTIC) {
    double vp = dev_p1(DIM, ep, npar, 0, ref);
    int k = 0;
    for (int a = 0; a < DIM; ++a)
      for (int b = a; b < DIM; ++b) {
        double s = 0;
        for (int c = 0; c < DIM; ++c) s += inv[c * DIM + a] * inv[c * DIM + b];
        g[(k++) * S::ND] = (float)(wq * vp * vp * det * s);
      }
  } else {
    for (int i = 0; i < DIM * DIM; ++i) g[i * S::ND] = (float)inv[i];
    g[DIM * DIM * S::ND] = (float)(wq * det);
    if (DIM == 2) {
      for (int q = 0; q < 4; ++q) g[(5 + q) * S::ND] = (float)dev_p1(DIM, ep, npar, 1 + q, ref);
    } else {
      double r = dev_p1(DIM, ep, npar, 0, ref), vpv = dev_p1(DIM, ep, npar, 1, ref),
             vph = dev_p1(DIM, ep, npar, 2, ref), vsv = dev_p1(DIM, ep, npar, 3, ref),
             vsh = dev_p1(DIM, ep, npar, 4, ref), eta = dev_p1(DIM, ep, npar, 5, ref);
      double c11 = r * vph * vph, c33 = r * vpv * vpv, c44 = r * vsv * vsv, c66 = r * vsh * vsh;
      double cc[9] = {c11, c11 - 2 * c66, eta * (c11 - 2 * c44), c11, eta * (c11 - 2 * c44), c33,
                      c44, c44, c66};  // c11 c12 c13 c22 c23 c33 c44 c55 c66
      for (int q = 0; q < 9; ++q) g[(10 + q) * S::ND] = (float)cc[q];
    }
  }
}

// Colour-ordered FP64 mass assembly: one launch per colour.
__global__ void k_mass_scatter(int64_t first, int64_t count, int nd, const uint32_t *__restrict__ conn,
                               const double *__restrict__ me, double *__restrict__ M) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * nd) return;
  int64_t e = first + t / nd;
  int p = (int)(t % nd);
  uint32_t g = conn[e * nd + p] & kNodeMask;
  M[g] += me[e * nd + p];
}

__global__ void k_reciprocal(int64_t G, const double *__restrict__ M, float *__restrict__ mi) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) mi[g] = (float)(1.0 / M[g]);
}

__global__ void k_fill_f64(int64_t n, double *p, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(int64_t n, float *p, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int64_t n, int32_t *p, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Race check: count[g] += 1 for each element of one colour; nonzero >1 flags.
__global__ void k_colour_touch(int64_t first, int64_t count, int nd, const uint32_t *__restrict__ conn,
                               int32_t *__restrict__ hits) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * nd) return;
  uint32_t g = conn[(first + t / nd) * nd + t % nd] & kNodeMask;
  atomicAdd(&hits[g], 1);
}
__global__ void k_colour_count(int64_t G, int32_t *__restrict__ hits, unsigned long long *viol) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  if (hits[g] > 1) atomicAdd(viol, (unsigned long long)(hits[g] - 1));
  hits[g] = 0;
}

// ------------------------------------------------------------------ element operator
// Per-thread cached 1-D operator rows/columns.
template <int DIM, int N>
struct NodeCtx {
  using S = Sh<DIM, N>;
  int p, P[3], base[3];
  float Drow[DIM][N + 1], Dcol[DIM][N + 1];
  float wq, xi[3];
  __device__ void init(int p_) {
    const DevTables &T = c_tab[N];
    p = p_;
    P[0] = p % S::n;
    P[1] = (p / S::n) % S::n;
    P[2] = DIM == 3 ? p / (S::n * S::n) : 0;
    const int stride[3] = {1, S::n, S::n * S::n};
    wq = 1.f;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      base[a] = p - P[a] * stride[a];
      wq *= T.w[P[a]];
      xi[a] = T.xi[P[a]];
#pragma unroll
      for (int m = 0; m < S::n; ++m) {
        Drow[a][m] = T.D[P[a] * S::n + m];
        Dcol[a][m] = T.D[m * S::n + P[a]];
      }
    }
    if (DIM == 2) xi[2] = 0.f;
  }
  // reference gradient of the LDS field f (nd floats)
  __device__ __forceinline__ void grad(const float *f, float *g) const {
    const int stride[3] = {1, S::n, S::n * S::n};
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      float s = 0.f;
#pragma unroll
      for (int m = 0; m < S::n; ++m) s = fmaf(Drow[a][m], f[base[a] + m * stride[a]], s);
      g[a] = s;
    }
  }
  // test-and-integrate with J^-1 applied at this (output) node, as
  // TensorQuad.cpp:350-397 does: F[a*ND + q] holds the physical-axis-a flux
  // times w det J (no J^-T), and the result is
  // sum_a sum_b Jinv[a][b] sum_m D[m][P_b] F_a(line b, m).
  __device__ __forceinline__ float div_ref(const float *F, const float *Ji) const {
    const int stride[3] = {1, S::n, S::n * S::n};
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) {
        float d = 0.f;
#pragma unroll
        for (int m = 0; m < S::n; ++m) d = fmaf(Dcol[b][m], F[a * S::ND + base[b] + m * stride[b]], d);
        s = fmaf(Ji[a * DIM + b], d, s);
      }
    return s;
  }
  // test-and-integrate of LDS fluxes F[a*ND + q]
  __device__ __forceinline__ float div(const float *F) const {
    const int stride[3] = {1, S::n, S::n * S::n};
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int m = 0; m < S::n; ++m) s = fmaf(Dcol[a][m], F[a * S::ND + base[a] + m * stride[a]], s);
    return s;
  }
};

// Material/geometry of one node, gathered once per element visit.
template <int DIM, int PHYS>
struct NodeMat {
  // acoustic: G[6] (already times w vp^2);
  // elastic : Jinv[DIM*DIM], wdet, c[9] (2D uses c[0..3] = C11,C13,C33,C55),
  //           the coefficients times wdet
  float G[6];
  float Jinv[9];
  float wdet;
  float c[9];
};

template <int DIM, int N, int PHYS, int GM>
__device__ __forceinline__ void load_mat(const NodeCtx<DIM, N> &X, int64_t e, gptr<const float> geo_e,
                                         gptr<const float> par_e, gptr<const float> geo_p,
                                         NodeMat<DIM, PHYS> &M) {
  using S = Sh<DIM, N>;
  constexpr bool AFFINE = GM == kGeoAffine;
  if (AFFINE) {
    float w8[8];
#pragma unroll
    for (int v = 0; v < S::NV; ++v) w8[v] = p1_weight(DIM, v, X.xi);
    if (PHYS == SEM_ACOUSTIC) {
      constexpr int ng = DIM * (DIM + 1) / 2;
      const gptr<const float> gp = geo_e + e * ng;
      float vp = 0.f;
#pragma unroll
      for (int v = 0; v < S::NV; ++v) vp = fmaf(w8[v], par_e[e * S::NV + v], vp);
      float s = X.wq * vp * vp;  // Scalar.cpp:45: vp^2 from P1-interpolated VP
#pragma unroll
      for (int k = 0; k < ng; ++k) M.G[k] = s * gp[k];
    } else {
      constexpr int ng = DIM * DIM + 1;
      const gptr<const float> gp = geo_e + e * ng;
#pragma unroll
      for (int k = 0; k < DIM * DIM; ++k) M.Jinv[k] = gp[k];
      M.wdet = X.wq * gp[DIM * DIM];
      if (DIM == 2) {
        const gptr<const float> pp = par_e + e * 4 * 4;  // 4 vertices x {C11,C13,C33,C55}
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          float s = 0.f;
#pragma unroll
          for (int v = 0; v < 4; ++v) s = fmaf(w8[v], pp[v * 4 + q], s);
          M.c[q] = s;
        }
      } else {
        const gptr<const float> pp = par_e + e * 8 * 6;  // 8 vertices x {RHO,VPV,VPH,VSV,VSH,ETA}
        float q6[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          float s = 0.f;
#pragma unroll
          for (int v = 0; v < 8; ++v) s = fmaf(w8[v], pp[v * 6 + q], s);
          q6[q] = s;
        }
        // Elastic3D.cpp:42-52
        float c11 = q6[0] * q6[2] * q6[2], c33 = q6[0] * q6[1] * q6[1];
        float c44 = q6[0] * q6[3] * q6[3], c66 = q6[0] * q6[4] * q6[4];
        float c13 = q6[5] * (c11 - 2.f * c44);
        M.c[0] = c11; M.c[1] = c11 - 2.f * c66; M.c[2] = c13; M.c[3] = c11; M.c[4] = c13;
        M.c[5] = c33; M.c[6] = c44; M.c[7] = c44; M.c[8] = c66;
      }
    }
  } else {
    constexpr int ngp = PHYS == SEM_ACOUSTIC ? (GM == kGeoQuadRef ? 5 : DIM * (DIM + 1) / 2)
                                             : (DIM == 2 ? 9 : 19);
    const gptr<const float> gp = geo_p + e * ngp * S::ND + X.p;
    if (PHYS == SEM_ACOUSTIC && GM == kGeoQuadRef) {
#pragma unroll
      for (int k = 0; k < 4; ++k) M.Jinv[k] = gp[k * S::ND];
      M.wdet = gp[4 * S::ND];  // w det J vp^2
    } else if (PHYS == SEM_ACOUSTIC) {
#pragma unroll
      for (int k = 0; k < ngp; ++k) M.G[k] = gp[k * S::ND];
    } else {
#pragma unroll
      for (int k = 0; k < DIM * DIM; ++k) M.Jinv[k] = gp[k * S::ND];
      M.wdet = gp[DIM * DIM * S::ND];
#pragma unroll
      for (int k = 0; k < (DIM == 2 ? 4 : 9); ++k) M.c[k] = gp[(DIM * DIM + 1 + k) * S::ND];
    }
  }
  if (PHYS != SEM_ACOUSTIC) {
    // the stress is linear in the coefficients: scaling them by w det J once
    // per element visit leaves the flux without a scaling of its own
#pragma unroll
    for (int k = 0; k < (DIM == 2 ? 4 : 9); ++k) M.c[k] *= M.wdet;
  }
}

// Stage 1: fluxes of this node into LDS.  su: C x ND field, sF: C x DIM x ND.
// QREF (reference quad operator): the physical-axis fluxes times w det J,
// without J^-T (div_ref applies J^-1 at the output node).
template <int DIM, int N, int PHYS, bool QREF = false>
__device__ __forceinline__ void node_flux(const NodeCtx<DIM, N> &X, const NodeMat<DIM, PHYS> &M,
                                          const float *su, float *sF) {
  using S = Sh<DIM, N>;
  if (PHYS == SEM_ACOUSTIC && QREF) {
    float g[3];
    X.grad(su, g);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      float d = 0.f;
#pragma unroll
      for (int b = 0; b < DIM; ++b) d = fmaf(M.Jinv[a * DIM + b], g[b], d);
      sF[a * S::ND + X.p] = M.wdet * d;  // Scalar.cpp:45-71: w det J vp^2 grad u
    }
  } else if (PHYS == SEM_ACOUSTIC) {
    float g[3];
    X.grad(su, g);
    if (DIM == 2) {
      sF[0 * S::ND + X.p] = M.G[0] * g[0] + M.G[1] * g[1];
      sF[1 * S::ND + X.p] = M.G[1] * g[0] + M.G[2] * g[1];
    } else {
      sF[0 * S::ND + X.p] = M.G[0] * g[0] + M.G[1] * g[1] + M.G[2] * g[2];
      sF[1 * S::ND + X.p] = M.G[1] * g[0] + M.G[3] * g[1] + M.G[4] * g[2];
      sF[2 * S::ND + X.p] = M.G[2] * g[0] + M.G[4] * g[1] + M.G[5] * g[2];
    }
  } else {
    // physical gradients du_c/dx_a = sum_b Jinv[a][b] g_b
    float du[3][3];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      float g[3];
      X.grad(su + c * S::ND, g);
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        float s = 0.f;
#pragma unroll
        for (int b = 0; b < DIM; ++b) s = fmaf(M.Jinv[a * DIM + b], g[b], s);
        du[c][a] = s;
      }
    }
    float sig[3][3];
    if (DIM == 2) {
      // Elastic2D.cpp:53-82 (mc13 = mc23 = 0)
      float exx = du[0][0], eyy = du[1][1], sh = du[0][1] + du[1][0];
      sig[0][0] = M.c[0] * exx + M.c[1] * eyy;
      sig[1][1] = M.c[1] * exx + M.c[2] * eyy;
      sig[0][1] = sig[1][0] = M.c[3] * sh;
    } else {
      // Elastic3D.cpp:76-83, :106-118
      float e0 = du[0][0], e1 = du[1][1], e2 = du[2][2];
      float e3 = du[1][2] + du[2][1], e4 = du[0][2] + du[2][0], e5 = du[0][1] + du[1][0];
      sig[0][0] = M.c[0] * e0 + M.c[1] * e1 + M.c[2] * e2;
      sig[1][1] = M.c[1] * e0 + M.c[3] * e1 + M.c[4] * e2;
      sig[2][2] = M.c[2] * e0 + M.c[4] * e1 + M.c[5] * e2;
      sig[1][2] = sig[2][1] = M.c[6] * e3;
      sig[0][2] = sig[2][0] = M.c[7] * e4;
      sig[0][1] = sig[1][0] = M.c[8] * e5;
    }
    if (QREF) {
      // rows (sigma_xx, sigma_xy) and (sigma_xy, sigma_yy), Elastic2D.cpp:84-107
#pragma unroll
      for (int c = 0; c < DIM; ++c)
#pragma unroll
        for (int a = 0; a < DIM; ++a) sF[(c * DIM + a) * S::ND + X.p] = sig[c][a];
      return;
    }
    // F_{c,a} = w detJ sum_b Jinv[b][a] sigma_{c,b}
#pragma unroll
    for (int c = 0; c < DIM; ++c)
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        float s = 0.f;
#pragma unroll
        for (int b = 0; b < DIM; ++b) s = fmaf(M.Jinv[b * DIM + a], sig[c][b], s);
        sF[(c * DIM + a) * S::ND + X.p] = s;
      }
  }
}

// ------------------------------------------------------------------ fused colour kernel
// One launch per colour.  Each block takes one work item (EPB elements of
// one group); for every shot of the group it gathers u, applies the element
// operator, and for every node either stores (first colour), accumulates,
// or -- on the node's last colour -- finishes the step there:
//   a = mi (acc + f - K u); Kahan v += dt/2 (a + a_); Kahan u += dt v + dt^2/2 a; a_ = a
// (Order2Newmark.cpp:68-115).  Sources add their nodal load times R(t_n).


__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
  float y = x - c;
  float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int DIM, int N, int PHYS, int GM>
__global__ void __launch_bounds__(512) k_fused(FusedArgs A) {
  using S = Sh<DIM, N>;
  constexpr int C = PHYS == SEM_ACOUSTIC ? 1 : DIM;
  constexpr bool QREF = GM == kGeoQuadRef;
  extern __shared__ float lds[];
  // XCD-aware item order: consecutive hardware block ids round-robin over
  // 8 XCDs; give each XCD a contiguous range of items.
  int bid = blockIdx.x, nb = gridDim.x;
  int item = bid;
  if (nb >= 16 && (nb & 7) == 0) item = (bid & 7) * (nb >> 3) + (bid >> 3);
  if (item >= A.nitems) return;
  const WorkItem W = glb(A.items)[item];
  const auto &G = glb(A.groups)[W.group];
  const int64_t step = *glb(G.step) + A.koff;
  if (step >= G.nsteps) return;  // uniform across the block
  const gptr<float> gu = glb(G.u), gw = glb(G.w), gap = glb(G.ap), gacc = glb(G.acc);
  const gptr<uint32_t> gkc = glb(G.kc);
  // descriptor fields the shot loop needs, read once: G sits in global memory,
  // so a field read after the loop's global stores would be reloaded (the
  // compiler cannot rule out aliasing), and waiting for that load would wait
  // for every store before it
  const int64_t GG = G.G;
  const float gdt = G.dt;
  const gptr<const float> gext = glb(G.ext), gstf = glb(G.stf), gload = glb(G.load);
  const bool has_ext = G.ext != nullptr;
  const gptr<const SrcDev> gsrc = glb(G.src);
  const int nsrc = G.nsrc;
  const bool q3 = (C >= 2) && (G.quirks & SEM_QUIRK_NEWMARK3D);  // {x, y, x}: x twice, z never
  const bool wap = A.koff == A.ap_koff || step + 1 >= G.nsteps || G.ap_always;

  const int t = threadIdx.x;
  // the group's first sources staged in LDS once per block (read after the
  // shot loop's barriers), so that a node's source scan is no chain of
  // dependent global loads
  constexpr int kMaxStaged = 64;
  __shared__ SrcDev ssrc[kMaxStaged];
  for (int k = t; k < nsrc && k < kMaxStaged; k += blockDim.x) ssrc[k] = gsrc[k];
  const int el = t / S::ND;
  const int p = t - el * S::ND;
  const bool active = el < W.count;
  float *su = lds + el * (C * S::ND);
  float *sF = lds + A.epb * C * S::ND + el * (C * DIM * S::ND);

  NodeCtx<DIM, N> X;
  NodeMat<DIM, PHYS> M;
  int64_t e = W.first + el;
  uint32_t word = 0, g = 0;
  float mi = 0.f;
  if (active) {
    X.init(p);
    word = glb(G.conn)[e * S::ND + p];
    g = word & kNodeMask;
    load_mat<DIM, N, PHYS, GM>(X, e, glb(G.geo_e), glb(G.par_e), glb(G.geo_p), M);
    if (word & kFlagLast) mi = glb(G.mi)[g];
  }
  const int Sg = G.S;
  const bool first = (word & kFlagFirst) != 0;
  const bool fin = (word & kFlagLast) && !A.ku_only;
  // shot 0's u gathered before the shot loop (every later shot's u is loaded
  // while the previous shot integrates and finishes)
  if (active) {
    float u0[C];
#pragma unroll
    for (int c = 0; c < C; ++c) u0[c] = gu[(int64_t)c * GG + g];
#pragma unroll
    for (int c = 0; c < C; ++c) su[c * S::ND + p] = u0[c];
  }
  for (int s = 0; s < Sg; ++s) {
    // this node's finish inputs, loaded ahead of the element work
    float pacc[C], pw[C];
    uint32_t pkc[C];
    if (active) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int64_t idx = ((int64_t)s * C + c) * GG + g;
        pacc[c] = first ? (has_ext ? gext[idx] : 0.f) : gacc[idx];
        pw[c] = fin ? gw[idx] : 0.f;
        pkc[c] = fin ? gkc[idx] : 0u;
      }
    }
    __syncthreads();
    if (active) node_flux<DIM, N, PHYS, QREF>(X, M, su, sF);
    __syncthreads();
    if (active) {
      float nu[C];
      const bool more = s + 1 < Sg;
#pragma unroll
      for (int c = 0; c < C; ++c) nu[c] = more ? gu[((int64_t)(s + 1) * C + c) * GG + g] : 0.f;
      float f[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        float ku = QREF ? X.div_ref(sF + c * DIM * S::ND, M.Jinv) : X.div(sF + c * DIM * S::ND);
        f[c] = (word & kFlagDirichlet) ? 0.f : -ku;
      }
      auto add_source = [&](const SrcDev &sd) {
        if (sd.elem == e && sd.shot == s) {
          float R = gstf[sd.stf_off + step];
#pragma unroll
          for (int c = 0; c < C; ++c) f[c] = fmaf(gload[sd.load_off + c * S::ND + p], R, f[c]);
        }
      };
      const int nsl = nsrc < kMaxStaged ? nsrc : kMaxStaged;
      for (int k = 0; k < nsl; ++k) add_source(ssrc[k]);           // LDS
      for (int k = kMaxStaged; k < nsrc; ++k) add_source(gsrc[k]);  // beyond the staged ones
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int64_t idx = ((int64_t)s * C + c) * GG + g;
        const float t = (first && !has_ext) ? f[c] : pacc[c] + f[c];
        if (!fin) gacc[idx] = t;
        else f[c] = t;
      }
      if (fin) {
        // Order2Newmark.cpp:68-115; with the field-list quirk every elastic
        // mesh advances x twice, y once and (3-D) z never
#pragma unroll
        for (int c = 0; c < C; ++c) {
          if (q3 && c == 2) continue;  // c == 2 exists only in 3-D
          const int64_t idx = ((int64_t)s * C + c) * GG + g;
          float a = mi * f[c];
          float w = pw[c], kw = comp_v(pkc[c]);
          float u = su[c * S::ND + p], ku = comp_u(pkc[c]);
          const int reps = (q3 && c == 0) ? 2 : 1;
          for (int r = 0; r < reps; ++r) {  // w = v + dt/2 a_: w += dt a, u += dt w
            kahan_add(w, kw, gdt * a);
            kahan_add(u, ku, gdt * w);
          }
          gu[idx] = u; gw[idx] = w; gkc[idx] = pack_comp(ku, kw);
          if (wap) gap[idx] = a;  // read out only
        }
      }
      // own slot only: every thread's flux stage has read su (barrier above)
      // and this thread's finish has read its u
      if (more) {
#pragma unroll
        for (int c = 0; c < C; ++c) su[c * S::ND + p] = nu[c];
      }
    }
    __syncthreads();
  }
}

__global__ void k_permute_f32(int64_t n, const int32_t *__restrict__ to, const float *__restrict__ src,
                              float *__restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[to[i]] = src[i];
}

// ------------------------------------------------------------------ plain path
// One shot per launch; S = 1 layout.  gather -> stiffness -> colour scatter
// -> update, each a separate kernel with element-local buffers in HBM.
__global__ void k_plain_gather(int64_t E, int nd, int C, int64_t G, const uint32_t *__restrict__ conn,
                               const float *__restrict__ u, float *__restrict__ uloc) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * nd) return;
  uint32_t g = conn[t] & kNodeMask;
  int64_t e = t / nd;
  int p = (int)(t % nd);
  for (int c = 0; c < C; ++c) uloc[(e * C + c) * nd + p] = u[(int64_t)c * G + g];
}



template <int DIM, int N, int PHYS, int GM>
__global__ void __launch_bounds__(512) k_plain_stiffness(PlainArgs A) {
  using S = Sh<DIM, N>;
  constexpr int C = PHYS == SEM_ACOUSTIC ? 1 : DIM;
  constexpr bool QREF = GM == kGeoQuadRef;
  extern __shared__ float lds[];
  const int t = threadIdx.x;
  const int el = t / S::ND;
  const int p = t - el * S::ND;
  const int64_t e = (int64_t)blockIdx.x * A.epb + el;
  const bool active = el < A.epb && e < A.E;
  float *su = lds + el * (C * S::ND);
  float *sF = lds + A.epb * C * S::ND + el * (C * DIM * S::ND);
  NodeCtx<DIM, N> X;
  NodeMat<DIM, PHYS> M;
  if (active) {
    X.init(p);
    load_mat<DIM, N, PHYS, GM>(X, e, glb((const float *)nullptr), glb((const float *)nullptr),
                               glb(A.geo_p), M);
#pragma unroll
    for (int c = 0; c < C; ++c) su[c * S::ND + p] = glb(A.uloc)[(e * C + c) * S::ND + p];
  }
  __syncthreads();
  if (active) node_flux<DIM, N, PHYS, QREF>(X, M, su, sF);
  __syncthreads();
  if (!active) return;
  const uint32_t word = glb(A.conn)[e * S::ND + p];
  const int64_t step = A.step ? *glb(A.step) + A.koff : 0;
  float f[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    float ku = QREF ? X.div_ref(sF + c * DIM * S::ND, M.Jinv) : X.div(sF + c * DIM * S::ND);
    f[c] = (word & kFlagDirichlet) ? 0.f : -ku;
  }
  for (int k = 0; k < A.nsrc; ++k) {
    const SrcDev sd = glb(A.src)[k];
    if (sd.elem == e) {
      float R = glb(A.stf)[sd.stf_off + step];
#pragma unroll
      for (int c = 0; c < C; ++c) f[c] = fmaf(glb(A.load)[sd.load_off + c * S::ND + p], R, f[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) glb(A.aloc)[(e * C + c) * S::ND + p] = f[c];
}

__global__ void k_plain_scatter(int64_t first, int64_t count, int nd, int C, int64_t G,
                                const uint32_t *__restrict__ conn, const float *__restrict__ aloc,
                                float *__restrict__ acc) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * nd) return;
  int64_t e = first + t / nd;
  int p = (int)(t % nd);
  uint32_t g = conn[e * nd + p] & kNodeMask;
  for (int c = 0; c < C; ++c) acc[(int64_t)c * G + g] += aloc[(e * C + c) * nd + p];
}

__global__ void k_plain_update(int64_t G, int C, int quirks, float dt,
                               const float *__restrict__ mi, const float *__restrict__ acc,
                               const float *__restrict__ ext, float *__restrict__ u,
                               float *__restrict__ w, float *__restrict__ ap,
                               uint32_t *__restrict__ kc) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const bool q3 = C >= 2 && (quirks & SEM_QUIRK_NEWMARK3D);  // {x, y, x}: x twice, z never
  float m = mi[g];
  for (int c = 0; c < C; ++c) {
    if (q3 && c == 2) continue;
    int64_t i = (int64_t)c * G + g;
    float a = m * (ext ? acc[i] + ext[i] : acc[i]);
    const uint32_t k2 = kc[i];
    float ww = w[i], kww = comp_v(k2), uu = u[i], kuu = comp_u(k2);
    int reps = (q3 && c == 0) ? 2 : 1;
    for (int r = 0; r < reps; ++r) {  // w = v + dt/2 a_: w += dt a, u += dt w
      kahan_add(ww, kww, dt * a);
      kahan_add(uu, kuu, dt * ww);
    }
    u[i] = uu; w[i] = ww; ap[i] = a; kc[i] = pack_comp(kuu, kww);
  }
}

// ------------------------------------------------------------------ receivers / steps


// Samples u after the step just taken (global step = *step + koff): one wave
// per receiver component, its lanes over the owning element's nodes, summed
// by a fixed butterfly.
__global__ void __launch_bounds__(256) k_receivers(RecArgs A, int C) {
  const int t = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= A.nrec * C) return;  // uniform per wave
  const int r = t / C, c = t % C;
  const auto &G = glb(A.groups)[glb(A.rec_group)[r]];
  const int64_t step = *glb(G.step) + A.koff;
  if (step >= G.nsteps) return;
  const RecDev R = glb(A.recs)[r];
  const gptr<const uint32_t> nodes = glb(A.nodes) + R.nodes_off;
  const gptr<const float> w = glb(A.weights) + R.w_off;
  const gptr<const float> u = glb(G.u) + ((int64_t)R.shot * C + c) * G.G;
  float s = 0.f;
  for (int p = lane; p < A.nd; p += 64) s = fmaf(w[p], u[nodes[p]], s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) glb(R.out)[step * R.stride + c] = s;
}

__global__ void k_advance(int ngroups, const GroupDev *groups, int64_t k) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ngroups) return;
  const gptr<int64_t> step = glb(glb(groups)[i].step);
  const int64_t s = *step + k;
  *step = s < glb(groups)[i].nsteps ? s : glb(groups)[i].nsteps;
}

}  // namespace sem

// ====================================================================== launchers
namespace sem {

namespace {
inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
int last_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return -2;
  }
  return 0;
}
}  // namespace

int init_device_tables(int device) {
  static bool done[64] = {false};
  if (device < 0 || device >= 64) { set_error("device index"); return -1; }
  if (done[device]) return 0;
  DevTables host[kMaxOrder2D + 1];
  memset(host, 0, sizeof host);
  for (int N = 1; N <= kMaxOrder2D; ++N) {
    Gll g = make_gll(N);
    for (int i = 0; i < g.n * g.n; ++i) host[N].D[i] = (float)g.D[i];
    for (int i = 0; i < g.n; ++i) {
      host[N].w[i] = (float)g.w[i];
      host[N].xi[i] = (float)g.xi[i];
      host[N].wd[i] = g.w[i];
      host[N].xid[i] = g.xi[i];
    }
  }
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), host, sizeof host, 0, hipMemcpyHostToDevice);
  if (e != hipSuccess) { set_error(std::string("GLL tables: ") + hipGetErrorString(e)); return -2; }
  done[device] = true;
  return 0;
}

int launch_affine_geometry(int dim, int physics, int64_t E, const double *vtx, float *geo,
                           hipStream_t st) {
  k_affine_geometry<<<nblocks(E, 128), 128, 0, st>>>(dim, physics, E, vtx, geo);
  return last_launch("affine geometry");
}

#define SEM_ORDERS_3D(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
#define SEM_ORDERS_2D(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)

int launch_point_geometry(int dim, int order, int physics, int gmode, int64_t E, const double *vtx,
                          const double *vpar, int npar, float *geo_p, double *me, hipStream_t st) {
  const int n = order + 1;
  const int64_t total = E * (dim == 2 ? n * n : n * n * n);
#define PG3(N) case N: k_point_geometry<3, N><<<nblocks(total, 128), 128, 0, st>>>(physics, gmode, E, vtx, vpar, npar, geo_p, me); break;
#define PG2(N) case N: k_point_geometry<2, N><<<nblocks(total, 128), 128, 0, st>>>(physics, gmode, E, vtx, vpar, npar, geo_p, me); break;
  if (dim == 3) {
    switch (order) { SEM_ORDERS_3D(PG3) default: set_error("order not compiled"); return -1; }
  } else {
    switch (order) { SEM_ORDERS_2D(PG2) default: set_error("order not compiled"); return -1; }
  }
  return last_launch("point geometry");
}

int launch_mass_scatter(int64_t first, int64_t count, int nd, const uint32_t *conn,
                        const double *me, double *M, hipStream_t st) {
  if (count <= 0) return 0;
  k_mass_scatter<<<nblocks(count * nd, 256), 256, 0, st>>>(first, count, nd, conn, me, M);
  return last_launch("mass scatter");
}
int launch_reciprocal(int64_t G, const double *M, float *mi, hipStream_t st) {
  k_reciprocal<<<nblocks(G, 256), 256, 0, st>>>(G, M, mi);
  return last_launch("reciprocal");
}
int launch_fill_f64(int64_t n, double *p, double v, hipStream_t st) {
  if (n <= 0) return 0;
  k_fill_f64<<<nblocks(n, 256), 256, 0, st>>>(n, p, v);
  return last_launch("fill");
}
int launch_fill_f32(int64_t n, float *p, float v, hipStream_t st) {
  if (n <= 0) return 0;
  k_fill_f32<<<nblocks(n, 256), 256, 0, st>>>(n, p, v);
  return last_launch("fill");
}
int launch_fill_i32(int64_t n, int32_t *p, int32_t v, hipStream_t st) {
  if (n <= 0) return 0;
  k_fill_i32<<<nblocks(n, 256), 256, 0, st>>>(n, p, v);
  return last_launch("fill");
}
int launch_colour_touch(int64_t first, int64_t count, int nd, const uint32_t *conn, int32_t *hits,
                        hipStream_t st) {
  if (count <= 0) return 0;
  k_colour_touch<<<nblocks(count * nd, 256), 256, 0, st>>>(first, count, nd, conn, hits);
  return last_launch("colour touch");
}
int launch_colour_count(int64_t G, int32_t *hits, unsigned long long *viol, hipStream_t st) {
  k_colour_count<<<nblocks(G, 256), 256, 0, st>>>(G, hits, viol);
  return last_launch("colour count");
}

int launch_fused(int dim, int order, int physics, int gmode, int nitems, int epb, int nd, int c,
                 const FusedArgs &a, hipStream_t st) {
  if (nitems <= 0) return 0;
  const int bt = block_threads(nd, epb);
  if (bt > 512) { set_error("elements per block too large"); return -1; }
  const size_t shm = fused_lds_bytes(dim, nd, c, epb);
  dim3 grid(nitems), block(bt);
#define FU(D, N, P, GMV) hipLaunchKernelGGL((k_fused<D, N, P, GMV>), grid, block, shm, st, a)
#define FUP(D, N, P)                                                          \
  if (gmode == kGeoAffine) FU(D, N, P, kGeoAffine);                           \
  else if (gmode == kGeoQuadRef) FU(D, N, P, (D == 2 ? kGeoQuadRef : kGeoPoint)); \
  else FU(D, N, P, kGeoPoint);
#define FU4(D, N)                                      \
  case N:                                              \
    if (physics == SEM_ACOUSTIC) {                     \
      FUP(D, N, 0)                                     \
    } else {                                           \
      FUP(D, N, 1)                                     \
    }                                                  \
    break;
#define FU3(N) FU4(3, N)
#define FU2(N) FU4(2, N)
  if (dim == 3) {
    switch (order) { SEM_ORDERS_3D(FU3) default: set_error("order not compiled"); return -1; }
  } else {
    switch (order) { SEM_ORDERS_2D(FU2) default: set_error("order not compiled"); return -1; }
  }
  return last_launch("fused step");
}

int launch_permute_f32(int64_t n, const int32_t *to, const float *src, float *dst, hipStream_t st) {
  if (n <= 0) return 0;
  k_permute_f32<<<nblocks(n, 256), 256, 0, st>>>(n, to, src, dst);
  return last_launch("permute");
}

int launch_plain_gather(int64_t E, int nd, int C, int64_t G, const uint32_t *conn, const float *u,
                        float *uloc, hipStream_t st) {
  k_plain_gather<<<nblocks(E * nd, 256), 256, 0, st>>>(E, nd, C, G, conn, u, uloc);
  return last_launch("plain gather");
}

int launch_plain_stiffness(int dim, int order, int physics, int gmode, int64_t E, int nd, int c,
                           const PlainArgs &a, hipStream_t st) {
  const int bt = block_threads(nd, a.epb);
  const size_t shm = fused_lds_bytes(dim, nd, c, a.epb);
  dim3 grid(nblocks(E, a.epb)), block(bt);
  // the plain path always takes pointwise factors
  const bool qref = dim == 2 && gmode == kGeoQuadRef;
#define PSG(D, N, P)                                                                            \
  if (qref) hipLaunchKernelGGL((k_plain_stiffness<D, N, P, (D == 2 ? kGeoQuadRef : kGeoPoint)>), \
                               grid, block, shm, st, a);                                         \
  else hipLaunchKernelGGL((k_plain_stiffness<D, N, P, kGeoPoint>), grid, block, shm, st, a);
#define PS(D, N) case N: \
    if (physics == SEM_ACOUSTIC) { PSG(D, N, 0) } \
    else { PSG(D, N, 1) } \
    break;
#define PS3(N) PS(3, N)
#define PS2(N) PS(2, N)
  if (dim == 3) {
    switch (order) { SEM_ORDERS_3D(PS3) default: set_error("order not compiled"); return -1; }
  } else {
    switch (order) { SEM_ORDERS_2D(PS2) default: set_error("order not compiled"); return -1; }
  }
  return last_launch("plain stiffness");
}

int launch_plain_scatter(int64_t first, int64_t count, int nd, int C, int64_t G,
                         const uint32_t *conn, const float *aloc, float *acc, hipStream_t st) {
  if (count <= 0) return 0;
  k_plain_scatter<<<nblocks(count * nd, 256), 256, 0, st>>>(first, count, nd, C, G, conn, aloc, acc);
  return last_launch("plain scatter");
}

int launch_plain_update(int64_t G, int C, int quirks, float dt, const float *mi, const float *acc,
                        const float *ext, float *u, float *w, float *ap, uint32_t *kc, hipStream_t st) {
  k_plain_update<<<nblocks(G, 256), 256, 0, st>>>(G, C, quirks, dt, mi, acc, ext, u, w, ap, kc);
  return last_launch("plain update");
}

int launch_receivers(const RecArgs &a, int C, hipStream_t st) {
  if (a.nrec <= 0) return 0;
  k_receivers<<<nblocks((int64_t)a.nrec * C, 4), 256, 0, st>>>(a, C);  // 4 waves per block
  return last_launch("receivers");
}

int launch_advance(int ngroups, const GroupDev *groups, int64_t k, hipStream_t st) {
  if (ngroups <= 0) return 0;
  k_advance<<<nblocks(ngroups, 64), 64, 0, st>>>(ngroups, groups, k);
  return last_launch("advance");
}

}  // namespace sem
