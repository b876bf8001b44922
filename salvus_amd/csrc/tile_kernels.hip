// Column-mapped tile kernel of the MI355X SEM path (gfx950).
//
// Structured (logically Cartesian, affine) meshes are cut into tiles of
// TX x TY (x TZ) elements whose node lattice is staged in LDS; one workgroup
// takes one tile, one launch takes the tiles of one lattice-parity colour.
// One thread owns one node column of one element: the column runs along the
// last reference axis (eta in 2-D, zeta in 3-D), so the derivative along it
// is a register-only product with wave-uniform D entries, and only the
// lateral derivatives read the LDS lattice.  Columns run in chunks of at most
// 512 threads, all chunks equally full.  Per shot:
//   gather  u of every lattice slot -> LDS (the first shot's with the tile's
//           words and element data; every later shot's, with the finish
//           inputs of the first slots, during the previous shot's face
//           passes, so each shot waits for one memory round trip);
//   column  reference gradient and J^-T flux per node; lateral fluxes go to
//           LDS, the column flux stays in registers; the test-and-integrate
//           result (-Ku, zero on Dirichlet rows, plus source loads) is
//           stored element-locally in LDS;
//   faces   element copies of shared nodes are summed in fixed passes along
//           x, y (and z), after which the element of highest lattice
//           position containing a node holds the tile's total for it;
//   finish  every lattice slot is stored / accumulated in HBM, or -- on the
//           last tile colour touching it -- finished: a = mi (acc + f);
//           Kahan w += dt a; Kahan u += dt w; a_ = a, where w = v + dt/2 a_
//           is the state the update reads (Order2Newmark.cpp:68-115:
//           v += dt/2 (a + a_), u += dt v + dt^2/2 a, a_ = a, rewritten so
//           that a_ is only written).
// The operator is the one of TensorQuad.cpp:176-210/:350-397 and
// Hexahedra.cpp:579-754 (sum-factorised gradient, J^-T contraction,
// test-and-integrate) with Scalar.cpp:42-71, Elastic2D.cpp:53-82 and
// Elastic3D.cpp:33-54/:70-118 material laws.  All sums are in a fixed order,
// so results are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "sem_internal.h"
#include "tile_layout.h"

namespace sem {

namespace {

constexpr int kTileMaxN = 8;

// GLL points, weights and derivative matrix of order N evaluated at compile
// time, so that every wave-uniform table entry becomes an instruction
// literal.  Points: interior roots of P_N' by Newton from the Chebyshev-
// Gauss-Lobatto guesses; w_j = 2 / (N (N+1) P_N(x_j)^2); D[i][m] = l'_m(x_i)
// = P_N(x_i) / (P_N(x_m) (x_i - x_m)) off the diagonal, -+N(N+1)/4 at the
// two end points and 0 elsewhere on the diagonal.
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  const double x2 = x * x;
  for (int k = 1; k < 40; ++k) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}

template <int N>
struct GllC {
  double x[N + 1], w[N + 1], D[(N + 1) * (N + 1)];
  constexpr GllC() : x(), w(), D() {
    const double pi = 3.14159265358979323846;
    double PN[N + 1] = {};
    for (int j = 0; j <= N; ++j) x[j] = -cx_cos(pi * j / N);
    for (int it = 0; it < 60; ++it)
      for (int j = 0; j <= N; ++j) {
        double p0 = 1.0, p1 = x[j];
        for (int k = 2; k <= N; ++k) {
          const double p2 = ((2.0 * k - 1.0) * x[j] * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        PN[j] = p1;
        if (j > 0 && j < N) x[j] -= (x[j] * p1 - p0) / ((N + 1) * p1);
      }
    for (int j = 0; j <= N; ++j) w[j] = 2.0 / (N * (N + 1.0) * PN[j] * PN[j]);
    for (int i = 0; i <= N; ++i)
      for (int m = 0; m <= N; ++m)
        D[i * (N + 1) + m] = i != m ? PN[i] / (PN[m] * (x[i] - x[m]))
                                    : (i == 0 ? -N * (N + 1.0) / 4.0 : (i == N ? N * (N + 1.0) / 4.0 : 0.0));
  }
};

// Compile-time table entries (arguments must be constants after unrolling).
template <int N>
__device__ __forceinline__ float gD(int i, int m) {
  constexpr GllC<N> g{};
  return (float)g.D[i * (N + 1) + m];
}
template <int N>
__device__ __forceinline__ float gW(int i) {
  constexpr GllC<N> g{};
  return (float)g.w[i];
}
template <int N>
__device__ __forceinline__ float gX(int i) {
  constexpr GllC<N> g{};
  return (float)g.x[i];
}

struct TileTables {
  float D[(kTileMaxN + 1) * (kTileMaxN + 1)];  // D[i*n+m] = l'_m(xi_i)
  float w[kTileMaxN + 1];
  float xi[kTileMaxN + 1];
};
__constant__ TileTables t_tab[kTileMaxN + 1];

// a + b never fused with a product feeding it (per-operation contract flag)
__device__ __forceinline__ float add_unfused(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}

__device__ __forceinline__ void kahan(float &s, float &c, float x) {
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// P1 vertices (QuadP1.h:98-108, HexP1.cpp:242-298): the multilinear weight
// of vertex v at (a, b[, t]) is prod (1 + s_i x_i) / 2^DIM with the vertex's
// signs s.  In 2-D the column coordinate t is eta.  Along the column the
// weights are linear in t, so the material is fixed by its values on the two
// end faces t = -1 and t = +1, where only that face's vertices take part.
template <int DIM>
__device__ __forceinline__ int p1_tsign(int v) {
  constexpr signed char ts[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  return DIM == 2 ? (v >= 2 ? 1 : -1) : ts[v];
}
// weight of vertex v on its own end face (t = its sign)
template <int DIM>
__device__ __forceinline__ float p1_face(int v, float a, float b) {
  if (DIM == 2) {
    const float sa = (v == 1 || v == 2) ? 1.f : -1.f;
    return 0.5f * (1.f + sa * a);
  }
  constexpr signed char hs[8][2] = {{-1, -1}, {-1, 1}, {1, 1}, {1, -1}, {-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  return 0.25f * (1.f + hs[v][0] * a) * (1.f + hs[v][1] * b);
}

// DG: every element's J^-1 (and, acoustic, its G = det J J^-1 J^-T) is
// diagonal (axis-aligned lattices), so the contractions keep their diagonal
// terms only; the terms dropped are exact zeros, so results are those of the
// general kernel.
template <int DIM, int N, int PHYS, int TX, int TY, int TZ, bool HO, bool DG>
__global__ void __launch_bounds__((TS<DIM, N, PHYS, TX, TY, TZ, HO>::BT))
    __attribute__((amdgpu_waves_per_eu(TS<DIM, N, PHYS, TX, TY, TZ, HO>::MINW))) k_tile(FusedArgs A) {
  using T = TS<DIM, N, PHYS, TX, TY, TZ, HO>;
  constexpr int n = T::n, C = T::C, ND = T::ND, NL = T::NL;
  extern __shared__ float lds[];
  float *su = lds;
  uint32_t *sw = reinterpret_cast<uint32_t *>(lds + T::O_SW);
  float *sF = lds + T::O_SF;
  float *sK = lds + T::O_SK;
  float *sE = lds + T::O_SE;
  int *sS = reinterpret_cast<int *>(lds + T::O_SS);  // {element, shot, stf_off, load_off} x k, count
  float *sD = lds + T::O_SD;  // GLL derivative matrix

  const int bid = blockIdx.x, nb = gridDim.x;
  int item = bid;
  if (nb >= 16 && (nb & 7) == 0) item = (bid & 7) * (nb >> 3) + (bid >> 3);
  if (item >= A.nitems) return;
  const WorkItem W = glb(A.items)[item];
  const auto &G = glb(A.groups)[W.group];
  const int64_t step = *glb(G.step) + A.koff;
  if (step >= G.nsteps) return;  // uniform across the block
  const TileInfo ti = glb(G.tiles)[W.first];
  const int tid = threadIdx.x;
  const TileTables &TT = t_tab[N];
  // descriptor fields, read once: G sits in global memory, so a field read
  // after a global store would be reloaded (the compiler cannot rule out
  // aliasing), and waiting for that load would wait for every store before it
  const gptr<float> gu = glb(G.u);
  const gptr<float> gw = glb(G.w);
  const gptr<float> gap = glb(G.ap);
  const gptr<uint32_t> gkc = glb(G.kc);
  const gptr<float> gacc = glb(G.acc);
  const gptr<const float> gmi = glb(G.mi);
  const uint32_t GG = (uint32_t)G.G, Sg = (uint32_t)G.S;
  const float gdt = G.dt;
  const gptr<const float> gstf = glb(G.stf), gload = glb(G.load);
  const gptr<const SrcDev> gsrc = glb(G.src);
  const int nsrc = G.nsrc;

  // ---- once per tile: node words, element data, sources inside the tile,
  // and the lattice of shot 0 (every later shot's lattice is loaded during
  // the previous shot's face passes).  Thread tid owns the tile words
  // k = tid + r BT, which come in node order, so a wave's global accesses
  // are coalesced; sl(r) is word k's lattice slot.  Outside the column phase
  // each thread writes and reads only its own slots of su.
  // inverse mass of the slots whose finish inputs are loaded ahead, read
  // once per tile for all S shots (not in the 80-register high-occupancy
  // variant, which reads it with the other finish inputs)
  constexpr bool kKeepMi = !HO;
  float mik[kKeepMi ? T::KF : 1];
  // lattice slot of the thread's word r: held in registers, in LDS, or
  // (lattice order) the word index itself
  constexpr bool kSlotRegs = T::NODE_ORDER && !T::SLOT_LDS;
  int slr[kSlotRegs ? T::NIT : 1];
  uint16_t *ssl = reinterpret_cast<uint16_t *>(lds + T::O_SL);
  const gptr<const uint16_t> slots = glb(G.tile_slot) + ti.node_off;
  auto sl = [&](int r) {
    const int k = tid + r * T::BT;
    if constexpr (!T::NODE_ORDER) {
      return k < T::NTL ? T::lds_slot(k) : -1;
    } else if constexpr (kSlotRegs) {
      return slr[r];
    } else {
      return k < T::NTL ? (int)ssl[k] : -1;  // written by this thread at staging
    }
  };
  {
    // every load that does not need the words goes out with them: element
    // data, the derivative table and the tile's sources; then the loads
    // that do (inverse mass, shot 0's lattice)
    const gptr<const uint32_t> words =
        (T::NODE_ORDER ? glb(G.tile_words) : glb(G.tile_words_lat)) + ti.node_off;
    uint32_t wv[T::NIT];
#pragma unroll
    for (int r = 0; r < T::NIT; ++r) {
      const int k = tid + r * T::BT;
      wv[r] = k < T::NTL ? words[k] : kSlotEmpty;
      if constexpr (T::NODE_ORDER) {
        const int slot = k < T::NTL ? T::lds_slot((int)slots[k]) : -1;
        if constexpr (kSlotRegs) slr[r] = slot;
        else if (k < T::NTL) ssl[k] = (uint16_t)slot;
      }
    }
    constexpr int NEI = (T::E * T::NED + T::BT - 1) / T::BT;
    static_assert(n * n <= T::BT, "derivative table staged one entry per thread");
    const gptr<const float> te = glb(G.telem) + (int64_t)ti.elem_off * T::NED;
    float ev[NEI];
#pragma unroll
    for (int q = 0; q < NEI; ++q) {
      const int i = tid + q * T::BT;
      ev[q] = i < T::E * T::NED ? te[i] : 0.f;
    }
    const float dv = tid < n * n ? TT.D[tid] : 0.f;
    if (tid < 64) {  // ordered compaction by wave 0 (ballot + prefix count)
      int cnt = 0;
      for (int j0 = 0; j0 < nsrc; j0 += 64) {
        const int j = j0 + tid;
        SrcDev sd{0, 0, 0, 0};
        bool in = false;
        if (j < nsrc) {
          sd = gsrc[j];
          in = sd.elem >= ti.elem_off && sd.elem < ti.elem_off + T::E;
        }
        const unsigned long long mask = __ballot(in);
        if (in) {
          const int pos = cnt + __popcll(mask & ((1ull << tid) - 1ull));
          if (pos < kTileMaxSrc) {
            sS[4 * pos + 0] = sd.elem - ti.elem_off;
            sS[4 * pos + 1] = sd.shot;
            sS[4 * pos + 2] = sd.stf_off;
            sS[4 * pos + 3] = sd.load_off;
          }
        }
        cnt += __popcll(mask);
      }
      if (tid == 0) sS[4 * kTileMaxSrc] = cnt < kTileMaxSrc ? cnt : kTileMaxSrc;
    }
    float uv[T::NIT][C];
#pragma unroll
    for (int r = 0; r < (kKeepMi ? T::KF : 0); ++r) {
      const bool fin0 = wv[r] != kSlotEmpty && (wv[r] & kFlagLast) && !A.ku_only;
      mik[r] = fin0 ? gmi[wv[r] & kNodeMask] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < T::NIT; ++r) {
      const bool ok = wv[r] != kSlotEmpty;
      const uint32_t g = ok ? (wv[r] & kNodeMask) : 0u;
#pragma unroll
      for (int c = 0; c < C; ++c) uv[r][c] = ok ? gu[c * GG + g] : 0.f;
    }
    if (tid < n * n) sD[tid] = dv;
#pragma unroll
    for (int q = 0; q < NEI; ++q) {
      const int i = tid + q * T::BT;
      if (i < T::E * T::NED) sE[i] = ev[q];
    }
#pragma unroll
    for (int r = 0; r < T::NIT; ++r) {
      const int l = sl(r);
      if (l >= 0) {
        sw[l] = wv[r];
#pragma unroll
        for (int c = 0; c < C; ++c) su[c * T::NTLP + l] = uv[r][c];
      }
    }
  }

  // ---- the column this thread owns in every chunk
  const int ec = tid / T::NCOL;
  const int col = tid - ec * T::NCOL;
  const int ci = col % n;
  const int cj = DIM == 3 ? col / n : 0;
  const bool lane = ec < T::EC;
  const float wl = DIM == 3 ? TT.w[ci] * TT.w[cj] : TT.w[ci];
  const float xa = TT.xi[ci], xb = DIM == 3 ? TT.xi[cj] : 0.f;
  const int pl = ci + (DIM == 3 ? n * cj : 0);  // element-local index of the column's node 0

  const bool q3 = (C >= 2) && (G.quirks & SEM_QUIRK_NEWMARK3D);  // {x, y, x}: x twice, z never
  // a_ is only read out (v = w - dt/2 a_, a_ itself): stored at the block's
  // last step and at the shot's last step, where a read-out can follow
  const bool wap = A.koff == A.ap_koff || step + 1 >= G.nsteps || G.ap_always;
  const gptr<const float> gext = glb(G.ext);  // coupled shots: external load
  const bool has_ext = G.ext != nullptr;

  __syncthreads();  // tile data and the first shot's lattice staged
  const int nts = sS[4 * kTileMaxSrc];

  for (uint32_t s = 0; s < Sg; ++s) {
    // ---- columns, one chunk of elements at a time
#pragma unroll 1
    for (int ch = 0; ch < T::NCH; ++ch) {
      if (ch > 0) __syncthreads();  // previous chunk has read its lateral fluxes
      const int e = ch * T::EC + ec;
      bool act = false;
      int ebase = 0;
      if (lane && e < T::E) {
        const int ea = e % TX, eb = (e / TX) % TY, ez = DIM == 3 ? e / (TX * TY) : 0;
        act = ea < ti.ex && eb < ti.ey && (DIM == 2 || ez < ti.ez);
        ebase = ea * N + T::LXP * (eb * N) + (DIM == 3 ? T::LXYP * (ez * N) : 0);
      }
      const int cb = ebase + ci + (DIM == 3 ? T::LXP * cj : 0);  // lattice slot of node 0
      float *myF = sF + ec * T::SFE;
      float Fz[C][n];
      if (act) {
        const float *ed = sE + e * T::NED;
        float Dx[n], Dy[n];  // rows ci and cj of D, from LDS
#pragma unroll
        for (int m = 0; m < n; ++m) {
          Dx[m] = sD[ci * n + m];
          Dy[m] = DIM == 3 ? sD[cj * n + m] : 0.f;
        }
        float uc[C][n];
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
          for (int k = 0; k < n; ++k) uc[c][k] = su[c * T::NTLP + cb + k * T::CS];
        // vertex material at t: mA + mB t, from its end-face values
        float mA[T::KP], mB[T::KP];
        {
          float wf[T::NV];
#pragma unroll
          for (int v = 0; v < T::NV; ++v) wf[v] = 0.5f * p1_face<DIM>(v, xa, xb);
#pragma unroll
          for (int q = 0; q < T::KP; ++q) {
            float lo = 0.f, hi = 0.f;
#pragma unroll
            for (int v = 0; v < T::NV; ++v) {
              const float P = ed[T::NG + v * T::KP + q];
              if (p1_tsign<DIM>(v) < 0) lo = fmaf(wf[v], P, lo);
              else hi = fmaf(wf[v], P, hi);
            }
            mA[q] = hi + lo;
            mB[q] = hi - lo;
          }
          if constexpr (PHYS != SEM_ACOUSTIC && DIM == 3) {
            // every Elastic3D coefficient is proportional to rho: scaling rho
            // by w_i w_j det J (and by w_k below) puts the quadrature weight
            // into the stress, so the flux needs no scaling of its own
            const float sw = wl * ed[DIM * DIM];
            mA[0] *= sw;
            mB[0] *= sw;
          }
        }
#pragma unroll
        for (int k = 0; k < n; ++k) {
          const float tk = gX<N>(k), wk = gW<N>(k);
          float g[C][DIM];
#pragma unroll
          for (int c = 0; c < C; ++c) {
            const float *rx = su + c * T::NTLP + (cb - ci) + k * T::CS;
            float s0 = 0.f;
#pragma unroll
            for (int m = 0; m < n; ++m) s0 = fmaf(Dx[m], rx[m], s0);
            g[c][0] = s0;
            if (DIM == 3) {
              const float *ry = su + c * T::NTLP + (cb - T::LXP * cj) + k * T::CS;
              float s1 = 0.f;
#pragma unroll
              for (int m = 0; m < n; ++m) s1 = fmaf(Dy[m], ry[m * T::LXP], s1);
              g[c][NL > 1 ? 1 : 0] = s1;
            }
            float sz = 0.f;
#pragma unroll
            for (int m = 0; m < n; ++m) sz = fmaf(gD<N>(k, m), uc[c][m], sz);
            g[c][DIM - 1] = sz;
          }
          float F[C][DIM];
          if constexpr (PHYS == SEM_ACOUSTIC) {
            const float vp = fmaf(mB[0], tk, mA[0]);
            const float sc = wl * wk * vp * vp;  // Scalar.cpp:45
            const float *Gs = ed;
            if constexpr (DG && DIM == 2) {
              F[0][0] = sc * (Gs[0] * g[0][0]);
              F[0][1] = sc * (Gs[2] * g[0][1]);
            } else if constexpr (DG) {
              F[0][0] = sc * (Gs[0] * g[0][0]);
              F[0][NL > 1 ? 1 : 0] = sc * (Gs[3] * g[0][1]);
              F[0][DIM - 1] = sc * (Gs[5] * g[0][DIM - 1]);
            } else if constexpr (DIM == 2) {
              F[0][0] = sc * (Gs[0] * g[0][0] + Gs[1] * g[0][1]);
              F[0][1] = sc * (Gs[1] * g[0][0] + Gs[2] * g[0][1]);
            } else {
              F[0][0] = sc * (Gs[0] * g[0][0] + Gs[1] * g[0][1] + Gs[2] * g[0][DIM - 1]);
              F[0][NL > 1 ? 1 : 0] = sc * (Gs[1] * g[0][0] + Gs[3] * g[0][1] + Gs[4] * g[0][DIM - 1]);
              F[0][DIM - 1] = sc * (Gs[2] * g[0][0] + Gs[4] * g[0][1] + Gs[5] * g[0][DIM - 1]);
            }
          } else {
            const float *Ji = ed;  // Jinv[a*DIM+b]
            const float wdet = wl * wk * ed[DIM * DIM];
            float cm[9];
            if constexpr (DIM == 2) {  // C11, C13, C33, C55
#pragma unroll
              for (int q = 0; q < 4; ++q) cm[q] = fmaf(mB[q], tk, mA[q]);
            } else {  // Elastic3D.cpp:42-52
              float q6[6];
#pragma unroll
              for (int q = 0; q < 6; ++q) q6[q] = fmaf(mB[q], tk, mA[q]);
              q6[0] *= wk;
              const float c11 = q6[0] * q6[2] * q6[2], c33 = q6[0] * q6[1] * q6[1];
              const float c44 = q6[0] * q6[3] * q6[3], c66 = q6[0] * q6[4] * q6[4];
              const float c13 = q6[5] * (c11 - 2.f * c44);
              cm[0] = c11; cm[1] = c11 - 2.f * c66; cm[2] = c13; cm[3] = c11; cm[4] = c13;
              cm[5] = c33; cm[6] = c44; cm[7] = c44; cm[8] = c66;
            }
            float gx[DIM][DIM];
#pragma unroll
            for (int c = 0; c < DIM; ++c)
#pragma unroll
              for (int a = 0; a < DIM; ++a) {
                float v = 0.f;
                if constexpr (DG) {
                  v = Ji[a * DIM + a] * g[c < C ? c : 0][a];
                } else {
#pragma unroll
                  for (int b = 0; b < DIM; ++b) v = fmaf(Ji[a * DIM + b], g[c < C ? c : 0][b], v);
                }
                gx[c][a] = v;
              }
            float sig[DIM][DIM];
            if constexpr (DIM == 2) {  // Elastic2D.cpp:53-82
              // shear sums unfused: with diagonal J^-1 the terms are plain
              // products the compiler could otherwise fuse into the sum, and
              // the variants would round differently
              const float exx = gx[0][0], eyy = gx[1][1], sh = add_unfused(gx[0][1], gx[1][0]);
              // explicit fmaf: the same fused products in every variant
              sig[0][0] = fmaf(cm[1], eyy, cm[0] * exx);
              sig[1][1] = fmaf(cm[2], eyy, cm[1] * exx);
              sig[0][1] = sig[1][0] = cm[3] * sh;
            } else {  // Elastic3D.cpp:76-83, :106-118
              constexpr int Y = DIM > 2 ? 2 : 1;
              const float e0 = gx[0][0], e1 = gx[1][1], e2 = gx[Y][Y];
              const float e3 = add_unfused(gx[1][Y], gx[Y][1]), e4 = add_unfused(gx[0][Y], gx[Y][0]);
              const float e5 = add_unfused(gx[0][1], gx[1][0]);  // unfused, as in 2-D
              sig[0][0] = fmaf(cm[2], e2, fmaf(cm[1], e1, cm[0] * e0));  // fixed fusion, as in 2-D
              sig[1][1] = fmaf(cm[4], e2, fmaf(cm[3], e1, cm[1] * e0));
              sig[Y][Y] = fmaf(cm[5], e2, fmaf(cm[4], e1, cm[2] * e0));
              sig[1][Y] = sig[Y][1] = cm[6] * e3;
              sig[0][Y] = sig[Y][0] = cm[7] * e4;
              sig[0][1] = sig[1][0] = cm[8] * e5;
            }
#pragma unroll
            for (int c = 0; c < DIM; ++c)
#pragma unroll
              for (int a = 0; a < DIM; ++a) {
                float v = 0.f;
                if constexpr (DG) {
                  v = Ji[a * DIM + a] * sig[c][a];
                } else {
#pragma unroll
                  for (int b = 0; b < DIM; ++b) v = fmaf(Ji[b * DIM + a], sig[c][b], v);
                }
                F[c < C ? c : 0][a] = DIM == 3 ? v : wdet * v;
              }
          }
          const int p = pl + T::ES * k;
#pragma unroll
          for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int a = 0; a < NL; ++a) myF[(c * NL + a) * ND + p] = F[c][a];
            Fz[c][k] = F[c][DIM - 1];
          }
        }
      }
      __syncthreads();
      float f[C][n];
      if (act) {
        float Dcx[n], Dcy[n];  // columns ci and cj of D
#pragma unroll
        for (int m = 0; m < n; ++m) {
          Dcx[m] = sD[m * n + ci];
          Dcy[m] = DIM == 3 ? sD[m * n + cj] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < n; ++k) {
          const int p = pl + T::ES * k;
          const bool dir = (sw[cb + k * T::CS] & kFlagDirichlet) != 0;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            const float *Fx = myF + (c * NL) * ND + (p - ci);  // F_x along the x line
            float ku = 0.f;
#pragma unroll
            for (int m = 0; m < n; ++m) ku = fmaf(Dcx[m], Fx[m], ku);
            if (DIM == 3) {
              const float *Fy = myF + (c * NL + (NL > 1 ? 1 : 0)) * ND + (p - n * cj);
#pragma unroll
              for (int m = 0; m < n; ++m) ku = fmaf(Dcy[m], Fy[m * n], ku);
            }
#pragma unroll
            for (int m = 0; m < n; ++m) ku = fmaf(gD<N>(m, k), Fz[c][m], ku);
            f[c][k] = dir ? 0.f : -ku;
          }
        }
        for (int j = 0; j < nts; ++j) {
          if (sS[4 * j + 0] != e || sS[4 * j + 1] != (int)s) continue;
          const float R = gstf[sS[4 * j + 2] + step];
          const gptr<const float> ld = gload + sS[4 * j + 3];
#pragma unroll
          for (int c = 0; c < C; ++c)
#pragma unroll
            for (int k = 0; k < n; ++k) f[c][k] = fmaf(ld[c * ND + pl + T::ES * k], R, f[c][k]);
        }
      }
      if (T::ALIAS && ch == T::NCH - 1) __syncthreads();  // lateral fluxes read before sK overwrites them
      if (act) {
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
          for (int k = 0; k < n; ++k) sK[e * T::SKE + c * ND + pl + T::ES * k] = f[c][k];
      }
    }
    __syncthreads();

    // ---- finish inputs of the first KF slots and the next shot's lattice
    // are loaded here so that their latency overlaps the face passes
    struct SlotIn {
      uint32_t w;
      float mi, prev[C], vw[C];  // vw: w = v + dt/2 a_
      uint32_t kc[C];
    };
    auto load_slot = [&](int l, SlotIn &in, bool keep) {
      in.w = l >= 0 ? sw[l] : kSlotEmpty;
      const bool valid = in.w != kSlotEmpty;
      const uint32_t g = valid ? (in.w & kNodeMask) : 0u;
      const bool first = (in.w & kFlagFirst) != 0;
      const bool fin = valid && (in.w & kFlagLast) && !A.ku_only;
      in.mi = keep ? 0.f : (fin ? gmi[g] : 0.f);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const uint32_t idx = (s * C + c) * GG + g;
        in.prev[c] = !valid ? 0.f : !first ? gacc[idx] : has_ext ? gext[idx] : 0.f;
        in.vw[c] = fin ? gw[idx] : 0.f;
        in.kc[c] = fin ? gkc[idx] : 0u;
      }
    };
    constexpr int KF = T::KF;
    SlotIn pre[KF];
#pragma unroll
    for (int r = 0; r < KF; ++r) {
      load_slot(sl(r), pre[r], kKeepMi);
      if constexpr (kKeepMi) pre[r].mi = mik[r];
    }
    const bool more = s + 1 < Sg;
    float nu[T::NIT][C];
#pragma unroll
    for (int r = 0; r < T::NIT; ++r) {
      const int l = sl(r);
      const uint32_t w = (more && l >= 0) ? sw[l] : kSlotEmpty;
      const uint32_t g = w != kSlotEmpty ? (w & kNodeMask) : 0u;
#pragma unroll
      for (int c = 0; c < C; ++c) nu[r][c] = w != kSlotEmpty ? gu[((s + 1) * C + c) * GG + g] : 0.f;
    }

    // ---- shared faces: x, then y, then z (fixed order)
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
      constexpr int NF = ND / n;
      for (int i = tid; i < T::E * NF; i += T::BT) {
        const int e = i / NF, q = i - e * NF;
        const int ea = e % TX, eb = (e / TX) % TY, ez = DIM == 3 ? e / (TX * TY) : 0;
        const bool valid = ea < ti.ex && eb < ti.ey && (DIM == 2 || ez < ti.ez);
        int lo, p, pn;
        if (ax == 0) {         // node (0, q) gains the x-neighbour's (N, q)
          lo = ea > 0 ? e - 1 : -1;
          p = q * n;
          pn = p + N;
        } else if (ax == 1) {  // (i, 0, k) gains (i, N, k)
          lo = eb > 0 ? e - TX : -1;
          p = DIM == 3 ? (q % n) + n * n * (q / n) : q;
          pn = p + n * N;
        } else {               // (i, j, 0) gains (i, j, N)
          lo = ez > 0 ? e - TX * TY : -1;
          p = q;
          pn = q + n * n * N;
        }
        if (valid && lo >= 0) {
#pragma unroll
          for (int c = 0; c < C; ++c) sK[e * T::SKE + c * ND + p] += sK[lo * T::SKE + c * ND + pn];
        }
      }
      __syncthreads();
    }

    // ---- finish every lattice slot: total of the owner element's face-summed
    // result, accumulation over colours, Newmark update at the last colour
    auto finish_slot = [&](int l, const SlotIn &in) {
      if (in.w == kSlotEmpty) return;
      const int z = DIM == 3 ? l / T::LXYP : 0, rl = l - z * T::LXYP;
      const int y = rl / T::LXP, x = rl - y * T::LXP;
      const int ea = min(x / N, ti.ex - 1), eb = min(y / N, ti.ey - 1);
      const int ez = DIM == 3 ? min(z / N, ti.ez - 1) : 0;
      const int e = ea + TX * (eb + TY * ez);
      const int p = (x - ea * N) + n * ((y - eb * N) + n * (z - ez * N));
      const uint32_t g = in.w & kNodeMask;
      const bool first = (in.w & kFlagFirst) != 0;
      const bool fin = (in.w & kFlagLast) && !A.ku_only;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const uint32_t idx = (s * C + c) * GG + g;
        const float tot = sK[e * T::SKE + c * ND + p];
        const float t = (first && !has_ext) ? tot : in.prev[c] + tot;
        if (!fin) {
          gacc[idx] = t;
          continue;
        }
        if (q3 && c == 2) continue;
        const float a = in.mi * t;
        float w = in.vw[c], kw = comp_v(in.kc[c]);
        float u = su[c * T::NTLP + l], ku = comp_u(in.kc[c]);
        const int reps = (q3 && c == 0) ? 2 : 1;
        for (int rep = 0; rep < reps; ++rep) {
          kahan(w, kw, gdt * a);
          kahan(u, ku, gdt * w);
        }
        gu[idx] = u; gw[idx] = w; gkc[idx] = pack_comp(ku, kw);
        if (wap) gap[idx] = a;
      }
    };
    // every finish input loaded ahead (and the next lattice) has arrived:
    // one wait on all paths here, so that no slot's finish below waits for
    // the stores of the slots before it (stores count on the same counter,
    // and the loads were issued under per-slot branches the compiler's
    // counting cannot follow)
    // (not in the high-occupancy variant, whose later slots load in the
    // finish loop: measured 0.5 % slower with the wait)
    if constexpr (!HO) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0) expcnt(7) lgkmcnt(15)
#pragma unroll
    for (int r = 0; r < KF; ++r) finish_slot(sl(r), pre[r]);
    constexpr int KB = C == 1 ? 4 : 2;
#pragma unroll 1
    for (int r0 = KF; r0 < T::NIT; r0 += KB) {
      SlotIn in[KB];
      int ls[KB];  // slots of the words k = tid + (r0 + r) BT (runtime r0: read again)
#pragma unroll
      for (int r = 0; r < KB; ++r) {
        const int k = tid + (r0 + r) * T::BT;
        ls[r] = (r0 + r < T::NIT && k < T::NTL)
                    ? (!T::NODE_ORDER ? T::lds_slot(k) : T::SLOT_LDS ? (int)ssl[k] : T::lds_slot((int)slots[k]))
                    : -1;
        load_slot(ls[r], in[r], false);
      }
#pragma unroll
      for (int r = 0; r < KB; ++r) finish_slot(ls[r], in[r]);
    }
    if (!more) break;
    // own slots only: u of this shot has been read above
#pragma unroll
    for (int r = 0; r < T::NIT; ++r) {
      const int l = sl(r);
      if (l >= 0) {
#pragma unroll
        for (int c = 0; c < C; ++c) su[c * T::NTLP + l] = nu[r][c];
      }
    }
    __syncthreads();  // next lattice staged; sK and sF free again
  }
}

template <int N>
void fill_tables(TileTables &t) {
  constexpr GllC<N> g{};
  for (int i = 0; i < (N + 1) * (N + 1); ++i) t.D[i] = (float)g.D[i];
  for (int i = 0; i <= N; ++i) {
    t.w[i] = (float)g.w[i];
    t.xi[i] = (float)g.x[i];
  }
}

}  // namespace

int init_tile_tables(int device) {
  static bool done[64] = {false};
  if (device < 0 || device >= 64) { set_error("device index"); return -1; }
  if (done[device]) return 0;
  TileTables host[kTileMaxN + 1];
  memset(host, 0, sizeof host);
  // the same compile-time tables the kernels use as literals
  fill_tables<1>(host[1]);
  fill_tables<2>(host[2]);
  fill_tables<3>(host[3]);
  fill_tables<4>(host[4]);
  fill_tables<5>(host[5]);
  fill_tables<6>(host[6]);
  fill_tables<7>(host[7]);
  fill_tables<8>(host[8]);
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(t_tab), host, sizeof host, 0, hipMemcpyHostToDevice);
  if (e != hipSuccess) { set_error(std::string("tile tables: ") + hipGetErrorString(e)); return -2; }
  done[device] = true;
  return 0;
}

static bool same_shape(const int *T, int tx, int ty, int tz) {
  return T[0] == tx && T[1] == ty && T[2] == tz;
}

TileKernelInfo tile_kernel_info(int dim, int order, int physics, const int *T) {
  TileKernelInfo r{0, 0};
#define TKI(D, N, X, Y, Z)                                                                  \
  if (dim == D && order == N && same_shape(T, X, Y, D == 3 ? Z : 1)) {                      \
    if (physics == SEM_ACOUSTIC) {                                                          \
      using S = TS<D, N, 0, X, Y, Z>;                                                       \
      r.lds_bytes = sizeof(float) * (size_t)S::LDS;                                         \
      r.threads = S::BT;                                                                    \
    } else {                                                                                \
      using S = TS<D, N, 1, X, Y, Z>;                                                       \
      r.lds_bytes = sizeof(float) * (size_t)S::LDS;                                         \
      r.threads = S::BT;                                                                    \
    }                                                                                       \
  }
  SEM_TILE_KERNELS(TKI)
#undef TKI
  return r;
}

int tile_kernel_shapes(int dim, int order, int physics, int *shapes, int max) {
  int k = 0;
#define TSH(D, N, X, Y, Z)                                                                  \
  if (dim == D && order == N && k < max) {                                                  \
    const int T[3] = {X, Y, D == 3 ? Z : 1};                                                \
    if (tile_kernel_info(dim, order, physics, T).lds_bytes <= 160 * 1024) {                 \
      shapes[3 * k] = T[0];                                                                 \
      shapes[3 * k + 1] = T[1];                                                             \
      shapes[3 * k + 2] = T[2];                                                             \
      ++k;                                                                                  \
    }                                                                                       \
  }
  SEM_TILE_KERNELS(TSH)
#undef TSH
  return k;
}

// Sets the dynamic-LDS limit of a kernel above 64 KiB and returns how many
// of its workgroups one CU holds.
template <int D, int N, int P, int X, int Y, int Z, bool HO>
static int prep_one(int *per_cu) {
  using S = TS<D, N, P, X, Y, Z, HO>;
  const size_t shm = sizeof(float) * (size_t)S::LDS;
  if (shm > 160 * 1024) { set_error("tile does not fit in LDS"); return -1; }
  const void *fn = (const void *)k_tile<D, N, P, X, Y, Z, HO, false>;
  if (shm > 64 * 1024) {
    const void *fns[2] = {fn, (const void *)k_tile<D, N, P, X, Y, Z, HO, true>};
    for (const void *f : fns) {
      hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
      if (e != hipSuccess) { set_error(std::string("tile LDS attribute: ") + hipGetErrorString(e)); return -2; }
    }
  }
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, S::BT, shm);
  if (e != hipSuccess) { set_error(std::string("tile occupancy: ") + hipGetErrorString(e)); return -2; }
  *per_cu = nb > 0 ? nb : 1;
  return 0;
}

// Workgroups per CU of the normal and the high-occupancy variant of one
// acoustic tile kernel (cached), and the device's CU count.
template <int D, int N, int P, int X, int Y, int Z>
static const int *tile_occupancy() {
  static int occ[3] = {0, 0, 0};
  if (!occ[0]) {
    int dev = 0, cus = 0, lo = 0, hi = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1 ||
        prep_one<D, N, P, X, Y, Z, false>(&lo) || prep_one<D, N, P, X, Y, Z, true>(&hi))
      return nullptr;
    occ[1] = lo;
    occ[2] = hi;
    occ[0] = cus;
  }
  return occ;
}

template <int D, int N, int P, int X, int Y, int Z>
static int prep_tile(int *per_cu) {
  const int rc = prep_one<D, N, P, X, Y, Z, false>(per_cu);
  if (rc) return rc;
  if constexpr (P == SEM_ACOUSTIC) {
    if (!tile_occupancy<D, N, P, X, Y, Z>()) { set_error("tile occupancy query failed"); return -2; }
  }
  return 0;
}

int tile_kernel_prepare(int dim, int order, int physics, const int *T, int *per_cu) {
  int rc = 1;
#define TPR(D, N, X, Y, Z)                                                                  \
  if (rc == 1 && dim == D && order == N && same_shape(T, X, Y, D == 3 ? Z : 1))             \
    rc = physics == SEM_ACOUSTIC ? prep_tile<D, N, 0, X, Y, Z>(per_cu)                      \
                                 : prep_tile<D, N, 1, X, Y, Z>(per_cu);
  SEM_TILE_KERNELS(TPR)
#undef TPR
  if (rc == 1) { set_error("tile kernel not compiled for this order/shape"); return -1; }
  return rc;
}

// Acoustic launches take the high-occupancy variant when it needs fewer
// rounds of workgroups over the CUs: one of its workgroups runs about 1.2x
// as long as a normal one (measured on 96^3 and 128^3 hex meshes), so it
// pays only where the extra resident workgroup saves a round (1.4x faster
// at 703 tiles per colour on 256 CUs, 4 % slower at 312).  Both variants
// do the same arithmetic in the same order: results are identical.
// SEM_TILE_OCCUPANCY=normal|high forces one variant (tests compare them).
template <int D, int N, int P, int X, int Y, int Z, bool DG>
static int go_tile(int grid, const FusedArgs &a, hipStream_t st) {
  using S = TS<D, N, P, X, Y, Z, false>;
  const size_t shm = sizeof(float) * (size_t)S::LDS;
  if constexpr (P == SEM_ACOUSTIC) {
    const int *occ = tile_occupancy<D, N, P, X, Y, Z>();
    if (!occ) { set_error("tile occupancy query failed"); return -2; }
    const int64_t lo = ((int64_t)grid + (int64_t)occ[0] * occ[1] - 1) / ((int64_t)occ[0] * occ[1]);
    const int64_t hi = ((int64_t)grid + (int64_t)occ[0] * occ[2] - 1) / ((int64_t)occ[0] * occ[2]);
    const char *force = getenv("SEM_TILE_OCCUPANCY");
    const bool high = force && !strcmp(force, "high")     ? true
                      : force && !strcmp(force, "normal") ? false
                                                          : 6 * hi < 5 * lo;  // 1.2 hi < lo
    if (high) {
      using SH = TS<D, N, P, X, Y, Z, true>;
      hipLaunchKernelGGL((k_tile<D, N, P, X, Y, Z, true, DG>), dim3(grid), dim3(SH::BT),
                         sizeof(float) * (size_t)SH::LDS, st, a);
      return 0;
    }
  }
  hipLaunchKernelGGL((k_tile<D, N, P, X, Y, Z, false, DG>), dim3(grid), dim3(S::BT), shm, st, a);
  return 0;
}

int launch_tile(int dim, int order, int physics, const int *T, bool diag, int grid, const FusedArgs &a,
                hipStream_t st) {
  if (a.nitems <= 0 || grid <= 0) return 0;
  int rc = 1;
#define TLN(D, N, X, Y, Z)                                                                  \
  if (rc == 1 && dim == D && order == N && same_shape(T, X, Y, D == 3 ? Z : 1))             \
    rc = physics == SEM_ACOUSTIC                                                            \
             ? (diag ? go_tile<D, N, 0, X, Y, Z, true>(grid, a, st)                         \
                     : go_tile<D, N, 0, X, Y, Z, false>(grid, a, st))                       \
             : (diag ? go_tile<D, N, 1, X, Y, Z, true>(grid, a, st)                         \
                     : go_tile<D, N, 1, X, Y, Z, false>(grid, a, st));
  SEM_TILE_KERNELS(TLN)
#undef TLN
  if (rc == 1) { set_error("tile kernel not compiled for this order/shape"); return -1; }
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { set_error(std::string("tile step: ") + hipGetErrorString(e)); return -2; }
  return 0;
}

}  // namespace sem
