// LDS layout and thread shape of the tile kernels (tile_kernels.hip).  Plain
// constexpr C++, so that host code and host tests can check the layout the
// kernels use.
#pragma once

#include "../../include/salvus_sem.h"

#if defined(__HIPCC__)
#define SEM_HD __host__ __device__ __forceinline__
#else
#define SEM_HD inline
#endif

namespace sem {

// Sources (over all shots of a group) one tile may hold; groups with more
// in some tile take the element kernel.
constexpr int kTileMaxSrc = 64;

// Instantiated tile kernels: (dim, order, TX, TY, TZ).
#define SEM_TILE_KERNELS(X)                                                               \
  X(3, 2, 4, 2, 2) X(3, 3, 4, 2, 2) X(3, 4, 4, 2, 2) X(3, 5, 4, 2, 2)                     \
  X(3, 2, 2, 2, 2) X(3, 3, 2, 2, 2) X(3, 4, 2, 2, 2) X(3, 5, 2, 2, 2)                     \
  X(3, 3, 4, 4, 2) X(3, 4, 4, 4, 2)                                                       \
  X(2, 2, 8, 8, 1) X(2, 3, 8, 8, 1) X(2, 4, 8, 8, 1) X(2, 5, 8, 8, 1) X(2, 6, 8, 8, 1)    \
  X(2, 2, 16, 8, 1) X(2, 3, 16, 8, 1) X(2, 4, 16, 8, 1) X(2, 5, 16, 8, 1) X(2, 6, 16, 8, 1)

// Padding of the LDS lattice of the order-4 3-D elastic tiles (extra words
// per lattice row and per plane), chosen by counting the bank conflicts of
// the column phase's LDS addresses (32 banks of 4 bytes per half-wave, and 64
// per wave): unpadded, most x-derivative and node-column reads of a wave hit
// one bank twice, because rows 17 words apart and elements 4 words apart
// land on the same banks.
constexpr int lattice_pad_row(int tx) { return tx == 4 ? 4 : tx == 2 ? 2 : 0; }
constexpr int lattice_pad_plane(int tx, int ty) { return tx == 4 && ty == 2 ? 7 : tx == 2 && ty == 2 ? 5 : 0; }
// smallest v >= a with v = r (mod m)
constexpr int round_to_residue(int a, int r, int m) { return a + (((r - a) % m) + m) % m; }

// LDS layout and thread shape of one tile kernel.  PADDED spreads the column
// phase's LDS accesses over the banks: padded lattice rows and planes (the
// lattice slot of node (x, y, z) is x + LXP y + LXYP z), and per-element
// strides of the lateral fluxes and the element results equal to NCOL modulo
// 32 and 64, so that column thread t writes and reads bank (t + const) mod 32.
// The slots in the tile tables are unpadded (x + LX y + LXY z).
template <int DIM, int N, int PHYS, int TX, int TY, int TZ, bool HO, bool PADDED>
struct TSL {
  static constexpr int n = N + 1;
  static constexpr int ND = DIM == 2 ? n * n : n * n * n;  // nodes per element
  static constexpr int NCOL = ND / n;                       // columns per element
  static constexpr int C = PHYS == SEM_ACOUSTIC ? 1 : DIM;
  static constexpr int NL = DIM - 1;                        // lateral axes
  static constexpr int LX = TX * N + 1, LY = TY * N + 1, LZ = DIM == 3 ? TZ * N + 1 : 1;
  static constexpr int LXY = LX * LY, NTL = LXY * LZ;       // lattice slots (tile words)
  static constexpr bool PAD = PADDED && DIM == 3 && C == 3 && N == 4 && !HO;
  static constexpr int LXP = PAD ? LX + lattice_pad_row(TX) : LX;
  static constexpr int LXYP = PAD ? LXP * LY + lattice_pad_plane(TX, TY) : LXY;
  static constexpr int NTLP = DIM == 3 ? LXYP * LZ : LXP * LY;  // LDS lattice words
  static constexpr int CS = DIM == 3 ? LXYP : LXP;              // lattice stride of the column axis
  static constexpr int SFE = PAD ? round_to_residue(C * NL * ND, NCOL % 32, 32) : C * NL * ND;
  static constexpr int SKE = PAD ? round_to_residue(C * ND, NCOL % 64, 64) : C * ND;
  static constexpr int ES = DIM == 3 ? n * n : n;           // element-local stride of the column axis
  static constexpr int E = TX * TY * (DIM == 3 ? TZ : 1);
  // column chunks of at most 512 threads, all chunks equally full; the
  // workgroup keeps the threads of the largest chunk (they all gather and
  // finish lattice slots)
  static constexpr int EC0 = E < 512 / NCOL ? E : 512 / NCOL;
  static constexpr int NCH = (E + EC0 - 1) / EC0;
  static constexpr int EC = (E + NCH - 1) / NCH;             // elements per column chunk
  static constexpr int BT0 = ((EC0 * NCOL + 63) / 64) * 64;
  // Threads: at least those of a chunk, and enough that every thread's
  // lattice slots fit the registers that hold finish inputs loaded ahead
  // (KFMAX slots), so that no slot waits for its loads in the finish.
  static constexpr int KFMAX = 64 / (5 * C + 2);  // registers per slot held ahead: budgeted as before
  static constexpr int BTF = (((NTL + KFMAX - 1) / KFMAX + 63) / 64) * 64;
  static constexpr int BT = BTF > BT0 ? (BTF < 1024 ? BTF : 1024) : BT0;
  static constexpr int NIT = (NTL + BT - 1) / BT;           // lattice slots per thread
  static constexpr int NG = PHYS == SEM_ACOUSTIC ? DIM * (DIM + 1) / 2 : DIM * DIM + 1;
  static constexpr int KP = PHYS == SEM_ACOUSTIC ? 1 : (DIM == 2 ? 4 : 6);
  static constexpr int NV = DIM == 2 ? 4 : 8;
  static constexpr int NED = NG + NV * KP;                  // floats of element data
  // LDS layout (floats): u lattice | words | lateral fluxes of one chunk |
  // element-local results | element data | tile source list (4 ints each).
  // With ALIAS the results of all chunks but the last come first and the
  // last chunk's results overwrite its own lateral fluxes once every column
  // has read them (one more barrier, less LDS): taken with a single chunk,
  // and with several where it fits more workgroups in a CU's 160 KiB.
  static constexpr int O_SW = C * NTLP;
  static constexpr int LDS_REST = O_SW + NTLP + E * NED + 4 * kTileMaxSrc + 1 + n * n;
  static constexpr int LDS_SEP = LDS_REST + EC * SFE + E * SKE;
  static constexpr int LDS_ALIAS = LDS_REST + EC * SFE + (NCH - 1) * EC * SKE;
  static constexpr bool ALIAS = (NCH == 1 && EC * SKE <= EC * SFE) || 40960 / LDS_ALIAS > 40960 / LDS_SEP;
  static constexpr int O_SF = ALIAS ? O_SW + NTLP + (NCH - 1) * EC * SKE : O_SW + NTLP;
  static constexpr int O_SK = ALIAS ? O_SW + NTLP : O_SF + EC * SFE;
  static constexpr int O_SE = ALIAS ? O_SF + EC * SFE : O_SK + E * SKE;
  // Register budget: 4 waves per SIMD (<= 128 VGPRs) with the finish
  // inputs of every slot loaded ahead (up to KFMAX); the high-occupancy
  // variant HO (acoustic tiles) takes 6 (<= 80 VGPRs, a third workgroup per
  // CU where LDS allows) and loads three slots per thread ahead.
  static constexpr int MINW = HO ? 6 : 4;
  static constexpr int KFR = HO ? 3 : KFMAX;
  static constexpr int KF = NIT < KFR ? NIT : KFR;
  static constexpr int O_SS = O_SE + E * NED;
  static constexpr int O_SD = O_SS + 4 * kTileMaxSrc + 1;
  // Threads take the tile words in node order (coalesced HBM accesses),
  // except in the high-occupancy variant, whose 80-register budget and
  // short finish favour lattice order (thread slot = word index, no slot
  // table).  The small 3-D elastic tiles, whose register budget the slots
  // would exceed, keep each word's lattice slot in LDS (16 bits per word).
  static constexpr bool NODE_ORDER = !HO;
  static constexpr bool SLOT_LDS = NODE_ORDER && C == 3 && E <= 8;
  static constexpr int O_SL = O_SD + n * n;
  static constexpr int LDS = O_SL + (SLOT_LDS ? (NTL + 1) / 2 : 0);
  // padded LDS lattice slot of the unpadded slot l
  SEM_HD static int lds_slot(int l) {
    if constexpr (!PAD) {
      return l;
    } else {
      const int z = l / LXY, r = l - z * LXY, y = r / LX;
      return (r - y * LX) + LXP * y + LXYP * z;
    }
  }
};

// The padded layout where it leaves at least as many workgroups per CU.
template <int DIM, int N, int PHYS, int TX, int TY, int TZ, bool HO = false>
struct TS : TSL<DIM, N, PHYS, TX, TY, TZ, HO,
                40960 / TSL<DIM, N, PHYS, TX, TY, TZ, HO, true>::LDS >=
                    40960 / TSL<DIM, N, PHYS, TX, TY, TZ, HO, false>::LDS> {};

}  // namespace sem
