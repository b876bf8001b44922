// Host check of the tile kernels' LDS layouts (salvus_amd/csrc/tile_layout.h)
// for every instantiated tile kernel, both physics and both occupancy
// variants:
//  - the regions (lattice, words, fluxes, results, element data, sources,
//    derivative table, slot table) fit the layout and do not overlap, except
//    the aliasing the layout states (the last chunk's results over its own
//    fluxes);
//  - the layout fits a CU's 160 KiB, and the padded layout is taken only
//    where it leaves at least as many workgroups per CU;
//  - lds_slot maps the unpadded lattice slots one to one onto padded slots of
//    the same node (x + LXP y + LXYP z), inside the padded lattice;
//  - padded layouts: column thread t stores its fluxes and results in bank
//    (t + const) mod 32 (no two lanes of a half-wave share a bank), and the
//    column phase's lattice reads (x-derivative rows, node columns) conflict
//    in fewer cycles than with the unpadded lattice, counted with 32 banks of
//    4 bytes per half-wave.
// Prints one line per layout and the first 20 failures; exits non-zero if any
// check fails.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>

#include "tile_layout.h"

using namespace sem;

static int failures = 0;
#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond) && ++failures <= 20) {                     \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
      std::printf(__VA_ARGS__);                            \
      std::printf("\n");                                   \
    }                                                      \
  } while (0)

// extra cycles of one LDS access of a wave (32-lane halves, 32 banks)
static int conflict_cycles(const std::vector<long> &addr) {
  int extra = 0;
  for (int h = 0; h < 64; h += 32) {
    std::map<long, std::set<long>> bank;
    for (int l = h; l < h + 32; ++l)
      if (addr[l] >= 0) bank[addr[l] % 32].insert(addr[l]);
    size_t deg = 1;
    for (auto &kv : bank) deg = std::max(deg, kv.second.size());
    extra += (int)deg - 1;
  }
  return extra;
}

// column-phase lattice reads of one shot over every wave: x-derivative rows
// and node columns, with lattice strides (lx, lxy)
template <class T>
static long lattice_read_conflicts(int lx, int lxy) {
  constexpr int N = T::n - 1, n = T::n;
  long total = 0;
  for (int ch = 0; ch < T::NCH; ++ch)
    for (int w = 0; w < T::BT / 64; ++w) {
      std::vector<long> base(64, -1), cb(64, -1);
      for (int l = 0; l < 64; ++l) {
        const int t = 64 * w + l, ec = t / T::NCOL, col = t % T::NCOL, e = ch * T::EC + ec;
        if (ec >= T::EC || e >= T::E) continue;
        const int ci = col % n, cj = T::NL > 1 ? col / n : 0;
        const int ea = e % (T::LX / N), eb = (e / (T::LX / N)) % (T::LY / N), ez = e / ((T::LX / N) * (T::LY / N));
        const long eb0 = (long)ea * N + (long)lx * (eb * N) + (T::NL > 1 ? (long)lxy * (ez * N) : 0);
        base[l] = eb0 + (T::NL > 1 ? (long)lx * cj : 0);
        cb[l] = base[l] + ci;
      }
      const long cs = T::NL > 1 ? lxy : lx;
      for (int k = 0; k < n; ++k) {
        std::vector<long> a(64, -1);
        for (int l = 0; l < 64; ++l) a[l] = cb[l] < 0 ? -1 : cb[l] + k * cs;
        total += conflict_cycles(a);
        for (int m = 0; m < n; ++m) {
          for (int l = 0; l < 64; ++l) a[l] = base[l] < 0 ? -1 : base[l] + k * cs + m;
          total += conflict_cycles(a);
        }
      }
    }
  return total;
}

template <int D, int N, int P, int X, int Y, int Z, bool HO>
static void check() {
  using T = TS<D, N, P, X, Y, Z, HO>;
  using U = TSL<D, N, P, X, Y, Z, HO, false>;
  using Pd = TSL<D, N, P, X, Y, Z, HO, true>;
  const char *ph = P == SEM_ACOUSTIC ? "acoustic" : "elastic";
  std::printf("dim %d order %d %s %dx%dx%d %s: %d LDS words (%d per CU), padded %d%s\n", D, N, ph, X, Y, Z,
              HO ? "high-occupancy" : "normal", T::LDS, 40960 / T::LDS, (int)T::PAD,
              T::PAD ? "" : (Pd::PAD ? " (padded layout would cost occupancy)" : ""));
  // occupancy rule and fit
  CHECK(T::LDS <= 40960, "layout exceeds 160 KiB");
  CHECK((bool)T::PAD == (Pd::PAD && 40960 / Pd::LDS >= 40960 / U::LDS), "padding choice");
  // regions: [start, end) in words
  struct R { const char *name; long a, b; };
  const long su_end = (long)T::C * T::NTLP;
  std::vector<R> regs = {{"lattice", 0, su_end},
                         {"words", T::O_SW, T::O_SW + T::NTLP},
                         {"fluxes", T::O_SF, T::O_SF + (long)T::EC * T::SFE},
                         {"element data", T::O_SE, T::O_SE + (long)T::E * T::NED},
                         {"sources", T::O_SS, T::O_SS + 4L * kTileMaxSrc + 1},
                         {"derivative table", T::O_SD, T::O_SD + (long)T::n * T::n},
                         {"slot table", T::O_SL, T::O_SL + (T::SLOT_LDS ? (T::NTL + 1) / 2 : 0)}};
  // results: with aliasing, chunks but the last before the fluxes, the last
  // chunk's over its own fluxes
  const long rk_end = T::O_SK + (long)T::E * T::SKE;
  if (T::ALIAS) {
    CHECK(T::O_SK + (long)(T::NCH - 1) * T::EC * T::SKE == T::O_SF, "aliased results end where fluxes start");
    CHECK(rk_end <= T::O_SF + (long)T::EC * T::SFE, "last chunk's results inside its fluxes");
  } else {
    regs.push_back({"results", T::O_SK, rk_end});
  }
  CHECK(T::SFE >= T::C * T::NL * T::ND && T::SKE >= T::C * T::ND, "element strides");
  for (size_t i = 0; i < regs.size(); ++i) {
    CHECK(regs[i].a >= 0 && regs[i].b <= T::LDS, "%s outside the layout", regs[i].name);
    for (size_t j = i + 1; j < regs.size(); ++j)
      CHECK(regs[i].b <= regs[j].a || regs[j].b <= regs[i].a, "%s overlaps %s", regs[i].name, regs[j].name);
  }
  // slot mapping
  std::set<int> seen;
  for (int l = 0; l < T::NTL; ++l) {
    const int z = l / T::LXY, y = (l % T::LXY) / T::LX, x = l % T::LX;
    const int p = T::lds_slot(l);
    CHECK(p == x + T::LXP * y + T::LXYP * z, "slot %d -> %d", l, p);
    CHECK(p >= 0 && p < T::NTLP, "slot %d outside the lattice", l);
    seen.insert(p);
  }
  CHECK((int)seen.size() == T::NTL, "slot mapping not one to one");
  if (T::PAD) {
    // column stores: bank of thread t is (t + const) mod 32
    for (int t = 0; t < T::EC * T::NCOL; ++t) {
      const int ec = t / T::NCOL, col = t % T::NCOL;
      CHECK((ec * T::SFE + col - t) % 32 == 0, "flux store bank of thread %d", t);
      CHECK((ec * T::SKE + col - t) % 64 == 0, "result store bank of thread %d", t);
    }
    if (T::LXP != T::LX) {
      const long pad = lattice_read_conflicts<T>(T::LXP, T::LXYP);
      const long unp = lattice_read_conflicts<T>(T::LX, T::LXY);
      std::printf("  lattice-read conflict cycles per shot and workgroup: %ld padded, %ld unpadded\n", pad, unp);
      CHECK(pad < unp, "padding does not reduce lattice-read conflicts");
    }
  }
}

int main() {
#define CHK(D, N, X, Y, Z)                  \
  check<D, N, 0, X, Y, Z, false>();         \
  check<D, N, 0, X, Y, Z, true>();          \
  check<D, N, 1, X, Y, Z, false>();
  SEM_TILE_KERNELS(CHK)
#undef CHK
  std::printf("%s\n", failures ? "FAILED" : "all layouts ok");
  return failures ? 1 : 0;
}
