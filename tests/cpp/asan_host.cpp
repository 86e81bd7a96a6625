// Host-code sanitizer check (CPU only; `make -C tests/cpp asan`): drives the
// library's pure-host paths under AddressSanitizer / LeakSanitizer -- the
// hipRTC source generator of the specialised H.v for every lattice kind and
// sector shape (single-model and batched), the bond lists and their argument
// errors (src/lattice.cpp:67-68, :153-154), the disorder RNG
// (src/tools/random_disorder.cpp:33-59), sector sizes, and the clean-up of
// lsqed_ham_create when no device is present.  Exits non-zero on a wrong
// result; the sanitizer aborts on a memory error or leak.
#include <cstdio>
#include <string>
#include <vector>

#include "lsqed_internal.hpp"

namespace lsqed {
std::string jit_source(const lsqed_ham* h, bool batch, bool cplx);
}

static int failures = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      std::printf("FAILED %s (line %d)\n", #c, __LINE__);           \
      ++failures;                                                   \
    }                                                               \
  } while (0)

static bool bonds(int lat, int N, int lx, int ly, lsqed_ham& h) {
  for (int k = 0; k < 2; ++k) {
    int c = 0;
    if (lsqed_lattice_bonds(lat, N, lx, ly, k, nullptr, nullptr, &c) != LSQED_OK) return false;
    h.bi[k].resize(c);
    h.bj[k].resize(c);
    if (lsqed_lattice_bonds(lat, N, lx, ly, k, h.bi[k].data(), h.bj[k].data(), &c) != LSQED_OK) return false;
  }
  return true;
}

static int generate(int lat, int N, int lx, int ly) {
  int n = 0;
  for (int nup : {0, 1, 2, N / 2, N - 1, N}) {
    lsqed_ham h;
    h.N = N;
    h.nup = nup;
    h.J[0] = 1.0;
    h.J[1] = 0.4;
    h.Dz[0] = 1.0;
    h.Dz[1] = 0.3;
    EXPECT(bonds(lat, N, lx, ly, h));
    if (!lsqed::split_plan_supported(&h)) continue;
    const std::string s = lsqed::jit_source(&h, false, false);
    EXPECT(s.find("hv_jit_nodot") != std::string::npos);
    const std::string c = lsqed::jit_source(&h, false, true);
    EXPECT(c.find("#define CPLX 1") != std::string::npos && c.find("hv_jit_nodot") != std::string::npos);
    ++n;
    if (N <= 32) {
      const std::string b = lsqed::jit_source(&h, true, false);
      EXPECT(b.find("hv_bjit_nodot") != std::string::npos);
      ++n;
    }
  }
  return n;
}

int main() {
  int nsrc = 0;
  for (int N = 6; N <= 40; N += 2) nsrc += generate(LSQED_CHAIN1D, N, 0, 0);
  const int sq[][2] = {{2, 4}, {3, 4}, {4, 4}, {4, 5}, {4, 6}, {2, 12}, {4, 7}, {4, 8}, {8, 4}, {6, 6}, {4, 10}, {6, 1}};
  for (const auto& d : sq) nsrc += generate(LSQED_SQUARE2D, d[0] * d[1], d[0], d[1]);
  const int hc[][2] = {{2, 4}, {6, 2}, {4, 4}, {4, 6}, {4, 8}, {6, 6}};
  for (const auto& d : hc) nsrc += generate(LSQED_HONEYCOMB2D, d[0] * d[1], d[0], d[1]);
  EXPECT(nsrc == 390);

  int c = 0;
  EXPECT(lsqed_lattice_bonds(LSQED_SQUARE2D, 16, 3, 4, 0, nullptr, nullptr, &c) == LSQED_ERR_ARG);
  EXPECT(lsqed_lattice_bonds(LSQED_HONEYCOMB2D, 15, 3, 5, 0, nullptr, nullptr, &c) == LSQED_ERR_ARG);

  int64_t seeds[10];
  double f[40];
  EXPECT(lsqed_rd_seeds(10, seeds) == LSQED_OK);
  EXPECT(lsqed_rd_fields(seeds[3], 40, -2.0, 2.0, f) == LSQED_OK);
  for (double v : f) EXPECT(v >= -2.0 && v < 2.0);

  EXPECT(lsqed_sector_dim(36, 18) == 9075135300LL);
  EXPECT(lsqed_sector_dim(62, 31) == 465428353255261088LL);
  for (int N = 1; N <= 62; ++N)
    for (int k = 0; k <= N; ++k) EXPECT(lsqed_sector_dim(N, k) > 0);

  lsqed_ham h;
  h.N = 16;
  EXPECT(bonds(LSQED_SQUARE2D, 16, 4, 4, h));
  const lsqed_params p{16, 8, (int)h.bi[0].size(), (int)h.bi[1].size(), h.bi[0].data(), h.bj[0].data(),
                       h.bi[1].data(), h.bj[1].data(), 1.0, 1.0, 0.5, 0.5};
  lsqed_ham* out = nullptr;
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
    EXPECT(lsqed_ham_create(&p, &out) != LSQED_OK);
    EXPECT(out == nullptr);
  }
  std::printf("%d specialised sources generated; %s\n", nsrc, failures ? "FAILED" : "ok");
  return failures ? 1 : 0;
}
