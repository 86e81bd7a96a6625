// Host-code sanitizer check (CPU only; `make -C tests/cpp asan`): drives the
// library's pure-host paths under AddressSanitizer / LeakSanitizer -- the
// hipRTC source generator of the specialised H.v for every lattice kind and
// sector shape (single-model and batched), the bond lists and their argument
// errors (src/lattice.cpp:67-68, :153-154), the disorder RNG
// (src/tools/random_disorder.cpp:33-59), sector sizes, and the clean-up of
// lsqed_ham_create when no device is present, and the low-bond records of
// the LOWTAB kernels against brute force.  Exits non-zero on a wrong
// result; the sanitizer aborts on a memory error or leak.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lsqed_internal.hpp"


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

// Low-bond records of a LOWTAB source against brute force: patterns of each
// popcount class listed in numeric order, a partner's index looked up in that
// list (no rank formula), the antiparallel counts recounted bond by bond.
static int check_records(int lat, int N, int lx, int ly, bool batch, double J2, int mode) {
  lsqed_ham h;
  h.N = N;
  h.nup = N / 2;
  h.J[0] = 1.0;
  h.J[1] = J2;
  h.Dz[0] = 1.0;
  h.Dz[1] = 0.3;
  EXPECT(bonds(lat, N, lx, ly, h));
  std::vector<uint32_t> rec;
  const std::string s = lsqed::jit_source(&h, batch, false, &rec);
  EXPECT(s.find(mode == 1 ? "#define LOWTAB 1\n" : "#define LOWTAB 2\n") != std::string::npos);
  const int L = lsqed::split_low_bits(&h, batch);
  std::vector<unsigned> m;
  std::vector<int> k, flipping;
  for (int kind = 0; kind < 2; ++kind)
    for (size_t b = 0; b < h.bi[kind].size(); ++b) {
      const int a = std::min(h.bi[kind][b], h.bj[kind][b]), c = std::max(h.bi[kind][b], h.bj[kind][b]);
      if (a == c || c >= L) continue;
      m.push_back((1u << a) | (1u << c));
      k.push_back(kind);
      // mode 1 keeps every flipping bond in the record, mode 2 only those with
      // one site in each byte of the pattern
      if (h.J[kind] != 0.0 && (mode == 1 || (a < 8 && c >= 8))) flipping.push_back((int)m.size() - 1);
    }
  const int nf = (int)flipping.size(), lword = (nf + 1) / 2, recq = (lword + 4) / 4;
  std::vector<std::vector<unsigned>> cls(L + 1);
  std::vector<int> pos(1u << L);
  for (unsigned v = 0; v < (1u << L); ++v) {
    const int c = __builtin_popcount(v);
    pos[v] = (int)cls[c].size();
    cls[c].push_back(v);
  }
  EXPECT(rec.size() == ((size_t)1 << L) * 4 * recq);
  if (rec.size() != ((size_t)1 << L) * 4 * recq) return 0;
  int rmax = 0;
  for (const auto& c : cls) rmax = std::max<int>(rmax, (int)c.size());
  size_t g = 0, bad = 0;
  for (int c = 0; c <= L; ++c)
    for (unsigned v : cls[c]) {
      const uint32_t* w = &rec[g * 4 * recq];
      bad += (w[lword] & 0xffffu) != v;
      int cc[2] = {0, 0};
      for (size_t b = 0; b < m.size(); ++b) cc[k[b]] += __builtin_popcount(v & m[b]) == 1;
      bad += (int)((w[lword] >> 16) & 0xffu) != cc[0] || (int)(w[lword] >> 24) != cc[1];
      for (int t = 0; t < nf; ++t) {
        const unsigned mb = m[flipping[t]];
        const int want = __builtin_popcount(v & mb) == 1 ? pos[v ^ mb] : rmax;
        const int got = (int)((t & 1) ? (w[t / 2] >> 16) : (w[t / 2] & 0xffffu));
        bad += got != want;
      }
      ++g;
    }
  EXPECT(bad == 0);
  return 1;
}

// Site renumbering plans: a bijection of the sites that puts at least 4 more
// flipping bonds among the row split's low sites, none where the numbering
// is already best; the 4x8 torus reaches the 43 low bonds that its 8x4
// numbering has.
static int low_bonds(const lsqed_ham& h, const std::vector<int>& perm, int L) {
  int n = 0;
  for (int k = 0; k < 2; ++k)
    for (size_t b = 0; b < h.bi[k].size(); ++b) {
      const int i = perm.empty() ? h.bi[k][b] : perm[h.bi[k][b]], j = perm.empty() ? h.bj[k][b] : perm[h.bj[k][b]];
      n += h.J[k] != 0.0 && i != j && i < L && j < L;
    }
  return n;
}

static int check_relabel(int lat, int N, int lx, int ly, double J2, int expect_low) {
  lsqed_ham h;
  h.N = N;
  h.nup = N / 2;
  h.D = lsqed_sector_dim(N, N / 2);
  h.J[0] = 1.0;
  h.J[1] = J2;
  h.Dz[0] = 1.0;
  h.Dz[1] = J2;
  EXPECT(bonds(lat, N, lx, ly, h));
  std::vector<int> perm;
  const bool on = lsqed::relabel_plan(&h, perm);
  const int L = lsqed::split_low_bits(&h, false);
  if (expect_low == -1) {
    EXPECT(!on);
    return 0;
  }
  if (expect_low >= 0) EXPECT(on);
  if (!on) return 0;
  std::vector<int> seen(N, 0);
  for (int v : perm) {
    EXPECT(v >= 0 && v < N);
    if (v >= 0 && v < N) ++seen[v];
  }
  for (int v = 0; v < N; ++v) EXPECT(seen[v] == 1);
  const int before = low_bonds(h, {}, L), after = low_bonds(h, perm, L);
  EXPECT(after >= before + 4);
  EXPECT(after >= expect_low);
  return 1;
}

int main() {
  setenv("LSQED_RELABEL", "1", 1);
  int nrel = check_relabel(LSQED_SQUARE2D, 32, 4, 8, 0.5, 43);
  nrel += check_relabel(LSQED_SQUARE2D, 36, 4, 9, 0.5, 43);
  nrel += check_relabel(LSQED_SQUARE2D, 24, 4, 6, 0.5, 43);
  nrel += check_relabel(LSQED_SQUARE2D, 24, 2, 12, 0.5, 50);
  nrel += check_relabel(LSQED_SQUARE2D, 32, 8, 4, 0.5, -1);     // numbering already best
  nrel += check_relabel(LSQED_SQUARE2D, 36, 6, 6, 0.5, -1);
  nrel += check_relabel(LSQED_CHAIN1D, 30, 0, 0, 0.2, -1);
  const int hcr = check_relabel(LSQED_HONEYCOMB2D, 32, 4, 8, 0.4, -2);  // either way; a valid plan if any
  unsetenv("LSQED_RELABEL");
  nrel += check_relabel(LSQED_SQUARE2D, 24, 4, 6, 0.5, -1);     // below 2^24 states by default
  setenv("LSQED_RELABEL", "0", 1);
  nrel += check_relabel(LSQED_SQUARE2D, 32, 4, 8, 0.5, -1);
  unsetenv("LSQED_RELABEL");
  EXPECT(nrel == 4);
  std::printf("renumbering plans checked (honeycomb 4x8: %s)\n", hcr ? "renumbered" : "kept");
  int nrec = check_records(LSQED_CHAIN1D, 22, 0, 0, true, 0.2, 2);  // the batched kernel's default
  nrec += check_records(LSQED_CHAIN1D, 18, 0, 0, true, 0.0, 2);      // NNN bonds do not flip, still counted
  setenv("LSQED_JIT_LOWTAB", "1", 1);
  nrec += check_records(LSQED_CHAIN1D, 22, 0, 0, true, 0.2, 1);
  nrec += check_records(LSQED_SQUARE2D, 32, 4, 8, false, 0.5, 1);
  nrec += check_records(LSQED_HONEYCOMB2D, 24, 4, 6, false, 0.4, 1);
  setenv("LSQED_JIT_LOWTAB", "2", 1);
  nrec += check_records(LSQED_SQUARE2D, 32, 4, 8, false, 0.5, 2);
  unsetenv("LSQED_JIT_LOWTAB");
  EXPECT(nrec == 6);
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
