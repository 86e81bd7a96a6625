/*
 * C++ host interface: the reference's dump tests (tests/unit_tests/
 * build_diag, build_off_diag, build_disorder; compare.sh diffs the MatView or
 * VecView(MatGetDiagonal) text of each against its .ref file).  Each
 * case builds the model through lsquantumed.hpp and must reproduce the
 * reference's dump character for character, in the rank layout its header
 * names.  build_disorder cases use h_i = i/N, as the reference's tests do.
 */
#include <unistd.h>

#include <fstream>
#include <sstream>

#include "check.hpp"
#include "lsquantumed.hpp"

using namespace lsqed;

static int _argc;
static char** _argv;
static char help[] = "C++ host interface dump tests (GPU)\n";

struct DumpCase {
  const char* ref;   // reference_data/<ref>
  const char* test;  // the reference test that writes it
  const char* lattice;
  int N, mag, lx, ly;
  double J1, D1, J2, D2;
  const char* builds;  // D = build_diag, O = build_off_diag, H = build_disorder
  bool mat;            // MatView (else VecView of the diagonal)
};

static const DumpCase kCases[] = {
    {"build_diag124.ref", "build_diag/square_2d/test_build_diag124.cpp", "square", 12, 4, 3, 4, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag125.ref", "build_diag/square_2d/test_build_diag125.cpp", "square", 12, 5, 3, 4, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag82.ref", "build_diag/square_2d/test_build_diag82.cpp", "square", 8, 2, 2, 4, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag60.ref", "build_diag/chain_1d/test_build_diag60.cpp", "chain", 6, 0, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag61.ref", "build_diag/chain_1d/test_build_diag61.cpp", "chain", 6, 1, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag62.ref", "build_diag/chain_1d/test_build_diag62.cpp", "chain", 6, 2, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag63.ref", "build_diag/chain_1d/test_build_diag63.cpp", "chain", 6, 3, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag6m1.ref", "build_diag/chain_1d/test_build_diag6m1.cpp", "chain", 6, -1, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag6m2.ref", "build_diag/chain_1d/test_build_diag6m2.cpp", "chain", 6, -2, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag6m3.ref", "build_diag/chain_1d/test_build_diag6m3.cpp", "chain", 6, -3, 0, 0, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag125h.ref", "build_diag/honeycomb_2d/test_build_diag125.cpp", "honeycomb", 12, 5, 6, 2, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_diag82h.ref", "build_diag/honeycomb_2d/test_build_diag82.cpp", "honeycomb", 8, 2, 2, 4, 1.0, 1.0, 0.0, 0.0, "D", false},
    {"build_disorder125.ref", "build_disorder/test_build_disorder125.cpp", "square", 12, 5, 3, 4, 1.0, 1.0, 1.0, 1.0, "OH", true},
    {"build_disorder60.ref", "build_disorder/test_build_disorder60.cpp", "chain", 6, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_disorder61.ref", "build_disorder/test_build_disorder61.cpp", "chain", 6, 1, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_disorder62.ref", "build_disorder/test_build_disorder62.cpp", "chain", 6, 2, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_disorder63.ref", "build_disorder/test_build_disorder63.cpp", "chain", 6, 3, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_disorder6m1.ref", "build_disorder/test_build_disorder6m1.cpp", "chain", 6, -1, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_disorder6m2.ref", "build_disorder/test_build_disorder6m2.cpp", "chain", 6, -2, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_disorder6m3.ref", "build_disorder/test_build_disorder6m3.cpp", "chain", 6, -3, 0, 0, 0.0, 0.0, 0.0, 0.0, "H", true},
    {"build_off_diag125.ref", "build_off_diag/square_2d/test_build_off_diag125.cpp", "square", 12, 5, 3, 4, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag82.ref", "build_off_diag/square_2d/test_build_off_diag82.cpp", "square", 8, 2, 2, 4, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag60.ref", "build_off_diag/chain_1d/test_build_off_diag60.cpp", "chain", 6, 0, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag61.ref", "build_off_diag/chain_1d/test_build_off_diag61.cpp", "chain", 6, 1, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag62.ref", "build_off_diag/chain_1d/test_build_off_diag62.cpp", "chain", 6, 2, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag63.ref", "build_off_diag/chain_1d/test_build_off_diag63.cpp", "chain", 6, 3, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag6m1.ref", "build_off_diag/chain_1d/test_build_off_diag6m1.cpp", "chain", 6, -1, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag6m2.ref", "build_off_diag/chain_1d/test_build_off_diag6m2.cpp", "chain", 6, -2, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag6m3.ref", "build_off_diag/chain_1d/test_build_off_diag6m3.cpp", "chain", 6, -3, 0, 0, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag125h.ref", "build_off_diag/honeycomb_2d/test_build_off_diag125.cpp", "honeycomb", 12, 5, 6, 2, 1.0, 1.0, 0.0, 0.0, "O", true},
    {"build_off_diag82h.ref", "build_off_diag/honeycomb_2d/test_build_off_diag82.cpp", "honeycomb", 8, 2, 2, 4, 1.0, 1.0, 0.0, 0.0, "O", true},
};

static std::string fixture_dir() {
  char buf[4096];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (n <= 0) return "";
  std::string exe(buf, size_t(n));
  return exe.substr(0, exe.rfind('/')) + "/../../golden/reference_data/";
}

static std::string read_file(const std::string& fn) {
  std::ifstream f(fn);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

/* Text written by one view call, captured through a temporary file. */
template <typename F>
static std::string capture(int nprocs, F&& view) {
  FILE* fp = std::tmpfile();
  if (!fp) return "";
  PetscViewer v = nullptr;
  PetscViewerASCIICreate(fp, nprocs, &v);
  const PetscErrorCode e = view(v);
  PetscViewerDestroy(&v);
  std::fflush(fp);
  std::rewind(fp);
  std::string out;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), fp)) > 0) out.append(buf, n);
  std::fclose(fp);
  return e ? std::string("error ") + std::to_string(e) : out;
}

template <typename L>
static std::string dump(const DumpCase& c, L& lat, Environment& env, int nprocs) {
  Basis b{env, c.mag};
  Hamiltonian<L> h{env, b, lat, c.J1, c.D1, c.J2, c.D2};
  const std::string builds = c.builds;
  if (builds.find('D') != std::string::npos && h.build_diag(env)) return "build_diag failed";
  if (builds.find('O') != std::string::npos && h.build_off_diag(env)) return "build_off_diag failed";
  if (builds.find('H') != std::string::npos) {
    std::vector<double> hi(size_t(c.N));
    for (int i = 0; i < c.N; ++i) hi[size_t(i)] = double(i) / double(c.N);
    if (h.build_disorder(env, hi)) return "build_disorder failed";
  }
  if (c.mat) return capture(nprocs, [&](PetscViewer v) { return MatView(h.hamilt, v); });
  Vec d = nullptr;
  if (MatCreateVecs(h.hamilt, &d, NULL)) return "MatCreateVecs failed";
  std::string out = "MatGetDiagonal failed";
  if (!MatGetDiagonal(h.hamilt, d)) out = capture(nprocs, [&](PetscViewer v) { return VecView(d, v); });
  VecDestroy(&d);
  return out;
}

TEST(CompareDumps, All31) {
  const std::string dir = fixture_dir();
  int done = 0;
  for (const DumpCase& c : kCases) {
    std::printf("  %s (%s)\n", c.ref, c.test);
    std::fflush(stdout);
    const std::string ref = read_file(dir + c.ref);
    ASSERT_TRUE(!ref.empty());
    // rank count from the header "... Object: P MPI processes"
    const size_t colon = ref.find(':');
    const int nprocs = std::atoi(ref.c_str() + colon + 1);
    ASSERT_GT(nprocs, 0);
    Environment env{_argc, _argv, c.N, help};
    std::string got;
    const std::string lat = c.lattice;
    if (lat == "chain") {
      chain1D l{env};
      got = dump(c, l, env, nprocs);
    } else if (lat == "square") {
      square2D l{env, c.lx, c.ly};
      got = dump(c, l, env, nprocs);
    } else {
      honeycomb2D l{env, c.lx, c.ly};
      got = dump(c, l, env, nprocs);
    }
    if (got != ref) check::fail(__FILE__, __LINE__, std::string(c.ref) + " differs:\n" + got.substr(0, 400));
    ++done;
  }
  EXPECT_EQ(31, done);
}

/* One process, chain N=4, J1 = D1 = 1, Sz = 0 (states 0011, 0101, 0110, 1001,
 * 1010, 1100 in colex order): a sequential vector; rows with as many
 * parallel as antiparallel bonds store no diagonal, and the Neel row 0101
 * flips to all four other states with one antiparallel pair moved. */
TEST(ViewFormat, OneProcess) {
  Environment env{_argc, _argv, 4, help};
  Basis b{env, 0};
  chain1D lat{env};
  Hamiltonian<chain1D> h{env, b, lat, 1.0, 1.0, 0.0, 0.0};
  ASSERT_EQ(0, h.build_diag(env));
  ASSERT_EQ(0, h.build_off_diag(env));
  Vec d = nullptr;
  ASSERT_EQ(0, MatCreateVecs(h.hamilt, &d, NULL));
  ASSERT_EQ(0, MatGetDiagonal(h.hamilt, d));
  const std::string v = capture(1, [&](PetscViewer vw) { return VecView(d, vw); });
  EXPECT_EQ(std::string("Vec Object: 1 MPI processes\n  type: seq\n0.\n-1.\n0.\n0.\n-1.\n0.\n"), v);
  const std::string m = capture(1, [&](PetscViewer vw) { return MatView(h.hamilt, vw); });
  EXPECT_EQ(0u, m.find("Mat Object: 1 MPI processes\n  type: mpiaij\nrow 0: (1, 0.5)  (4, 0.5) \nrow 1: (0, 0.5)  (1, -1.)  (2, 0.5)  (3, 0.5)  (5, 0.5) \n"));
  VecDestroy(&d);
}

int main(int argc, char** argv) {
  _argc = argc;
  _argv = argv;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    std::printf("no GPU\n");
    return 2;
  }
  return check::run(argc, argv);
}
