/*
 * The MATSHELL adapter (lsquantumed_amd/petsc/hamiltonian_shell.hpp) on the
 * GPU, compiled against the PETSc test double in petsc_double/ -- once with
 * complex scalars (as the reference's PETSc, README.md:18) and once with real
 * ones.  The shell's MatMult / MatMultTranspose / MatGetDiagonal equal the
 * C-ABI products bit for bit, follow the build_* calls (an empty operator
 * after MatInit, then diagonal, off-diagonal, fields), and reject host
 * vectors, wrong lengths and aliased vectors with PETSc error codes.
 *
 *   test_shell_{complex,real}                     run the cases
 *   test_shell_{complex,real} dump LAT N MAG LX LY J2
 *       print the full-build operator (J1 = D1 = 1, J2 = D2) column by column
 *       from MatMult(e_j), as "row i: (c, v)" lines, for comparison with the
 *       reference's build_hamilt*.ref dumps (tests/test_petsc_shell.py)
 */
#include <cmath>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "check.hpp"
#include "hamiltonian_shell.hpp"

using namespace lsqed_petsc;

/* A lattice with the reference's neighbour layout (include/lattice.hpp):
 * nn / nnn flat, stride num_nn() + 1, slots 1..num_nnXsite(i) of site i. */
struct RefLayoutLattice {
  PetscInt N = 0, stride[2] = {1, 1};
  std::vector<PetscInt> per_site[2];
  std::vector<PetscInt> nn, nnn;
  RefLayoutLattice(int lat, int n, int lx, int ly) : N(n) {
    for (int k = 0; k < 2; ++k) {
      int c = 0;
      lsqed_lattice_bonds(lat, n, lx, ly, k, nullptr, nullptr, &c);
      std::vector<int> bi((size_t)c), bj((size_t)c);
      lsqed_lattice_bonds(lat, n, lx, ly, k, bi.data(), bj.data(), &c);
      per_site[k].assign(size_t(n), 0);
      for (int b = 0; b < c; ++b) ++per_site[k][size_t(bi[size_t(b)])];
      PetscInt mx = 1;
      for (PetscInt v : per_site[k]) mx = std::max(mx, v);
      stride[k] = mx;
      std::vector<PetscInt>& arr = k ? nnn : nn;
      arr.assign(size_t((mx + 1) * n), -1);
      std::vector<PetscInt> fill(size_t(n), 0);
      for (int b = 0; b < c; ++b) {
        const int i = bi[size_t(b)];
        arr[size_t((mx + 1) * i)] = i;
        arr[size_t((mx + 1) * i + ++fill[size_t(i)])] = bj[size_t(b)];
      }
    }
  }
  PetscInt num_nn() { return stride[0]; }
  PetscInt num_nnn() { return stride[1]; }
  PetscInt num_nnXsite(PetscInt i) { return per_site[0][size_t(i)]; }
  PetscInt num_nnnXsite(PetscInt i) { return per_site[1][size_t(i)]; }
};

static std::vector<PetscScalar> to_host(Vec v) {
  std::vector<PetscScalar> h(size_t(v->n));
  (void)hipMemcpy(h.data(), v->a, sizeof(PetscScalar) * size_t(v->n), hipMemcpyDeviceToHost);
  return h;
}
static void from_host(Vec v, const std::vector<PetscScalar>& h) {
  (void)hipMemcpy(v->a, h.data(), sizeof(PetscScalar) * size_t(v->n), hipMemcpyHostToDevice);
}
static double re(PetscScalar s) { return std::real(s); }
static double im(PetscScalar s) { return std::imag(s); }

static Mat make_shell(int lat, int N, int mag, int lx, int ly, double J2) {
  RefLayoutLattice L(lat, N, lx, ly);
  Mat A = nullptr;
  if (HamiltonianShellFromLattice(PETSC_COMM_WORLD, N, N / 2 + mag, L, 1.0, 1.0, J2, J2, &A)) return nullptr;
  return A;
}

static int dump(int argc, char** argv) {
  if (argc < 8) return 2;
  const std::string lat = argv[2];
  const int kind = lat == "chain" ? LSQED_CHAIN1D : lat == "square" ? LSQED_SQUARE2D : LSQED_HONEYCOMB2D;
  Mat A = make_shell(kind, atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), atof(argv[7]));
  if (!A || HamiltonianShellBuild(A, LSQED_BUILD_DIAG) || HamiltonianShellBuild(A, LSQED_BUILD_OFFDIAG)) return 3;
  Vec e = nullptr, col = nullptr;
  MatCreateVecs(A, &e, &col);
  const PetscInt D = e->n;
  std::vector<std::vector<PetscScalar>> cols;
  for (PetscInt j = 0; j < D; ++j) {
    std::vector<PetscScalar> h(size_t(D), PetscScalar(0.0));
    h[size_t(j)] = 1.0;
    from_host(e, h);
    if (MatMult(A, e, col)) return 4;
    cols.push_back(to_host(col));
  }
  for (PetscInt i = 0; i < D; ++i) {
    std::printf("row %lld:", (long long)i);
    for (PetscInt j = 0; j < D; ++j) {
      const PetscScalar v = cols[size_t(j)][size_t(i)];
      if (im(v) != 0.0) return 5;
      if (re(v) != 0.0) std::printf(" (%lld, %.17g)", (long long)j, re(v));
    }
    std::printf("\n");
  }
  VecDestroy(&e);
  VecDestroy(&col);
  MatDestroy(&A);
  return 0;
}

/* Shell products against the C-ABI on a 4x4 J1-J2 lattice with fields. */
TEST(Shell, MatchesLibrary) {
  Mat A = make_shell(LSQED_SQUARE2D, 16, 0, 4, 4, 0.5);
  ASSERT_TRUE(A != nullptr);
  lsqed_ham* h = nullptr;
  ASSERT_EQ(0, HamiltonianShellHandle(A, &h));
  Vec x = nullptr, y = nullptr, z = nullptr;
  ASSERT_EQ(0, MatCreateVecs(A, &x, &y));
  ASSERT_EQ(0, MatCreateVecs(A, &z, nullptr));
  EXPECT_EQ(std::string(VECHIP), x->type);
  const PetscInt D = x->n;
  EXPECT_EQ(PetscInt(12870), D);
  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<PetscScalar> xh((size_t)D);
  for (auto& v : xh) {
#if defined(PETSC_USE_COMPLEX)
    v = PetscScalar(U(rng), U(rng));
#else
    v = U(rng);
#endif
  }
  from_host(x, xh);
  // after MatInit nothing is built: the operator is zero
  ASSERT_EQ(0, MatMult(A, x, y));
  bool zero = true;
  for (auto v : to_host(y)) zero = zero && v == PetscScalar(0.0);
  EXPECT_TRUE(zero);
  // build_diag, build_off_diag, build_disorder
  ASSERT_EQ(0, HamiltonianShellBuild(A, LSQED_BUILD_DIAG));
  ASSERT_EQ(0, HamiltonianShellBuild(A, LSQED_BUILD_OFFDIAG));
  std::vector<PetscReal> hi(16);
  for (int i = 0; i < 16; ++i) hi[size_t(i)] = 0.1 * i - 0.7;
  ASSERT_EQ(0, HamiltonianShellSetFields(A, hi.data()));
  ASSERT_EQ(0, MatMult(A, x, y));
#if defined(PETSC_USE_COMPLEX)
  ASSERT_EQ(0, lsqed_hv_c128(h, (const double*)x->a, (double*)z->a, nullptr));
#else
  ASSERT_EQ(0, lsqed_hv(h, x->a, z->a, nullptr));
#endif
  const auto yh = to_host(y), zh = to_host(z);
  size_t same = 0;
  for (size_t i = 0; i < yh.size(); ++i) same += (yh[i] == zh[i]);
  EXPECT_EQ(yh.size(), same);
  // MatMultTranspose: H is symmetric
  ASSERT_EQ(0, MatMultTranspose(A, x, z));
  const auto th = to_host(z);
  same = 0;
  for (size_t i = 0; i < yh.size(); ++i) same += (yh[i] == th[i]);
  EXPECT_EQ(yh.size(), same);
  // the diagonal
  ASSERT_EQ(0, MatGetDiagonal(A, z));
  double* dd = nullptr;
  ASSERT_EQ(0, (int)hipMalloc((void**)&dd, sizeof(double) * size_t(D)));
  ASSERT_EQ(0, lsqed_diag(h, dd, nullptr));
  std::vector<double> d((size_t)D);
  (void)hipMemcpy(d.data(), dd, sizeof(double) * size_t(D), hipMemcpyDeviceToHost);
  (void)hipFree(dd);
  const auto dh = to_host(z);
  same = 0;
  for (size_t i = 0; i < dh.size(); ++i) same += (re(dh[i]) == d[i] && im(dh[i]) == 0.0);
  EXPECT_EQ(dh.size(), same);
  // errors: a host vector, a wrong length, x aliased with y
  Vec hostv = nullptr, shortv = nullptr;
  ASSERT_EQ(0, VecCreateType(D, VECSEQ, &hostv));
  ASSERT_EQ(0, VecCreateType(D - 1, VECHIP, &shortv));
  EXPECT_EQ(PETSC_ERR_ARG_WRONG, MatMult(A, hostv, y));
  EXPECT_EQ(PETSC_ERR_ARG_SIZ, MatMult(A, x, shortv));
  EXPECT_EQ(PETSC_ERR_ARG_IDN, MatMult(A, x, x));
  EXPECT_EQ(PETSC_ERR_ARG_WRONG, MatGetDiagonal(A, hostv));
  VecDestroy(&hostv);
  VecDestroy(&shortv);
  VecDestroy(&x);
  VecDestroy(&y);
  VecDestroy(&z);
  EXPECT_EQ(0, MatDestroy(&A));
}

/* The 4x8 sector (D = 601,080,390) through the shell: one product equals the
 * C-ABI's, spot-checked on rows; no host copy happens in MatMult (the
 * vectors never leave the device). */
TEST(Shell, Square4x8) {
  Mat A = make_shell(LSQED_SQUARE2D, 32, 0, 4, 8, 0.5);
  ASSERT_TRUE(A != nullptr);
  ASSERT_EQ(0, HamiltonianShellBuild(A, LSQED_BUILD_DIAG | LSQED_BUILD_OFFDIAG));
  lsqed_ham* h = nullptr;
  ASSERT_EQ(0, HamiltonianShellHandle(A, &h));
  Vec x = nullptr, y = nullptr, z = nullptr;
  ASSERT_EQ(0, MatCreateVecs(A, &x, &y));
  ASSERT_EQ(0, MatCreateVecs(A, &z, nullptr));
  const PetscInt D = x->n;
  EXPECT_EQ(PetscInt(601080390), D);
  // x_r = 1 / (1 + r mod 97) (+ i r mod 89 / 89): deterministic, host-free to check
  std::vector<PetscScalar> chunk(size_t(1) << 24);
  for (PetscInt s = 0; s < D; s += PetscInt(chunk.size())) {
    const PetscInt n = std::min<PetscInt>(PetscInt(chunk.size()), D - s);
    for (PetscInt r = 0; r < n; ++r) {
      const PetscInt g = s + r;
#if defined(PETSC_USE_COMPLEX)
      chunk[size_t(r)] = PetscScalar(1.0 / double(1 + g % 97), double(g % 89) / 89.0);
#else
      chunk[size_t(r)] = 1.0 / double(1 + g % 97);
#endif
    }
    (void)hipMemcpy(x->a + s, chunk.data(), sizeof(PetscScalar) * size_t(n), hipMemcpyHostToDevice);
  }
  ASSERT_EQ(0, MatMult(A, x, y));
#if defined(PETSC_USE_COMPLEX)
  ASSERT_EQ(0, lsqed_hv_c128(h, (const double*)x->a, (double*)z->a, nullptr));
#else
  ASSERT_EQ(0, lsqed_hv(h, x->a, z->a, nullptr));
#endif
  size_t diff = 0;
  for (PetscInt row0 : {PetscInt(0), D / 3, D - 100000}) {
    std::vector<PetscScalar> a(100000), b(100000);
    (void)hipMemcpy(a.data(), y->a + row0, sizeof(PetscScalar) * a.size(), hipMemcpyDeviceToHost);
    (void)hipMemcpy(b.data(), z->a + row0, sizeof(PetscScalar) * b.size(), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < a.size(); ++i) diff += (a[i] != b[i]);
  }
  EXPECT_EQ(size_t(0), diff);
  VecDestroy(&x);
  VecDestroy(&y);
  VecDestroy(&z);
  EXPECT_EQ(0, MatDestroy(&A));
}

int main(int argc, char** argv) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    std::printf("no GPU\n");
    return 2;
  }
  if (argc > 1 && std::string(argv[1]) == "dump") return dump(argc, argv);
  return check::run(argc, argv);
}
