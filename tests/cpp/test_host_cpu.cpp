/*
 * C++ host interface, parts that need no GPU: Environment, Basis, lattices,
 * sublattices and Tools::RandomDisorder, with the literal values of the
 * reference's unit tests (tests/unit_tests/environment, sublattice,
 * tools_random) and the sector sizes and row splits of src/basis.cpp.
 */
#include "check.hpp"
#include "lsquantumed.hpp"

using namespace lsqed;

static char help[] = "C++ host interface tests (no GPU)\n";
static int _argc;
static char** _argv;

/* tests/unit_tests/environment/test_env4.cpp, with one process. */
TEST(EnvironmentTest, ConstrDestr8) {
  Environment testEnv(_argc, _argv, 8, help);
  EXPECT_EQ(1, testEnv.mpi_size);
  EXPECT_EQ(0, testEnv.mpi_rank);
  EXPECT_EQ(8, testEnv.nspins);
  EXPECT_EQ(false, testEnv.disorder_flg);
  Environment later(_argc, _argv, help);
  EXPECT_EQ(0, later.nspins);
}

/* src/environment.cpp:31-35: odd counts are rejected first, then
 * non-positive ones. */
TEST(EnvironmentTest, InvalidSpinCounts) {
  EXPECT_THROW_MSG(Environment(_argc, _argv, 7, help), std::invalid_argument,
                   "The total number of spins has to be EVEN.");
  EXPECT_THROW_MSG(Environment(_argc, _argv, -3, help), std::invalid_argument,
                   "The total number of spins has to be EVEN.");
  EXPECT_THROW_MSG(Environment(_argc, _argv, PetscInt(0), help), std::invalid_argument,
                   "The total number of spins has to be POSITIVE.");
  EXPECT_THROW_MSG(Environment(_argc, _argv, -2, help), std::invalid_argument,
                   "The total number of spins has to be POSITIVE.");
}

TEST(BasisTest, SectorSizes) {
  Environment env(_argc, _argv, 8, help);
  Basis b0{env, 0};
  EXPECT_EQ(4, b0.nspins_up);
  EXPECT_EQ(70, b0.size);
  EXPECT_EQ(70, b0.local_size);
  EXPECT_EQ(0, b0.global_start_index);
  EXPECT_EQ(4, b0.max_mag);
  Basis b1{env, 1};
  EXPECT_EQ(56, b1.size);
  Basis bm4{env, -4};
  EXPECT_EQ(0, bm4.nspins_up);
  EXPECT_EQ(1, bm4.size);
  EXPECT_THROW_MSG((Basis{env, 5}), std::invalid_argument, "The required constant magnetization has to be <= nspins/2.");
  EXPECT_THROW_MSG((Basis{env, -5}), std::invalid_argument,
                   "The required constant magnetization has to be <= nspins/2.");
  Environment env36(_argc, _argv, 36, help);
  Basis big{env36, 0};
  EXPECT_EQ(9075135300LL, big.size);
}

/* The row split of src/basis.cpp:78-80 when the environment reports two
 * ranks: 15 rows -> 8 + 7. */
TEST(BasisTest, RowSplitTwoRanks) {
  Environment env(_argc, _argv, 6, help);
  env.mpi_size = 2;
  env.mpi_rank = 0;
  Basis r0{env, 1};
  EXPECT_EQ(15, r0.size);
  EXPECT_EQ(8, r0.local_size);
  EXPECT_EQ(0, r0.global_start_index);
  env.mpi_rank = 1;
  Basis r1{env, 1};
  EXPECT_EQ(7, r1.local_size);
  EXPECT_EQ(8, r1.global_start_index);
  env.mpi_size = 2;
  EXPECT_THROW_MSG((Basis{env, 3}), std::invalid_argument,
                   "The size of the basis has to be at least equal to the number of MPI processes used.");
}

TEST(LatticeTest, Chain1D) {
  Environment env(_argc, _argv, 8, help);
  chain1D lat{env};
  EXPECT_EQ(lattice_type::chain1D, lat.get_type());
  EXPECT_EQ(1, lat.num_nn());
  EXPECT_EQ(1, lat.num_nnn());
  EXPECT_EQ(8, lat.num_spins());
  for (PetscInt i = 0; i < 8; ++i) {
    EXPECT_EQ(1, lat.num_nnXsite(i));
    EXPECT_EQ(1, lat.num_nnnXsite(i));
  }
  EXPECT_SEQ(lat.nn_i, 0, 1, 2, 3, 4, 5, 6, 7);
  EXPECT_SEQ(lat.nn_j, 1, 2, 3, 4, 5, 6, 7, 0);
  EXPECT_SEQ(lat.nnn_j, 2, 3, 4, 5, 6, 7, 0, 1);
}

TEST(LatticeTest, Square2D) {
  Environment env(_argc, _argv, 12, help);
  square2D lat{env, 3, 4};
  EXPECT_EQ(lattice_type::square2D, lat.get_type());
  EXPECT_EQ(2, lat.num_nn());
  EXPECT_EQ(24u, lat.nn_i.size());
  EXPECT_EQ(24u, lat.nnn_i.size());
  for (PetscInt i = 0; i < 12; ++i) {
    EXPECT_EQ(2, lat.num_nnXsite(i));
    EXPECT_EQ(2, lat.num_nnnXsite(i));
  }
  // every site has four distinct nearest neighbours on the periodic 3x4 torus
  std::vector<int> degree(12, 0);
  for (size_t b = 0; b < lat.nn_i.size(); ++b) {
    ++degree[size_t(lat.nn_i[b])];
    ++degree[size_t(lat.nn_j[b])];
  }
  for (int d : degree) EXPECT_EQ(4, d);
  EXPECT_THROW_MSG((square2D{env, 3, 3}), std::invalid_argument,
                   "The product between the number of spins in x direction and the numer of spins in the y direction "
                   "has to be equal to the total number of spins.");
}

TEST(LatticeTest, Honeycomb2D) {
  Environment env(_argc, _argv, 16, help);
  honeycomb2D lat{env, 4, 4};
  EXPECT_EQ(lattice_type::honeycomb2D, lat.get_type());
  EXPECT_EQ(2, lat.num_nn());
  EXPECT_EQ(3, lat.num_nnn());
  PetscInt nn = 0, nnn = 0;
  for (PetscInt i = 0; i < 16; ++i) {
    nn += lat.num_nnXsite(i);
    nnn += lat.num_nnnXsite(i);
    EXPECT_EQ(3, lat.num_nnnXsite(i));
  }
  EXPECT_EQ(24, nn);  // coordination 3: 3N/2 bonds
  EXPECT_EQ(48, nnn);
  std::vector<int> degree(16, 0);
  for (size_t b = 0; b < lat.nn_i.size(); ++b) {
    ++degree[size_t(lat.nn_i[b])];
    ++degree[size_t(lat.nn_j[b])];
  }
  for (int d : degree) EXPECT_EQ(3, d);
  Environment env12(_argc, _argv, 12, help);
  EXPECT_THROW_MSG((honeycomb2D{env12, 3, 4}), std::invalid_argument,
                   "The number of spins in each direction has to be even.");
}

/* tests/unit_tests/sublattice/test_sublattice1.cpp */
TEST(SublatticeTest, GetSublatsChain1D) {
  Environment env(_argc, _argv, 6, help);
  chain1D l1{env};
  AF<chain1D> sublat{l1, 3};
  EXPECT_EQ(2, sublat.get_size());
  EXPECT_SEQ(sublat.get_sl(0), 0, 2, 4);
  EXPECT_SEQ(sublat.get_sl(1), 1, 3, 5);
}

/* tests/unit_tests/sublattice/test_sublattice2.cpp */
TEST(SublatticeTest, GetSublatsSquare2D) {
  Environment env(_argc, _argv, 16, help);
  square2D l1{env, 4, 4};
  AF<square2D> sublat1{l1, 5};
  EXPECT_EQ(2, sublat1.get_size());
  EXPECT_SEQ(sublat1.get_sl(0), 0, 2, 5, 7, 8, 10, 13, 15);
  EXPECT_SEQ(sublat1.get_sl(1), 1, 3, 4, 6, 9, 11, 12, 14);
  Striped<square2D> sublat11{l1, 5};
  EXPECT_EQ(4, sublat11.get_size());
  EXPECT_SEQ(sublat11.get_sl(0), 0, 2, 4, 6, 8, 10, 12, 14);
  EXPECT_SEQ(sublat11.get_sl(1), 1, 3, 5, 7, 9, 11, 13, 15);
  EXPECT_SEQ(sublat11.get_sl(2), 0, 1, 2, 3, 8, 9, 10, 11);
  EXPECT_SEQ(sublat11.get_sl(3), 4, 5, 6, 7, 12, 13, 14, 15);

  square2D l2{env, 2, 8};
  AF<square2D> sublat2{l2, 46};
  EXPECT_EQ(2, sublat2.get_size());
  EXPECT_SEQ(sublat2.get_sl(0), 0, 2, 4, 6, 9, 11, 13, 15);
  EXPECT_SEQ(sublat2.get_sl(1), 1, 3, 5, 7, 8, 10, 12, 14);
  Striped<square2D> sublat22{l2, 54};
  EXPECT_EQ(4, sublat22.get_size());
  EXPECT_SEQ(sublat22.get_sl(0), 0, 2, 4, 6, 8, 10, 12, 14);
  EXPECT_SEQ(sublat22.get_sl(1), 1, 3, 5, 7, 9, 11, 13, 15);
  EXPECT_SEQ(sublat22.get_sl(2), 0, 1, 2, 3, 4, 5, 6, 7);
  EXPECT_SEQ(sublat22.get_sl(3), 8, 9, 10, 11, 12, 13, 14, 15);
}

/* tests/unit_tests/sublattice/test_sublattice3.cpp */
TEST(SublatticeTest, GetSublatsHoneycomb2D) {
  Environment env(_argc, _argv, 16, help);
  honeycomb2D l1{env, 4, 4};
  AF<honeycomb2D> sublat1{l1, 2};
  EXPECT_EQ(2, sublat1.get_size());
  EXPECT_SEQ(sublat1.get_sl(0), 0, 2, 5, 7, 8, 10, 13, 15);
  EXPECT_SEQ(sublat1.get_sl(1), 1, 3, 4, 6, 9, 11, 12, 14);
  Striped<honeycomb2D> sublat11{l1, 6};
  EXPECT_EQ(6, sublat11.get_size());
  EXPECT_SEQ(sublat11.get_sl(0), 0, 2, 4, 6, 8, 10, 12, 14);
  EXPECT_SEQ(sublat11.get_sl(1), 1, 3, 5, 7, 9, 11, 13, 15);
  honeycomb2D l2{env, 2, 8};
  AF<honeycomb2D> sublat2{l2, 4};
  EXPECT_SEQ(sublat2.get_sl(0), 0, 2, 4, 6, 9, 11, 13, 15);
  EXPECT_SEQ(sublat2.get_sl(1), 1, 3, 5, 7, 8, 10, 12, 14);
}

/* tests/unit_tests/tools_random/test_tools_random1.cpp */
TEST(ToolsRandom, Constructor) {
  Tools::RandomDisorder rd{6, PETSC_TRUE, -0.005, 0.005, 10};
  EXPECT_EQ(6, rd.nspins);
  EXPECT_NEAR(-0.005, rd.min, 1e-13);
  EXPECT_NEAR(0.005, rd.max, 1e-13);
  EXPECT_EQ(10, rd.rep);
  EXPECT_SEQ(rd.srand_seq, 134, 136, 451, 21, 351, 912, 471, 74, 570, 635);
  rd.hi_init(2);
  EXPECT_NEAR(0.00033296339931027097, rd.hi[0], 1e-13);
  EXPECT_NEAR(-0.0024819526282803856, rd.hi[1], 1e-13);
  EXPECT_NEAR(-0.0043726977344249567, rd.hi[2], 1e-13);
  EXPECT_NEAR(-0.0043282870609637177, rd.hi[3], 1e-13);
  EXPECT_NEAR(-0.0018033511217890774, rd.hi[4], 1e-13);
  EXPECT_NEAR(-0.0002136219343576908, rd.hi[5], 1e-13);
}

/* Without reprod the seeds come from std::random_device: still in [0, 1000]. */
TEST(ToolsRandom, NotReproducible) {
  Tools::RandomDisorder rd{4, PETSC_FALSE, -1.0, 1.0, 64};
  for (PetscInt s : rd.srand_seq) {
    EXPECT_GE(s, 0);
    EXPECT_LE(s, 1000);
  }
  rd.hi_init(63);
  for (double h : rd.hi) {
    EXPECT_GE(h, -1.0);
    EXPECT_LE(h, 1.0);
  }
}

int main(int argc, char** argv) {
  _argc = argc;
  _argv = argv;
  return check::run(argc, argv);
}
