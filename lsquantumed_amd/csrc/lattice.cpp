// lattice.cpp -- bond lists and random fields for liblsqed.so (host C++).
//
// Bond lists reproduce the forward-neighbour conventions of src/lattice.cpp
// (chain1D :27-53, square2D :86-136, honeycomb2D :172-264) flattened in the
// order HamiltHelper::get_coup_elems visits them (site ascending, then the
// site's neighbour slots; include/hamiltonian.tmpl_func.hpp:350-387).  That
// order is the emission order of a row's column list, which
// tests/unit_tests/get_coup_elems pins.  Periodic wrap-around on short sides
// produces repeated bonds (e.g. L=2); they are kept, as the reference keeps
// them (its matrix entries and zz counts double).
#include <random>
#include <string>
#include <vector>

#include "../../include/lsqed.h"

namespace lsqed {
int fail(int code, const std::string& msg);
}

namespace {

struct Bond {
  int i, j;
};

// site index of (ix, iy) with periodic wrap, row-major in y (src/lattice.cpp:94)
inline int site(int ix, int iy, int lx, int ly) { return ((ix % lx + lx) % lx) * ly + ((iy % ly + ly) % ly); }

void chain_bonds(int N, int kind, std::vector<Bond>& out) {
  const int step = kind == 0 ? 1 : 2;
  for (int i = 0; i < N; ++i) out.push_back({i, (i + step) % N});
}

void square_bonds(int lx, int ly, int kind, std::vector<Bond>& out) {
  for (int ix = 0; ix < lx; ++ix)
    for (int iy = 0; iy < ly; ++iy) {
      const int s = site(ix, iy, lx, ly);
      if (kind == 0) {
        out.push_back({s, site(ix, iy + 1, lx, ly)});   // along y
        out.push_back({s, site(ix + 1, iy, lx, ly)});   // along x
      } else {
        out.push_back({s, site(ix + 1, iy + 1, lx, ly)});
        out.push_back({s, site(ix + 1, iy - 1, lx, ly)});
      }
    }
}

void honeycomb_bonds(int lx, int ly, int kind, std::vector<Bond>& out) {
  for (int ix = 0; ix < lx; ++ix)
    for (int iy = 0; iy < ly; ++iy) {
      const int s = site(ix, iy, lx, ly);
      if (kind == 0) {
        // every site bonds along y; sites whose parity matches the row parity
        // also bond along x (the other slot is padding, never visited)
        out.push_back({s, site(ix, iy + 1, lx, ly)});
        const bool xbond = (ix % 2 == 0) ? (s % 2 == 0) : (s % 2 != 0);
        if (xbond) out.push_back({s, site(ix + 1, iy, lx, ly)});
      } else {
        out.push_back({s, site(ix, iy + 2, lx, ly)});
        out.push_back({s, site(ix + 1, iy + 1, lx, ly)});
        out.push_back({s, site(ix + 1, iy - 1, lx, ly)});
      }
    }
}

}  // namespace

extern "C" {

int lsqed_lattice_bonds(int lattice, int nspins, int lx, int ly, int kind, int* bi, int* bj, int* count) {
  using lsqed::fail;
  if (!count) return fail(LSQED_ERR_ARG, "null count");
  if (kind != 0 && kind != 1) return fail(LSQED_ERR_ARG, "kind must be 0 (nn) or 1 (nnn)");
  if (nspins <= 0 || nspins % 2 != 0) return fail(LSQED_ERR_ARG, "nspins must be even and positive");
  std::vector<Bond> b;
  switch (lattice) {
    case LSQED_CHAIN1D:
      chain_bonds(nspins, kind, b);
      break;
    case LSQED_SQUARE2D:
      if (lx <= 0 || ly <= 0 || lx * ly != nspins)
        return fail(LSQED_ERR_ARG,
                    "The product between the number of spins in x direction and the numer of spins in the y "
                    "direction has to be equal to the total number of spins.");
      square_bonds(lx, ly, kind, b);
      break;
    case LSQED_HONEYCOMB2D:
      if (lx <= 0 || ly <= 0 || lx * ly != nspins)
        return fail(LSQED_ERR_ARG,
                    "The product between the number of spins in x direction and the numer of spins in the y "
                    "direction has to be equal to the total number of spins.");
      if (lx % 2 != 0 || ly % 2 != 0)
        return fail(LSQED_ERR_ARG, "The number of spins in each direction has to be even.");
      honeycomb_bonds(lx, ly, kind, b);
      break;
    default:
      return fail(LSQED_ERR_ARG, "unknown lattice type");
  }
  *count = (int)b.size();
  if (bi && bj)
    for (size_t k = 0; k < b.size(); ++k) {
      bi[k] = b[k].i;
      bj[k] = b[k].j;
    }
  return LSQED_OK;
}

// Tools::RandomDisorder::srand_seq_init with reprod = true
// (src/tools/random_disorder.cpp:33-49): libstdc++'s mt19937_64 seeded with 1
// and uniform_int_distribution<PetscInt>(0, 1000).
int lsqed_rd_seeds(int rep, int64_t* seeds) {
  if (rep < 0 || (rep > 0 && !seeds)) return lsqed::fail(LSQED_ERR_ARG, "bad seed buffer");
  std::uniform_int_distribution<long> dist(0, 1000);
  std::mt19937_64 gen(1);
  for (int i = 0; i < rep; ++i) seeds[i] = dist(gen);
  return LSQED_OK;
}

// Tools::RandomDisorder::hi_init (src/tools/random_disorder.cpp:53-59).
int lsqed_rd_fields(int64_t seed, int nspins, double lo, double hi, double* out) {
  if (nspins < 0 || (nspins > 0 && !out)) return lsqed::fail(LSQED_ERR_ARG, "bad field buffer");
  std::mt19937_64 gen((uint64_t)seed);
  std::uniform_real_distribution<double> dist(lo, hi);
  for (int i = 0; i < nspins; ++i) out[i] = dist(gen);
  return LSQED_OK;
}

}  // extern "C"
