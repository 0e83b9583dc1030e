// cheb_setup.hpp -- host-side set-up of the Chebyshev 1/r operators (FP64):
// roots, polynomial tables, child/parent interpolators, the 343 symmetry
// permutations and the 16 compressed M2L operators.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace hfmm {

constexpr int kMinOrder = 2;
constexpr int kMaxOrder = 14;

struct ChebOperators {
    int p = 0;        // interpolation order
    int nn = 0;       // p^3 nodes
    double eps = 0;   // compression accuracy
    std::array<double, kMaxOrder> roots{};
    std::vector<double> Troots;        // [o][j] = T_o(root_j), o = 0..p-1 (row 0 = 1)
    std::vector<double> child;         // [child][axis][n][m] = S(x_child_m, root_n)
    std::array<uint32_t, 343> pidx{};  // canonical operator index of every relative position
    std::vector<uint16_t> pvec;        // [343][nn] node permutation of every relative position
    std::array<int, 343> rank{};       // rank of the operator stored at a canonical index (0 = none)
    std::array<std::vector<double>, 343> K;  // [U | V], column-major nn x rank each, unweighted
};

// Chebyshev roots of order p as the reference tabulates them
// (Src/Kernels/Chebyshev/FChebRoots.hpp:90-297).
bool cheb_roots(int p, double* out);

// Builds the roots, interpolators and symmetry tables (not the compressed M2L
// operators: compress_operators_device); returns an HFMM_* status.
int build_cheb_operators(int order, double eps, ChebOperators& ops, std::string& err);

// The uniform (Lagrange) interpolation of FUnifKernel: the equispaced roots of
// FUnifRootsCore and the child/parent tables built with FUnifRoots::L; returns
// an HFMM_* status (orders 2..8).
double unif_lagrange(int p, int n, double x);
int make_uniform_tables(ChebOperators& op);

// The 189 far-field slots of a cell of parity e (= Morton child index 0..7):
// slot j <-> (parent-neighbour offset, cousin) in the reference enumeration
// order of FTreeCoordinate::getInteractionNeighbors (FTreeCoordinate.hpp:219-264).
struct FarSlot {
    int8_t pdx, pdy, pdz;   // parent-neighbour offset (-1..1)
    int8_t cousin;          // 0..7
    int16_t pos;            // relative position 0..342
};
void far_slots(int parity, FarSlot out[189]);

}  // namespace hfmm
