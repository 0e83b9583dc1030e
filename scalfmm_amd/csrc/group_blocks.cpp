// group_blocks.cpp -- evaluation of the particles held in the particle blocks
// of the reference's group tree.
//
// The byte layout of a block (GroupTree/Core/FGroupOfParticles.hpp:19-130, in
// 64-bit builds with FSize = MortonIndex = long long, FGlobal.hpp:34, 47):
//   BlockHeader (alignas 8, 64 bytes): startingIndex, endingIndex (int64),
//     numberOfLeavesInBlock (int32 + 4 bytes of padding),
//     nbParticlesAllocatedInGroup (int64), offsetPosition,
//     positionsLeadingDim, attributeLeadingDim (size_t), nbParticlesInGroup (int64)
//   LeafHeader[numberOfLeavesInBlock] (24 bytes each): mindex, nbParticles, offSet
//   at offsetPosition: x, y, z, then the NbSymbAttributes physical values, each
//     an array of nbParticlesAllocatedInGroup doubles (leading dimension
//     positionsLeadingDim bytes); leaf i owns entries offSet .. offSet+nbParticles-1
// The results live in a separate buffer (getRawAttributesBuffer,
// FGroupOfParticles.hpp:222-231): NbAttributesPerParticle arrays of
// nbParticlesAllocatedInGroup doubles, attributeLeadingDim bytes apart.  For
// FP2PGroupParticleContainer<double, 1, 1, NVALS> (FP2PGroupParticleContainer.hpp:8-57)
// those are potential[v], forceX[v], forceY[v], forceZ[v] at v, NVALS + v,
// 2 NVALS + v, 3 NVALS + v; the physical values are the NVALS symbolic ones.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "hfmm_internal.hpp"

namespace {

struct BlockHeader {
    int64_t starting_index;
    int64_t ending_index;
    int32_t number_of_leaves;
    int32_t padding;
    int64_t nb_allocated;
    uint64_t offset_position;
    uint64_t positions_leading_dim;
    uint64_t attribute_leading_dim;
    int64_t nb_particles;
};
static_assert(sizeof(BlockHeader) == 64, "BlockHeader of FGroupOfParticles");

struct LeafHeader {
    int64_t mindex;
    int64_t nb_particles;
    uint64_t offset;
};
static_assert(sizeof(LeafHeader) == 24, "LeafHeader of FGroupOfParticles");

struct Block {
    BlockHeader h;
    std::vector<LeafHeader> leaves;
    const unsigned char* raw;
    int64_t first;  // index of the block's first particle in the gathered arrays
};

std::string where(int32_t b) { return "particle block " + std::to_string(b) + ": "; }

// Header checks: the assertions of the reading constructor
// (FGroupOfParticles.hpp:99-124) and of newLeaf (:300-312), plus the sorted
// leaf order that getLeafIndex's binary search relies on (:277-292).
int read_block(hfmm::Context* c, int32_t b, const unsigned char* raw, int64_t max_key, Block& out) {
    std::memcpy(&out.h, raw, sizeof(BlockHeader));
    const BlockHeader& h = out.h;
    if (h.number_of_leaves < 0 || h.nb_allocated < 0 || h.nb_particles < 0 || h.nb_particles > h.nb_allocated)
        return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "bad counts in the block header");
    const uint64_t ld = sizeof(double) * static_cast<uint64_t>(h.nb_allocated);
    if (h.positions_leading_dim != ld || h.attribute_leading_dim != ld)
        return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "leading dimensions differ from 8 * nbParticlesAllocatedInGroup");
    const uint64_t headers = sizeof(BlockHeader) + sizeof(LeafHeader) * static_cast<uint64_t>(h.number_of_leaves);
    if (h.offset_position < headers || h.offset_position % sizeof(double))
        return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "positions overlap the leaf headers");
    out.leaves.resize(static_cast<size_t>(h.number_of_leaves));
    if (h.number_of_leaves)
        std::memcpy(out.leaves.data(), raw + sizeof(BlockHeader), sizeof(LeafHeader) * out.leaves.size());
    int64_t total = 0, prev = -1;
    for (size_t i = 0; i < out.leaves.size(); ++i) {
        const LeafHeader& l = out.leaves[i];
        if (l.mindex <= prev || l.mindex < h.starting_index || l.mindex >= h.ending_index || l.mindex > max_key)
            return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "leaf " + std::to_string(i) +
                                                   ": Morton index out of order or outside the block's interval");
        if (l.nb_particles < 0 || l.offset > static_cast<uint64_t>(h.nb_allocated) ||
            static_cast<uint64_t>(l.nb_particles) > static_cast<uint64_t>(h.nb_allocated) - l.offset)
            return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "leaf " + std::to_string(i) + ": particles outside the block");
        prev = l.mindex;
        total += l.nb_particles;
    }
    if (total != h.nb_particles)
        return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "leaf particle counts do not add up to nbParticlesInGroup");
    out.raw = raw;
    return HFMM_OK;
}

}  // namespace

extern "C" {

int hfmm_evaluate_particle_blocks(hfmm_ctx* h, int32_t nblocks, const void* const* blocks, double* const* attributes,
                                  int32_t nvals, uint32_t ops, int32_t upper, int32_t lower) {
    if (!h) return HFMM_ERR_ARG;
    hfmm::Context* c = hfmm::context_of(h);
    if (nblocks < 0 || (nblocks > 0 && (!blocks || !attributes)))
        return hfmm::fail(c, HFMM_ERR_ARG, "particle blocks: bad arguments");
    if (nvals < 1 || nvals > HFMM_MAX_RHS) return hfmm::fail(c, HFMM_ERR_ARG, "particle blocks: bad nvals");
    if (ops & ~HFMM_NEAR_AND_FAR) return hfmm::fail(c, HFMM_ERR_ARG, "unknown operation bits");
    if (upper < 2 || upper > lower || lower > c->P.tree_height)
        return hfmm::fail(c, HFMM_ERR_ARG, "working levels must satisfy 2 <= upper <= lower <= height");
    const int64_t max_key = (int64_t(1) << (3 * (c->P.tree_height - 1))) - 1;
    std::vector<Block> bl(static_cast<size_t>(nblocks));
    int64_t n = 0;
    for (int32_t b = 0; b < nblocks; ++b) {
        if (!blocks[b] || !attributes[b]) return hfmm::fail(c, HFMM_ERR_ARG, where(b) + "null buffer");
        HFMM_TRY(read_block(c, b, static_cast<const unsigned char*>(blocks[b]), max_key, bl[b]));
        bl[b].first = n;
        n += bl[b].h.nb_particles;
    }
    if (n == 0) return HFMM_OK;  // nothing to evaluate, nothing to add

    // gather the particles in block and leaf order (already sorted by leaf)
    std::vector<double> x(n), y(n), z(n), q(static_cast<size_t>(n) * nvals);
    std::vector<int64_t> key(n);
    for (const Block& k : bl) {
        const double* pos = reinterpret_cast<const double*>(k.raw + k.h.offset_position);
        const size_t na = static_cast<size_t>(k.h.nb_allocated);
        int64_t i = k.first;
        for (const LeafHeader& l : k.leaves)
            for (int64_t p = 0; p < l.nb_particles; ++p, ++i) {
                const size_t e = l.offset + p;
                // unaligned-safe reads: the caller's buffer need not be 8-byte aligned
                std::memcpy(&x[i], pos + e, sizeof(double));
                std::memcpy(&y[i], pos + na + e, sizeof(double));
                std::memcpy(&z[i], pos + 2 * na + e, sizeof(double));
                for (int v = 0; v < nvals; ++v)
                    std::memcpy(&q[static_cast<size_t>(v) * n + i], pos + (3 + v) * na + e, sizeof(double));
                key[i] = l.mindex;
            }
    }
    // With kept cells, blocks holding the same positions as at the last call
    // keep the tree and its cells (the reference's tree is the same object
    // between its executions); only the physical values are refreshed.
    std::vector<double> xyz(static_cast<size_t>(3) * n);
    std::memcpy(xyz.data(), x.data(), sizeof(double) * n);
    std::memcpy(xyz.data() + n, y.data(), sizeof(double) * n);
    std::memcpy(xyz.data() + 2 * n, z.data(), sizeof(double) * n);
    const bool reuse = c->keep_cells && c->tree_built && c->nsets == 1 && c->nrhs == nvals && c->n == n &&
                       c->block_xyz == xyz;
    if (reuse) {
        HFMM_TRY(hfmm_set_charges(h, q.data()));
    } else {
        c->block_xyz.clear();
        HFMM_TRY(hfmm_set_particle_sets(h, 1, nullptr, nvals, n, x.data(), y.data(), z.data(), q.data()));
        HFMM_TRY(hfmm_build_tree(h));
        // every particle must lie in the leaf its block files it under
        std::vector<int64_t> skeys(n), perm(n);
        HFMM_TRY(hfmm_get_sorted_keys(h, skeys.data(), perm.data()));
        for (int64_t s = 0; s < n; ++s)
            if (skeys[s] != key[perm[s]]) {
                const int64_t i = perm[s];
                int32_t b = 0;
                while (b + 1 < nblocks && bl[b + 1].first <= i) ++b;
                c->tree_built = false;  // the tree does not describe the caller's blocks
                return hfmm::fail(c, HFMM_ERR_ARG,
                                  where(b) + "particle " + std::to_string(i - bl[b].first) + " lies in leaf " +
                                      std::to_string(skeys[s]) + ", not in the leaf " + std::to_string(key[i]) +
                                      " it is stored under");
            }
        c->block_xyz.swap(xyz);
    }
    HFMM_TRY(hfmm_evaluate_levels(h, ops, upper, lower));
    std::vector<double> res(static_cast<size_t>(4) * n * nvals);
    const size_t rn = static_cast<size_t>(n) * nvals;
    HFMM_TRY(hfmm_get_results(h, res.data(), res.data() + rn, res.data() + 2 * rn, res.data() + 3 * rn));
    // add (P2P and L2P accumulate into the containers, FChebSymKernel.hpp:432-486)
    for (const Block& k : bl) {
        double* attr = attributes[&k - bl.data()];
        const size_t na = static_cast<size_t>(k.h.nb_allocated);
        int64_t i = k.first;
        for (const LeafHeader& l : k.leaves)
            for (int64_t p = 0; p < l.nb_particles; ++p, ++i) {
                const size_t e = l.offset + p;
                for (int a = 0; a < 4; ++a)
                    for (int v = 0; v < nvals; ++v) {
                        double* dst = attr + (static_cast<size_t>(a) * nvals + v) * na + e;
                        double acc;
                        std::memcpy(&acc, dst, sizeof(double));
                        acc += res[a * rn + static_cast<size_t>(v) * n + i];
                        std::memcpy(dst, &acc, sizeof(double));
                    }
            }
    }
    return HFMM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cell blocks of one level (GroupTree/Core/FGroupOfCells.hpp:18-242 with the
// cells of GroupTree/Chebyshev/FChebCellPOD.hpp:11-112, alignment 8,
// StarPUUtils/FStarPUDefaultAlign.hpp:4-6):
//   raw buffer:  BlockHeader {startingIndex, endingIndex (int64),
//                numberOfCellsInBlock (int32 + 4 bytes of padding)},
//                the cells' Morton indexes (int64 each), then one
//                FBasicCellPOD {mortonIndex (int64), coordinates[3], level
//                (int32)} per cell (GroupTree/Core/FBasicCellPOD.hpp:9-13)
//   multipoles:  per cell FChebCellPODPole: nvals x 2 p^3 doubles, RHS v at
//                v 2 p^3 (FChebCellPOD.hpp:85-100); the symmetric kernel uses
//                the first p^3 values of each RHS slot, FChebKernel keeps its
//                compressed expansion behind them (FChebKernel.hpp:103, 148)
//   locals:      the same with FChebCellPODLocal
namespace {

struct CellBlockHeader {
    int64_t starting_index;
    int64_t ending_index;
    int32_t number_of_cells;
    int32_t padding;
};
static_assert(sizeof(CellBlockHeader) == 24, "BlockHeader of FGroupOfCells");

struct CellBlock {
    CellBlockHeader h;
    std::vector<int64_t> keys;
    std::vector<int64_t> cell;  // the context's cell index (in its level) of each block cell
};

std::string cwhere(int32_t b) { return "cell block " + std::to_string(b) + ": "; }

// Headers, index order and interval (the assertions of FGroupOfCells'
// constructors and newCell, FGroupOfCells.hpp:88-112, 130-168, 347-354), and
// every cell found among the context's cells of the level.
int read_cell_blocks(hfmm::Context* c, int32_t level, int32_t nblocks, const void* const* blocks,
                     std::vector<CellBlock>& out) {
    const int64_t nc = c->ncell[level];
    std::vector<int64_t> keys(static_cast<size_t>(nc));
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    HFMM_HIP(c, hipMemcpy(keys.data(), c->d_cellkey.as<uint64_t>() + c->lvl_off[level], sizeof(int64_t) * nc,
                          hipMemcpyDeviceToHost));
    out.resize(static_cast<size_t>(nblocks));
    for (int32_t b = 0; b < nblocks; ++b) {
        if (!blocks[b]) return hfmm::fail(c, HFMM_ERR_ARG, cwhere(b) + "null buffer");
        const unsigned char* raw = static_cast<const unsigned char*>(blocks[b]);
        CellBlock& k = out[b];
        std::memcpy(&k.h, raw, sizeof(CellBlockHeader));
        if (k.h.number_of_cells < 0 || k.h.ending_index < k.h.starting_index ||
            k.h.ending_index - k.h.starting_index < k.h.number_of_cells)
            return hfmm::fail(c, HFMM_ERR_ARG, cwhere(b) + "bad header");
        k.keys.resize(static_cast<size_t>(k.h.number_of_cells));
        if (k.h.number_of_cells)
            std::memcpy(k.keys.data(), raw + sizeof(CellBlockHeader), sizeof(int64_t) * k.keys.size());
        k.cell.resize(k.keys.size());
        int64_t prev = -1;
        for (size_t i = 0; i < k.keys.size(); ++i) {
            const int64_t m = k.keys[i];
            if (m <= prev || m < k.h.starting_index || m >= k.h.ending_index)
                return hfmm::fail(c, HFMM_ERR_ARG, cwhere(b) + "cell " + std::to_string(i) +
                                                       ": Morton index out of order or outside the block's interval");
            prev = m;
            const auto it = std::lower_bound(keys.begin(), keys.end(), m);
            if (it == keys.end() || *it != m)
                return hfmm::fail(c, HFMM_ERR_ARG, cwhere(b) + "cell " + std::to_string(m) + " of level " +
                                                       std::to_string(level) + " is not a cell of the context's tree");
            k.cell[i] = it - keys.begin();
        }
    }
    return HFMM_OK;
}

// Device arrays of charge vector r: nodal M and L [cells][LD], and for the
// global-SVD scheme the compressed M~ and L~ [cells][k] (one vector only).
struct ExpArrays {
    double* M;
    double* L;
    double* Mt;
    double* Lt;
};

ExpArrays exp_arrays(hfmm::Context* c, int r) {
    const int64_t words = c->total_cells * static_cast<int64_t>(c->LD);
    ExpArrays a{c->d_M.as<double>(), c->d_L.as<double>(), nullptr, nullptr};
    if (c->keep_cells && c->nrhs > 1) {
        a.M = c->d_kept.as<double>() + 2 * words * r;
        a.L = a.M + words;
    }
    if (c->m2l_scheme != HFMM_M2L_SYMMETRIC && c->glob.k > 0) {
        a.Mt = c->d_Mt.as<double>();
        a.Lt = c->d_Lt.as<double>();
    }
    return a;
}

int check_cell_call(hfmm::Context* c, int32_t level, int32_t nblocks, const void* const* blocks, int32_t nvals) {
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "cell blocks before build_tree");
    if (c->nsets != 1) return hfmm::fail(c, HFMM_ERR_ARG, "cell blocks: the context holds several particle sets");
    if (level < 2 || level >= c->P.tree_height)
        return hfmm::fail(c, HFMM_ERR_ARG, "cell blocks: level must lie in 2 .. height-1 (the kernel's levels)");
    if (nblocks < 0 || (nblocks > 0 && !blocks)) return hfmm::fail(c, HFMM_ERR_ARG, "cell blocks: bad arguments");
    if (nvals != c->nrhs) return hfmm::fail(c, HFMM_ERR_ARG, "cell blocks: nvals differs from the charge vectors");
    if (nvals > 1 && !c->keep_cells)
        return hfmm::fail(c, HFMM_ERR_STATE, "cell blocks of several charge vectors need kept cells");
    return HFMM_OK;
}

}  // namespace

extern "C" {

int hfmm_get_cell_blocks(hfmm_ctx* h, int32_t level, int32_t nblocks, const void* const* blocks,
                         double* const* multipoles, double* const* locals, int32_t nvals) {
    if (!h) return HFMM_ERR_ARG;
    hfmm::Context* c = hfmm::context_of(h);
    HFMM_TRY(check_cell_call(c, level, nblocks, blocks, nvals));
    if (!c->evaluated && !(c->keep_cells && !c->cells_dirty))
        return hfmm::fail(c, HFMM_ERR_STATE, "get_cell_blocks before an evaluation");
    std::vector<CellBlock> bl;
    HFMM_TRY(read_cell_blocks(c, level, nblocks, blocks, bl));
    const int nn = c->ops.nn, LD = c->LD;
    const int64_t nc = c->ncell[level], first = c->lvl_off[level];
    const int k = (c->m2l_scheme != HFMM_M2L_SYMMETRIC) ? c->glob.k : 0;
    const bool compressed = k > 0 && k <= nn && c->m2l_scheme != HFMM_M2L_UNIFORM;
    std::vector<double> M(static_cast<size_t>(nc) * LD), L(M.size()), Mt, Lt;
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < nvals; ++r) {
        const ExpArrays a = exp_arrays(c, r);
        HFMM_HIP(c, hipMemcpy(M.data(), a.M + first * LD, sizeof(double) * M.size(), hipMemcpyDeviceToHost));
        HFMM_HIP(c, hipMemcpy(L.data(), a.L + first * LD, sizeof(double) * L.size(), hipMemcpyDeviceToHost));
        if (compressed && a.Mt) {
            Mt.resize(static_cast<size_t>(nc) * k);
            Lt.resize(Mt.size());
            HFMM_HIP(c, hipMemcpy(Mt.data(), a.Mt + first * k, sizeof(double) * Mt.size(), hipMemcpyDeviceToHost));
            HFMM_HIP(c, hipMemcpy(Lt.data(), a.Lt + first * k, sizeof(double) * Lt.size(), hipMemcpyDeviceToHost));
        }
        for (int32_t b = 0; b < nblocks; ++b)
            for (size_t i = 0; i < bl[b].cell.size(); ++i) {
                const size_t src = static_cast<size_t>(bl[b].cell[i]);
                const size_t dst = (i * nvals + r) * 2 * static_cast<size_t>(nn);
                if (multipoles && multipoles[b]) {
                    std::memcpy(multipoles[b] + dst, M.data() + src * LD, sizeof(double) * nn);
                    if (!Mt.empty()) std::memcpy(multipoles[b] + dst + nn, Mt.data() + src * k, sizeof(double) * k);
                }
                if (locals && locals[b]) {
                    std::memcpy(locals[b] + dst, L.data() + src * LD, sizeof(double) * nn);
                    if (!Lt.empty()) std::memcpy(locals[b] + dst + nn, Lt.data() + src * k, sizeof(double) * k);
                }
            }
    }
    return HFMM_OK;
}

int hfmm_set_cell_blocks(hfmm_ctx* h, int32_t level, int32_t nblocks, const void* const* blocks,
                         const double* const* multipoles, const double* const* locals, int32_t nvals) {
    if (!h) return HFMM_ERR_ARG;
    hfmm::Context* c = hfmm::context_of(h);
    HFMM_TRY(check_cell_call(c, level, nblocks, blocks, nvals));
    if (!c->keep_cells) return hfmm::fail(c, HFMM_ERR_STATE, "set_cell_blocks needs kept cells (HFMM_CELLS_KEPT)");
    if (c->m2l_scheme != HFMM_M2L_SYMMETRIC)
        return hfmm::fail(c, HFMM_ERR_ARG, "set_cell_blocks: the symmetric M2L scheme only");
    std::vector<CellBlock> bl;
    HFMM_TRY(read_cell_blocks(c, level, nblocks, blocks, bl));
    HFMM_HIP(c, hipSetDevice(c->device));
    HFMM_TRY(hfmm::prepare_kept_cells(c));  // expansion buffers, initial state of a new tree
    const int nn = c->ops.nn, LD = c->LD;
    const int64_t nc = c->ncell[level], first = c->lvl_off[level];
    std::vector<double> M(static_cast<size_t>(nc) * LD), L(M.size());
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < nvals; ++r) {
        const ExpArrays a = exp_arrays(c, r);
        HFMM_HIP(c, hipMemcpy(M.data(), a.M + first * LD, sizeof(double) * M.size(), hipMemcpyDeviceToHost));
        HFMM_HIP(c, hipMemcpy(L.data(), a.L + first * LD, sizeof(double) * L.size(), hipMemcpyDeviceToHost));
        for (int32_t b = 0; b < nblocks; ++b)
            for (size_t i = 0; i < bl[b].cell.size(); ++i) {
                const size_t dst = static_cast<size_t>(bl[b].cell[i]) * LD;
                const size_t src = (i * nvals + r) * 2 * static_cast<size_t>(nn);
                if (multipoles && multipoles[b]) std::memcpy(M.data() + dst, multipoles[b] + src, sizeof(double) * nn);
                if (locals && locals[b]) std::memcpy(L.data() + dst, locals[b] + src, sizeof(double) * nn);
            }
        HFMM_HIP(c, hipMemcpy(a.M + first * LD, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice));
        HFMM_HIP(c, hipMemcpy(a.L + first * LD, L.data(), sizeof(double) * L.size(), hipMemcpyHostToDevice));
    }
    return HFMM_OK;
}

}  // extern "C"
