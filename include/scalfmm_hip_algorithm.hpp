/*
 * scalfmm_hip_algorithm.hpp -- the engine of scalfmm_hip.h behind the
 * reference's C++ call shapes (header only; link libscalfmm_hip.so).
 *
 *   AlgorithmShape            the execute() overloads and working levels of
 *                             FAbstractAlgorithm (Src/Core/FCoreCommon.hpp:36-94),
 *                             for builds without the reference's headers; inside
 *                             the reference pass ::FAbstractAlgorithm as Base
 *   FHipChebSymKernel<...>    the kernel object of the Chebyshev 1/r path: the
 *                             FChebSymKernel constructors (FChebSymKernel.hpp:
 *                             110-140) and template parameters, owning an engine
 *                             context of order ORDER
 *   FGroupHipAlgorithm<...>   FGroupSeqAlgorithm / FGroupTaskAlgorithm over an
 *                             FGroupTree (GroupTree/Core/FGroupSeqAlgorithm.hpp:
 *                             26-83): the tree's particle blocks are evaluated in
 *                             place and its cell blocks hold the expansions
 *   hip_execute(...)          FFmmAlgorithm(Thread)::execute over a classic
 *                             octree (Core/FFmmAlgorithm.hpp, FOctree.hpp) with
 *                             FP2PParticleContainerIndexed leaves
 *
 * The reference calls its kernels one cell or leaf at a time through
 * FAbstractKernels (Components/FAbstractKernels.hpp:36-208); the engine
 * evaluates whole levels per launch, so the drop-in unit is the algorithm:
 * these classes take the reference's trees as they are and run every pass of
 * an execute() on the GPU.  Errors of the engine become std::runtime_error
 * (the reference throws the same type from its operator set-up,
 * FChebSymM2LHandler.hpp:156-160, and aborts on bad trees through FAssertLF).
 *
 * Requirements on the tree types (the reference's names):
 *   group tree:  getHeight(), getBoxWidth(), getBoxCenter() (getX/getY/getZ),
 *                getNbParticleGroup(), getParticleGroup(i) -> getRawBuffer(),
 *                getRawAttributesBuffer(); getNbCellGroupAtLevel(level),
 *                getCellGroup(level, i) -> getRawBuffer(),
 *                getRawMultipoleBuffer(), getRawLocalBuffer() (FChebCellPOD
 *                layout, GroupTree/Chebyshev/FChebCellPOD.hpp:11-100)
 *   classic octree: getHeight(), getBoxWidth(), getBoxCenter(), forEachLeaf(fn)
 *                with leaf->getTargets() -> getNbParticles(), getPositions(),
 *                getPhysicalValues(), getPotentials(), getForcesX/Y/Z()
 */
#ifndef SCALFMM_HIP_ALGORITHM_HPP
#define SCALFMM_HIP_ALGORITHM_HPP

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "scalfmm_hip.h"

namespace scalfmm_hip {

// Operation bits: the values of FFmmOperations (Core/FCoreCommon.hpp:17-29).
enum Operations : unsigned {
    OpP2P = HFMM_P2P,
    OpP2M = HFMM_P2M,
    OpM2M = HFMM_M2M,
    OpM2L = HFMM_M2L,
    OpL2L = HFMM_L2L,
    OpL2P = HFMM_L2P,
    OpNearField = HFMM_NEAR_FIELD,
    OpFarField = HFMM_FAR_FIELD,
    OpNearAndFarFields = HFMM_NEAR_AND_FAR
};

inline void check(hfmm_ctx* ctx, int status, const char* what) {
    if (status != HFMM_OK)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? hfmm_last_error(ctx) : "no context"));
}

// execute(), execute(ops), execute(upper, lower), execute(ops, upper, lower)
// with FAbstractAlgorithm's level bookkeeping: the whole tree is [2, height),
// validateLevels() requires 2 <= upper <= lower.
class AlgorithmShape {
protected:
    int upperWorkingLevel = 2;
    int lowerWorkingLevel = 0;
    int nbLevelsInTree = -1;

    void setNbLevelsInTree(const int levels) {
        nbLevelsInTree = levels;
        lowerWorkingLevel = levels;
    }
    void validateLevels() const {
        if (!(upperWorkingLevel <= lowerWorkingLevel && 2 <= upperWorkingLevel))
            throw std::runtime_error("working levels must satisfy 2 <= upper <= lower");
    }
    virtual void executeCore(const unsigned operationsToProceed) = 0;

public:
    virtual ~AlgorithmShape() {}
    void execute(const unsigned ops = OpNearAndFarFields) {
        upperWorkingLevel = 2;
        lowerWorkingLevel = nbLevelsInTree;
        validateLevels();
        executeCore(ops);
    }
    void execute(const int upper, const int lower) { execute(OpNearAndFarFields, upper, lower); }
    void execute(const unsigned ops, const int upper, const int lower) {
        upperWorkingLevel = upper;
        lowerWorkingLevel = lower;
        validateLevels();
        executeCore(ops);
    }
};

// The kernel object: FChebSymKernel<FReal, CellClass, ContainerClass,
// MatrixKernelClass, ORDER, NVALS>(treeHeight, boxWidth, boxCenter,
// matrixKernel[, epsilon]), epsilon 10^-ORDER by default.  The matrix kernel
// is the 1/r kernel (FInterpMatrixKernelR); the engine evaluates it itself,
// so the pointer is kept only as the reference's kernel keeps it (not owned).
// precision: HFMM_PRECISION_FP64 (the reference's FReal = double) or a mixed
// form.
template <class FReal, class CellClass, class ContainerClass, class MatrixKernelClass, int ORDER, int NVALS = 1>
class FHipChebSymKernel {
    hfmm_ctx* ctx_ = nullptr;
    const MatrixKernelClass* matrix_kernel_;
    int height_;

public:
    static_assert(ORDER >= 5 && ORDER <= 8, "the engine's orders are 5 to 8");
    static_assert(NVALS >= 1 && NVALS <= HFMM_MAX_RHS, "NVALS charge vectors per particle");

    template <class PointClass>
    FHipChebSymKernel(const int treeHeight, const FReal boxWidth, const PointClass& boxCenter,
                      const MatrixKernelClass* const matrixKernel, const FReal epsilon = -1,
                      const int precision = HFMM_PRECISION_FP64, const int device = 0)
        : matrix_kernel_(matrixKernel), height_(treeHeight) {
        hfmm_params p = {};
        p.tree_height = treeHeight;
        p.order = ORDER;
        p.box_width = static_cast<double>(boxWidth);
        p.box_center[0] = static_cast<double>(boxCenter.getX());
        p.box_center[1] = static_cast<double>(boxCenter.getY());
        p.box_center[2] = static_cast<double>(boxCenter.getZ());
        p.epsilon = epsilon > 0 ? static_cast<double>(epsilon) : 0.0;
        p.precision = precision;
        p.device = device;
        check(nullptr, hfmm_create(&p, &ctx_), "hfmm_create");
    }
    FHipChebSymKernel(const FHipChebSymKernel&) = delete;
    FHipChebSymKernel& operator=(const FHipChebSymKernel&) = delete;
    ~FHipChebSymKernel() {
        if (ctx_) hfmm_destroy(ctx_);
    }

    hfmm_ctx* context() const { return ctx_; }
    int treeHeight() const { return height_; }
    const MatrixKernelClass* matrixKernel() const { return matrix_kernel_; }
    static constexpr int order() { return ORDER; }
    static constexpr int nvals() { return NVALS; }
};

// FGroupSeqAlgorithm(tree, kernels) shape over an FGroupTree.  executeCore:
//   1. the particle blocks (getRawBuffer / getRawAttributesBuffer) go to the
//      engine as they are (hfmm_evaluate_particle_blocks with no pass): it
//      reads the particles and builds or keeps its tree;
//   2. the cell blocks of levels 2..height-1 go into the engine's kept cells
//      (hfmm_set_cell_blocks): passes add into what the tree's cells hold;
//   3. the passes run over the working levels and the results are added into
//      the attribute buffers, as P2P and L2P add into the containers;
//   4. the expansions go back into the cell blocks (hfmm_get_cell_blocks).
// Cells persist between executions, as the tree's do; resetting them is the
// tree's forEachCell(resetToInitialState) on the blocks.
template <class OctreeClass, class KernelClass, class Base = AlgorithmShape>
class FGroupHipAlgorithm : public Base {
    OctreeClass* const tree_;
    KernelClass* const kernel_;

    template <class Group>
    static const void* raw(Group* g) {
        return static_cast<const void*>(g->getRawBuffer());
    }

    void cell_blocks(const int level, std::vector<const void*>& blocks, std::vector<double*>& multipoles,
                     std::vector<double*>& locals) const {
        blocks.clear();
        multipoles.clear();
        locals.clear();
        for (int i = 0; i < tree_->getNbCellGroupAtLevel(level); ++i) {
            auto* g = tree_->getCellGroup(level, i);
            blocks.push_back(raw(g));
            multipoles.push_back(reinterpret_cast<double*>(g->getRawMultipoleBuffer()));
            locals.push_back(reinterpret_cast<double*>(g->getRawLocalBuffer()));
        }
    }

protected:
    void executeCore(const unsigned ops) override {
        hfmm_ctx* ctx = kernel_->context();
        std::vector<const void*> pblocks;
        std::vector<double*> attributes;
        for (int i = 0; i < tree_->getNbParticleGroup(); ++i) {
            auto* g = tree_->getParticleGroup(i);
            pblocks.push_back(raw(g));
            attributes.push_back(reinterpret_cast<double*>(g->getRawAttributesBuffer()));
        }
        const int n = static_cast<int>(pblocks.size());
        const int nv = KernelClass::nvals();
        check(ctx, hfmm_evaluate_particle_blocks(ctx, n, pblocks.data(), attributes.data(), nv, 0, 2,
                                                 tree_->getHeight()),
              "particle blocks");
        std::vector<const void*> cblocks;
        std::vector<double*> mult, loc;
        for (int level = 2; level < tree_->getHeight(); ++level) {
            cell_blocks(level, cblocks, mult, loc);
            std::vector<const double*> cm(mult.begin(), mult.end()), cl(loc.begin(), loc.end());
            check(ctx, hfmm_set_cell_blocks(ctx, level, static_cast<int32_t>(cblocks.size()), cblocks.data(),
                                            cm.data(), cl.data(), nv),
                  "set_cell_blocks");
        }
        check(ctx, hfmm_evaluate_particle_blocks(ctx, n, pblocks.data(), attributes.data(), nv, ops,
                                                 Base::upperWorkingLevel, Base::lowerWorkingLevel),
              "evaluate");
        for (int level = 2; level < tree_->getHeight(); ++level) {
            cell_blocks(level, cblocks, mult, loc);
            check(ctx, hfmm_get_cell_blocks(ctx, level, static_cast<int32_t>(cblocks.size()), cblocks.data(),
                                            mult.data(), loc.data(), nv),
                  "get_cell_blocks");
        }
    }

public:
    FGroupHipAlgorithm(OctreeClass* const tree, KernelClass* const kernel) : tree_(tree), kernel_(kernel) {
        if (!tree || !kernel) throw std::runtime_error("tree and kernels cannot be null");
        if (tree->getHeight() != kernel->treeHeight()) throw std::runtime_error("tree and kernel heights differ");
        check(kernel->context(), hfmm_set_cell_mode(kernel->context(), HFMM_CELLS_KEPT), "cell mode");
        Base::setNbLevelsInTree(tree->getHeight());
    }
};

// FFmmAlgorithm(Thread)(tree, kernels).execute(ops) over a classic octree
// whose leaves hold FP2PParticleContainerIndexed targets: the leaves'
// particles in forEachLeaf order go to the engine, one evaluation, and the
// potentials and forces are added back leaf by leaf (as the kernel's P2P and
// L2P add into the containers).
template <class OctreeClass, class KernelClass>
void hip_execute(OctreeClass& tree, KernelClass& kernel, const unsigned ops = OpNearAndFarFields) {
    hfmm_ctx* ctx = kernel.context();
    std::vector<double> x, y, z, q;
    tree.forEachLeaf([&](auto* leaf) {
        auto* c = leaf->getTargets();
        for (int64_t i = 0; i < static_cast<int64_t>(c->getNbParticles()); ++i) {
            x.push_back(c->getPositions()[0][i]);
            y.push_back(c->getPositions()[1][i]);
            z.push_back(c->getPositions()[2][i]);
            q.push_back(c->getPhysicalValues()[i]);
        }
    });
    const int64_t n = static_cast<int64_t>(x.size());
    if (n == 0) return;
    std::vector<double> pot(n), fx(n), fy(n), fz(n);
    check(ctx, hfmm_set_particles(ctx, n, x.data(), y.data(), z.data(), q.data()), "set_particles");
    check(ctx, hfmm_build_tree(ctx), "build_tree");
    check(ctx, hfmm_evaluate(ctx, ops), "evaluate");
    check(ctx, hfmm_get_results(ctx, pot.data(), fx.data(), fy.data(), fz.data()), "get_results");
    int64_t k = 0;
    tree.forEachLeaf([&](auto* leaf) {
        auto* c = leaf->getTargets();
        for (int64_t i = 0; i < static_cast<int64_t>(c->getNbParticles()); ++i, ++k) {
            c->getPotentials()[i] += pot[k];
            c->getForcesX()[i] += fx[k];
            c->getForcesY()[i] += fy[k];
            c->getForcesZ()[i] += fz[k];
        }
    });
}

}  // namespace scalfmm_hip

#endif
