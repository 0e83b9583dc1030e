/*
 * scalfmm_hip.h -- C-ABI of the MI355X Chebyshev 1/r FMM engine.
 *
 * The engine replaces the reference's Chebyshev 1/r path: the kernel class
 * FChebSymKernel<FReal,Cell,Container,FInterpMatrixKernelR<FReal>,ORDER,NVALS>
 * (Src/Kernels/Chebyshev/FChebSymKernel.hpp:45-488) driven through the plugin
 * interface FAbstractKernels (Src/Components/FAbstractKernels.hpp:23-210) over
 * the blocked octree FGroupTree (Src/GroupTree/Core/FGroupTree.hpp:178-360)
 * by FGroupTaskAlgorithm / FGroupSeqAlgorithm.  Instead of one virtual call per
 * cell, a context owns the device tree and runs whole levels per launch.
 *
 * Conventions
 *   - every entry point returns an int status (HFMM_OK = 0); no exception
 *     crosses the ABI; hfmm_last_error() gives the message of the last failure;
 *   - the context owns all device memory; the caller owns every host array;
 *   - one context per host thread; a context runs on one HIP stream of its
 *     device; calls on one context are not re-entrant;
 *   - particle inputs and results are in the caller's (input) order.
 *
 * Reference file:line citations are relative to the reference's source tree.
 */
#ifndef SCALFMM_HIP_H
#define SCALFMM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define HFMM_OK 0
#define HFMM_ERR_ARG 1         /* invalid argument (order, height, NULL, size) */
#define HFMM_ERR_OUTSIDE_BOX 2 /* a particle lies outside the box: the reference
                                  aborts in FCoordinateComputer.hpp:18 (FAssertLF) */
#define HFMM_ERR_NUMERIC 3     /* operator compression failed (the reference throws
                                  std::runtime_error, FChebSymM2LHandler.hpp:156-160) */
#define HFMM_ERR_NOMEM 4       /* host or device allocation failed */
#define HFMM_ERR_DEVICE 5      /* a HIP runtime call failed */
#define HFMM_ERR_STATE 6       /* call out of order (e.g. evaluate before build) */
#define HFMM_ERR_IO 7          /* a file could not be opened, read or written, or has a bad format
                                  (the reference's loaders exit the process) */

/* ---- operation bits: the values of FFmmOperations (Src/Core/FCoreCommon.hpp:17-29) */
#define HFMM_P2P (1u << 0)
#define HFMM_P2M (1u << 1)
#define HFMM_M2M (1u << 2)
#define HFMM_M2L (1u << 3)
#define HFMM_L2L (1u << 4)
#define HFMM_L2P (1u << 5)
#define HFMM_NEAR_FIELD HFMM_P2P
#define HFMM_FAR_FIELD (HFMM_P2M | HFMM_M2M | HFMM_M2L | HFMM_L2L | HFMM_L2P)
#define HFMM_NEAR_AND_FAR (HFMM_NEAR_FIELD | HFMM_FAR_FIELD)

/* ---- precision modes ------------------------------------------------------ */
#define HFMM_PRECISION_FP64 0 /* every kernel in FP64 (the reference's FReal=double) */
#define HFMM_PRECISION_MIXED 1 /* M2L on bfloat16 matrix cores with three-part operand
                                  splits and FP32 accumulation, P2P in FP32 with
                                  compensated sums; P2M/M2M/L2L/L2P and all results in
                                  FP64.  Relative error about 1e-7 on top of the
                                  FP64 path's; no counterpart in the reference (its
                                  FReal=float instantiation is the closest). */
#define HFMM_PRECISION_MIXED_HILO 2 /* as HFMM_PRECISION_MIXED with two-part (hi/lo)
                                  bfloat16 splits in the M2L: three partial
                                  products instead of six, relative M2L error
                                  about 1e-5; for orders whose interpolation
                                  error is far above that (order 5: ~1e-4). */

typedef struct hfmm_ctx hfmm_ctx;

/* Kernel parameters: the arguments of the FChebSymKernel constructor
 * (FChebSymKernel.hpp:110-140: treeHeight, boxWidth, boxCenter, epsilon) plus
 * the template parameter ORDER and the device to run on. */
typedef struct hfmm_params {
    int32_t tree_height;   /* 3..20 (levels 0..height-1; leaves at height-1) */
    int32_t order;         /* Chebyshev interpolation order ORDER: 5, 6, 7 or 8 */
    double box_width;      /* width of the cubic simulation box */
    double box_center[3];  /* centre of the box */
    double epsilon;        /* M2L compression accuracy; <= 0 selects 10^-order */
    int32_t precision;     /* HFMM_PRECISION_* */
    int32_t device;        /* HIP device ordinal */
    int32_t m2l_scheme;    /* HFMM_M2L_* the context starts with (0, the symmetric
                              scheme, when zero-initialised); see hfmm_set_m2l_scheme */
} hfmm_params;

/* Creates a context: builds the Chebyshev roots, child/parent interpolators,
 * the 343 symmetry permutations and the 16 compressed M2L operators
 * (replaces the FChebSymKernel constructor, FChebSymKernel.hpp:110-126 and
 * SymmetryHandler, FChebSymM2LHandler.hpp:321-344).  A context created with
 * another M2L scheme builds that scheme's operators instead (the FChebKernel,
 * FChebDenseKernel and FUnifKernel constructors) and compresses the 16
 * symmetric operators only when it switches to the symmetric scheme or
 * hfmm_get_operator / hfmm_save_operators ask for them. */
int hfmm_create(const hfmm_params* params, hfmm_ctx** out);
int hfmm_destroy(hfmm_ctx* ctx);
/* Message of the last failure on this context (never NULL). */
const char* hfmm_last_error(const hfmm_ctx* ctx);

/* Copies n particles (host SoA, one charge per particle) to the device.
 * Replaces filling the FP2PParticleContainer handed to the FGroupTree
 * constructor (Tests/GroupTree/testBlockedChebyshev.cpp:98-114). */
int hfmm_set_particles(hfmm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                       const double* q);
/* Several independent particle sets in one context, each with nrhs charge
 * vectors.  Set s holds particles set_offsets[s] .. set_offsets[s+1]-1 of the
 * arrays (set_offsets[0] = 0, set_offsets[nsets] = n, non-decreasing; NULL when
 * nsets = 1) and lives in its own copy of the context's box with the context's
 * tree height: no interaction crosses two sets.  q holds nrhs vectors of n
 * charges (RHS-major).  With nrhs > 1, hfmm_get_results and hfmm_get_field
 * write nrhs blocks of n values per output (RHS-major) and the device pointers
 * of hfmm_get_results_device lead RHS r's arrays by 4 n r doubles.  The
 * reference's counterparts are one FGroupTree per set and the NVALS template
 * parameter of FChebSymKernel (FChebSymKernel.hpp:45-488, RHS stride 2 p^3 in
 * FChebCellPOD.hpp:85-100); the sets share one device tree whose keys carry
 * the set index above the set-local Morton key.  nsets * 8^(tree_height-1)
 * must stay below 8^20. */
#define HFMM_MAX_RHS 64
int hfmm_set_particle_sets(hfmm_ctx* ctx, int32_t nsets, const int64_t* set_offsets, int32_t nrhs, int64_t n,
                           const double* x, const double* y, const double* z, const double* q);
/* Same with arrays already resident in device memory of the context's device;
 * the context reads them during hfmm_build_tree and hfmm_evaluate, so they must
 * stay valid until the next hfmm_set_particles*. */
int hfmm_set_particles_device(hfmm_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                              const double* q);

/* New charges for the particles of the built tree (input order; host array, or
 * a device pointer that must stay valid until the next evaluation).  The tree
 * depends only on the positions, so evaluating several charge vectors on one
 * tree skips the rebuild; the reference's NVALS right-hand sides
 * (FChebSymKernel.hpp:45-488, template parameter NVALS) serve the same use. */
int hfmm_set_charges(hfmm_ctx* ctx, const double* q);
int hfmm_set_charges_device(hfmm_ctx* ctx, const double* q);

/* Builds the octree on the device: leaf coordinates and Morton keys
 * (FCoordinateComputer.hpp:42-54, FTreeCoordinate.hpp:91-116), the sort by key
 * (FGroupTree.hpp:215-219), leaves and parent levels (FGroupTree.hpp:224-360),
 * the P2P neighbour lists (FTreeCoordinate.hpp:189-214) and the M2L
 * interaction lists (FTreeCoordinate.hpp:219-264). */
int hfmm_build_tree(hfmm_ctx* ctx);

/* Runs the passes selected by ops (HFMM_* bits) over the built tree; replaces
 * FAbstractAlgorithm::execute(operationsToProceed) (Src/Core/FCoreCommon.hpp:73-78)
 * with FChebSymKernel's P2M/M2M/M2L/L2L/L2P/P2P (FChebSymKernel.hpp:181-486).
 * Potentials and forces are reset, then accumulated by the selected passes. */
int hfmm_evaluate(hfmm_ctx* ctx, uint32_t ops);
/* Same over the working levels [upper, lower), 2 <= upper <= lower <= height:
 * FAbstractAlgorithm::execute(operationsToProceed, upperWorkingLevel,
 * lowerWorkingLevel) (Src/Core/FCoreCommon.hpp:81-86) with the loop bounds of
 * FGroupSeqAlgorithm (FGroupSeqAlgorithm.hpp:88, 137, 238): M2M parents from
 * min(height-2, lower-1) down to upper, M2L at levels lower-1 .. upper, L2L
 * parents upper .. lower-2; P2M, L2P and P2P on the leaves.  hfmm_evaluate is
 * hfmm_evaluate_levels(ctx, ops, 2, height).  With particle sets the levels
 * are those of each set's tree. */
int hfmm_evaluate_levels(hfmm_ctx* ctx, uint32_t ops, int32_t upper, int32_t lower);

/* The group tree's particle blocks as the boundary: FGroupSeqAlgorithm /
 * FGroupTaskAlgorithm::execute(ops, upper, lower) (FGroupSeqAlgorithm.hpp:61-
 * 83, FGroupTaskAlgorithm.hpp:82-111) over an FGroupTree whose leaves are
 * FP2PGroupParticleContainer<double, 1, 1, nvals>.  blocks[b] is particle
 * group b's getRawBuffer() (GroupTree/Core/FGroupOfParticles.hpp:211-214: a
 * 64-byte header, the leaf headers, then x, y, z and the nvals physical values
 * as arrays of nbParticlesAllocatedInGroup doubles), attributes[b] its
 * getRawAttributesBuffer() (:222-231: potential[v], forceX[v], forceY[v],
 * forceZ[v] at arrays v, nvals+v, 2 nvals+v, 3 nvals+v).  The particles are
 * evaluated as one set (the blocks of one tree, in any block order) and the
 * results are ADDED to the attribute buffers, as the kernel's P2P and L2P add
 * into the containers (FChebSymKernel.hpp:432-486).  Fails with HFMM_ERR_ARG
 * on inconsistent headers (the reference asserts on the same conditions) or on
 * a particle whose position does not lie in the leaf it is stored under; the
 * attribute buffers are then untouched.  The context's tree height, box and
 * order must be the group tree's.  The cells of the group tree are not read or
 * written: multipoles and locals stay inside the engine. */
int hfmm_evaluate_particle_blocks(hfmm_ctx* ctx, int32_t nblocks, const void* const* blocks, double* const* attributes,
                                  int32_t nvals, uint32_t ops, int32_t upper, int32_t lower);

/* The group tree's cell blocks of one level (2 .. height-1) as the boundary:
 * blocks[b] is cell group b's getRawBuffer() (GroupTree/Core/FGroupOfCells.hpp:
 * 189-197: a 24-byte header {startingIndex, endingIndex, numberOfCellsInBlock},
 * the cells' Morton indexes, then one FBasicCellPOD {mortonIndex,
 * coordinates[3], level} per cell), multipoles[b] its getRawMultipoleBuffer()
 * and locals[b] its getRawLocalBuffer() (:199-242): per cell nvals x 2 p^3
 * doubles, RHS v at offset 2 p^3 v, the FChebCellPODPole / FChebCellPODLocal
 * layout (GroupTree/Chebyshev/FChebCellPOD.hpp:11-21, 85-100).
 * hfmm_get_cell_blocks writes the context's multipole and local expansions of
 * the cells of the blocks into the first p^3 values of each RHS slot (the
 * global-SVD scheme also writes its compressed expansions, k values, at
 * offset p^3, where FChebKernel keeps them, FChebKernel.hpp:103, 148); the
 * other values are untouched.  hfmm_set_cell_blocks copies them into the
 * context's cells (kept-cell mode, symmetric scheme): the next evaluation adds
 * into them, as the reference's passes add into the tree's cells.  Every cell
 * of the blocks must be a cell of the context's tree at that level (one
 * particle set); nvals is the context's charge-vector count (above 1 with
 * kept cells only); NULL buffers skip.  Fails with HFMM_ERR_ARG on a bad
 * header, indexes out of order or outside the block's interval, or an unknown
 * cell. */
int hfmm_get_cell_blocks(hfmm_ctx* ctx, int32_t level, int32_t nblocks, const void* const* blocks,
                         double* const* multipoles, double* const* locals, int32_t nvals);
int hfmm_set_cell_blocks(hfmm_ctx* ctx, int32_t level, int32_t nblocks, const void* const* blocks,
                         const double* const* multipoles, const double* const* locals, int32_t nvals);

/* M2L scheme.  HFMM_M2L_SYMMETRIC (the default) is FChebSymKernel's: 16
 * canonical operators compressed one by one, applied through the 343
 * symmetries.  HFMM_M2L_GLOBAL_SVD is FChebKernel's
 * (Kernels/Chebyshev/FChebKernel.hpp:93-200, FChebM2LHandler.hpp:216-523):
 * the 316 far-field operators share one pair of bases U, B of rank k from one
 * SVD of all of them, multipoles are compressed (B^T M) after P2M and M2M,
 * the M2L works on the compressed expansions (L~ += 2/w C_t M~) and U L~ is
 * added to the locals before L2L and L2P.  Selecting it builds the operators
 * (host, FP64: a Householder QR and a Jacobi SVD of two 316 p^3 x p^3
 * matrices; seconds at order 6, about a minute at order 8).  FP64 precision
 * only; with kept cells, one charge vector.  hfmm_get_global_operators
 * returns the rank, U and B (p^3 x k, column-major), C ([343][k x k],
 * column-major, zero at the near positions) and the singular values of K_col
 * and K_row (p^3 each; zero for the dense scheme) of the scheme selected
 * last; any pointer may be NULL. */
#define HFMM_M2L_SYMMETRIC 0
#define HFMM_M2L_GLOBAL_SVD 1
#define HFMM_M2L_DENSE 2 /* FChebDenseKernel (FChebDenseKernel.hpp:93-165,
                            FChebDenseM2LHandler.hpp:180-215): the 316 operators
                            uncompressed (rank p^3, identity bases) */
#define HFMM_M2L_UNIFORM 3 /* FUnifKernel (Kernels/Uniform/FUnifKernel.hpp:72-175):
                              P2M, M2M, L2L and L2P interpolate on the equispaced
                              roots of FUnifRoots with Lagrange polynomials; the M2L
                              works in Fourier space (FUnifM2LHandler.hpp:26-297):
                              expansions zero-padded to (2p-1)^3 and DFT'd after
                              P2M / M2M, L^ += 2/w K^_t . M^ entrywise, the inverse
                              DFT added to the locals before L2L / L2P */
int hfmm_set_m2l_scheme(hfmm_ctx* ctx, int32_t scheme);
int hfmm_get_global_operators(hfmm_ctx* ctx, int32_t* rank, double* U, double* B, double* C, double* S1, double* S2);

/* Near field beside the far field (default off): a context that owns its
 * stream runs the P2P on a second stream of its own, concurrently with P2M,
 * M2M, M2L and L2L, and the L2P waits for both (the reference runs these
 * passes as independent tasks too, FGroupTaskAlgorithm.hpp:82-111).  Results
 * are the same bit for bit either way; contexts of a batch and profiled
 * evaluations run one kernel at a time.  Off by default: on MI355X the
 * persistent P2P waves leave no register room for the M2L waves beside them,
 * so the passes do not overlap and the M2L autotuner, timing under the P2P,
 * picks worse tile shapes.  No counterpart in the reference. */
int hfmm_set_stream_overlap(hfmm_ctx* ctx, int32_t enable);

/* Near field of the big leaves in the mixed forms (default on): each unordered
 * pair of two neighbour leaves with more than 64 particles is evaluated once
 * and added to both particles, as the reference's P2POuter and its algorithms'
 * opposite-direction calls do (FChebSymKernel.hpp:468-478,
 * FGroupSeqAlgorithm.hpp:343-394, FP2P.hpp:103-171).  The source-side sums of
 * a pair go through per-(particle, neighbour work item) slots in device memory
 * that a second kernel adds in a fixed order, so results do not depend on the
 * launch.  Off: every ordered pair is evaluated by its target's kernel.  The
 * FP64 form always evaluates ordered pairs.  A tree whose slots would need
 * more than a quarter of the free device memory falls back to ordered pairs. */
int hfmm_set_p2p_mutual(hfmm_ctx* ctx, int32_t enable);

/* The state of the cells (multipoles and locals) between evaluations.
 * HFMM_CELLS_FRESH (the default): every evaluation starts from cells in their
 * initial state, as after tree.forEachCell(resetToInitialState)
 * (FChebCell.hpp:151-152), so evaluating twice gives the same results.
 * HFMM_CELLS_KEPT: the cells keep their expansions from one evaluation to the
 * next and every pass adds into them, as the reference's P2M, M2M, M2L and
 * L2L add into the tree's cells (FChebInterpolator.hpp:627-700, 560-582;
 * FChebSymKernel.hpp:193-206, 210-328, 331-350): execute(P2M|M2M) followed by
 * execute(M2L|L2L|L2P) computes the far field, and executing twice without a
 * reset accumulates as the reference does.  The particle results of an
 * evaluation are its own contributions in both modes (the containers of the
 * reference's trees add them up).  A new tree (hfmm_build_tree), a change of
 * mode and hfmm_reset_cells put the cells back in their initial state.  With
 * several charge vectors every vector keeps its own cells; hfmm_get_expansions
 * shows the last vector's. */
#define HFMM_CELLS_FRESH 0
#define HFMM_CELLS_KEPT 1
int hfmm_set_cell_mode(hfmm_ctx* ctx, int32_t mode);
int hfmm_reset_cells(hfmm_ctx* ctx);

/* Waits for the context's stream. */
int hfmm_synchronize(hfmm_ctx* ctx);

/* Copies potentials and forces (input order) to host arrays of n doubles.
 * force = q * grad(potential), the sign convention of the reference's P2P and
 * L2P (FP2P.hpp:139-162, FChebInterpolator.hpp:1405-1410). Any pointer may be NULL. */
int hfmm_get_results(hfmm_ctx* ctx, double* pot, double* fx, double* fy, double* fz);
/* Device pointers (owned by the context, input order) of the same results. */
int hfmm_get_results_device(hfmm_ctx* ctx, const double** pot, const double** fx, const double** fy,
                            const double** fz);

/* FP64 direct sum over all particles of the context for the given targets
 * (input-order indices, host array): potential and force as the FMM reports
 * them, the O(N^2) reference the reference's test programs compute with
 * FP2P::NonMutualParticles (Tests/Kernels/DirectAlgorithm.cpp:203-215).
 * Needs hfmm_set_particles*; results are host arrays of ntargets entries
 * (NULL skips one). */
int hfmm_direct(hfmm_ctx* ctx, int64_t ntargets, const int64_t* targets, double* pot, double* fx, double* fy,
                double* fz);

/* hfmm_direct without device allocation: the same kernel and the same results,
 * bit for bit, with the target indices and results staged in one fixed
 * device array per GPU (HFMM_DIRECT_NOALLOC_MAX_TARGETS x 40 bytes, 2.5 MiB,
 * held until the process ends) instead of buffers allocated per call.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  More targets than the maximum fail with HFMM_ERR_ARG.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * threads, two contexts) race and must be serialised by the caller. */
#define HFMM_DIRECT_NOALLOC_MAX_TARGETS 65536
int hfmm_direct_noalloc(hfmm_ctx* ctx, int64_t ntargets, const int64_t* targets, double* pot, double* fx, double* fy,
                        double* fz);

/* The field grad(phi) at every particle (input order): the force without the
 * target charge, F = q grad(phi) (FChebSymKernel.hpp:432-486 accumulate q times
 * it).  Needed by the adjoint of the potential: for L(phi), dL/dx_j =
 * g_j grad(phi)(x_j) + q_j grad(psi)(x_j) with g = dL/dphi and psi the
 * potential of charges g.  Host arrays (NULL skips one), or device pointers
 * owned by the context. */
int hfmm_get_field(hfmm_ctx* ctx, double* ex, double* ey, double* ez);
int hfmm_get_field_device(hfmm_ctx* ctx, const double** ex, const double** ey, const double** ez);

/* ---- inspection (used by the parity tests) -------------------------------- */
typedef struct hfmm_tree_info {
    int64_t n_particles;
    int32_t height;
    int32_t order;
    int64_t n_cells[21];       /* per level; n_cells[height-1] = leaves */
    int64_t p2p_entries;       /* total P2P list entries over all leaves */
    int64_t m2l_entries[21];   /* total M2L list entries per level */
} hfmm_tree_info;
int hfmm_get_tree_info(hfmm_ctx* ctx, hfmm_tree_info* info);

/* Sorted particle keys and the original index of each sorted particle. */
int hfmm_get_sorted_keys(hfmm_ctx* ctx, int64_t* keys, int64_t* perm);
/* Morton keys of the cells of one level, ascending. */
int hfmm_get_cell_keys(hfmm_ctx* ctx, int32_t level, int64_t* keys);
/* Leaf particle ranges in sorted order. */
int hfmm_get_leaf_ranges(hfmm_ctx* ctx, int64_t* start, int64_t* count);
/* P2P list in CSR form and reference order: offsets[n_leaves+1], nbr[] = leaf
 * index, pos[] = 0..26 (the neighborPositions of FAbstractKernels::P2P). */
int hfmm_get_p2p_list(hfmm_ctx* ctx, int64_t* offsets, int64_t* nbr, int32_t* pos);
/* M2L list of a level (2..height-1) in CSR form and reference order:
 * offsets[n_cells+1], src[] = cell index in the level, pos[] = 0..342. */
int hfmm_get_m2l_list(hfmm_ctx* ctx, int32_t level, int64_t* offsets, int64_t* src, int32_t* pos);
/* Interactions of one level per relative position: counts[343] (position
 * ((dx+3)*7 + dy+3)*7 + dz+3), counted on the device; the reference's flop
 * count of the level is sum_pos counts[pos] (4 p^3 k_pos - k_pos - p^3)
 * (FChebFlopsSymKernel.hpp:62-63). */
int hfmm_get_m2l_counts(hfmm_ctx* ctx, int32_t level, int64_t* counts);
/* Compressed operator of canonical position idx (FChebSymM2LHandler.hpp:165-189):
 * *rank, and K = [U | V] column-major nnodes x rank each (2*rank*nnodes doubles,
 * K may be NULL to query the rank). */
int hfmm_get_operator(hfmm_ctx* ctx, int32_t idx, int32_t* rank, double* K);
/* Multipole and local expansions of a level after hfmm_evaluate:
 * [n_cells][order^3] each, either pointer may be NULL. */
int hfmm_get_expansions(hfmm_ctx* ctx, int32_t level, double* M, double* L);

/* ---- timing ------------------------------------------------------------- */
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream; hfmm_get_kernel_times() synchronizes and reports the sum of
 * event durations and the launch count per kernel name. */
typedef struct hfmm_kernel_time {
    char name[48];
    double total_ms;
    int64_t launches;
} hfmm_kernel_time;
int hfmm_set_profiling(hfmm_ctx* ctx, int32_t enable);
/* Restricts the timed kernels to a comma-separated list of the names
 * hfmm_get_kernel_times reports (e.g. "p2p_mixed_big,m2l_mixed"); NULL or ""
 * times every kernel.  With a short list a profiled evaluation records only a
 * few events per step.  No counterpart in the reference. */
int hfmm_set_profiling_filter(hfmm_ctx* ctx, const char* names);
int hfmm_get_kernel_times(hfmm_ctx* ctx, hfmm_kernel_time* out, int32_t max, int32_t* count);
int hfmm_reset_kernel_times(hfmm_ctx* ctx);

/* ---- diagnostics ------------------------------------------------------------ */
/* Runs one v_mfma_f64_16x16x4f64 on A[i][k] = 1 + 4i + k, B[k][j] = 0.5 + 16k + j
 * and writes D = A B (16 x 16, row-major) as read back through the operand
 * layout the M2L kernel assumes; no counterpart in the reference (a check of
 * the CDNA4 matrix-core lane layout). */
int hfmm_selftest_mfma_f64(hfmm_ctx* ctx, double* out);
/* Runs one v_mfma_f32_16x16x32_bf16 on A[i][k] = (7i + 3k) mod 11 - 5,
 * B[k][j] = (5k + 2j) mod 13 - 6 (exact in bfloat16 and FP32) and writes
 * D = A B (16 x 16, row-major) through the operand layout the mixed-precision
 * M2L kernel assumes; no counterpart in the reference. */
int hfmm_selftest_mfma_bf16(hfmm_ctx* ctx, float* out);

/* ---- files -------------------------------------------------------------------- */
/* Particle files of FFmaGenericLoader / FFmaGenericWriter
 * (Src/Files/FFmaGenericLoader.hpp:34-850): ".bfma" in the name selects the
 * binary format, else ".fma" the text format.  The header holds the element
 * size (8), the values per particle (>= 4: x, y, z, q, then e.g. potential and
 * forces), the particle count, HALF the box width and the box centre.
 * hfmm_fma_read with data = NULL reads the header only; data then takes
 * n * nvals values, particle-major.  box_width is the full width. */
int hfmm_fma_read(const char* path, int64_t* n, int32_t* nvals, double* box_width, double* center, double* data);
int hfmm_fma_write(const char* path, int64_t n, int32_t nvals, double box_width, const double* center,
                   const double* data);
/* The compressed M2L operators in the reference's cache file format
 * (ComputeAndCompressAndStoreInBinaryFile / ReadFromBinaryFile,
 * FChebSymM2LHandler.hpp:471-567): per stored operator, int index, int rank,
 * U then V (rank * p^3 doubles each, column-major).  Loading replaces the
 * context's operators (the 16 canonical ones must be present; ranks <= 64) and
 * re-packs the device tables.  hfmm_operator_file_name writes the reference's
 * file name, sym2l_d_o<ORDER>_e<epsilon>.bin (epsilon <= 0: 10^-order). */
int hfmm_save_operators(hfmm_ctx* ctx, const char* path);
int hfmm_load_operators(hfmm_ctx* ctx, const char* path);
int hfmm_operator_file_name(int32_t order, double epsilon, char* buf, int32_t size);

/* ---- M2L kernel variants (autotuned) ---------------------------------------- */
/* The mixed-precision M2L has several tile shapes: 0 = four waves share 64
 * targets (every order), 1 = one 16-target tile per wave (orders 5 and 6),
 * 2 = two 16-target tiles per wave (order 5), 3 = four waves share 32
 * targets (every order).  Variant -1 (the default) times
 * every available variant once at the next M2L on the actual tree and keeps
 * the fastest for the context's later evaluations.  The FP64 form has one
 * kernel (its variant stays -1).  hfmm_get_m2l_variant also returns the
 * times (ms) the last tuning measured per variant (4 floats, 0 = not run).
 * No counterpart in the reference (its M2L has one CPU implementation). */
int hfmm_set_m2l_variant(hfmm_ctx* ctx, int32_t variant);
int hfmm_get_m2l_variant(hfmm_ctx* ctx, int32_t* variant, float* tune_ms);

/* ---- batches of particle sets ----------------------------------------------- */
/* Many independent particle sets of uneven sizes, each with nrhs charge
 * vectors, evaluated by one launch of a HIP graph.  Sets of equal tree height
 * share one device tree (hfmm_set_particle_sets); the per-height trees share
 * one stream, and the first hfmm_batch_evaluate after a build runs the passes
 * and captures them, every RHS of every set, into one graph that later calls
 * launch.  No counterpart in the reference (one FGroupTree and one algorithm
 * run per set there). */
typedef struct hfmm_batch hfmm_batch;
/* Default height of a set of n particles: the smallest h in [3, 20] with
 * n <= 32 * 8^(h-1) (at most 32 particles per leaf on average). */
int32_t hfmm_batch_default_height(int64_t n);
/* params: order, box (each set's own box), epsilon, precision and device;
 * params->tree_height is ignored.  set_offsets: nsets+1 entries as in
 * hfmm_set_particle_sets; set_heights: nsets heights in [3, 20], or NULL for
 * hfmm_batch_default_height of each set's size. */
int hfmm_batch_create(const hfmm_params* params, int32_t nsets, const int64_t* set_offsets,
                      const int32_t* set_heights, int32_t nrhs, hfmm_batch** out);
int hfmm_batch_destroy(hfmm_batch* b);
const char* hfmm_batch_last_error(const hfmm_batch* b);
/* Host arrays of all sets, concatenated in set order; q holds nrhs vectors of
 * n charges (RHS-major). */
int hfmm_batch_set_particles(hfmm_batch* b, const double* x, const double* y, const double* z, const double* q);
int hfmm_batch_build(hfmm_batch* b);
/* Evaluates every set and RHS (ops as in hfmm_evaluate).  The first call after
 * a build, or with other ops, runs the passes eagerly and captures them; later
 * calls launch the captured graph.  Returns without waiting. */
int hfmm_batch_evaluate(hfmm_batch* b, uint32_t ops);
int hfmm_batch_synchronize(hfmm_batch* b);
/* nrhs blocks of n values per output (RHS-major, input order); NULL skips one. */
int hfmm_batch_get_results(hfmm_batch* b, double* pot, double* fx, double* fy, double* fz);
/* Number of per-height trees and whether the next evaluate launches a graph. */
int hfmm_batch_info(const hfmm_batch* b, int32_t* ntrees, int32_t* graph_ready);
/* The context of tree i (0 .. ntrees-1, increasing height), owned by the batch,
 * for inspection (tree info, lists, counts) and profiling: while a tree has
 * profiling enabled, hfmm_batch_evaluate runs the passes eagerly so that the
 * kernel events are recorded. */
int hfmm_batch_tree(hfmm_batch* b, int32_t i, hfmm_ctx** tree, int32_t* height, int32_t* nsets);

/* Library version string. */
const char* hfmm_version(void);

#ifdef __cplusplus
}
#endif
#endif
