// hfmm_internal.hpp -- context and device-buffer plumbing shared by the host
// API (hfmm_api.cpp) and the device code (tree_device.hip, fmm_device.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/scalfmm_hip.h"
#include "cheb_setup.hpp"

namespace hfmm {

constexpr int kMaxLevels = 21;
constexpr int kNbrSlots = 26;   // P2P neighbour slots per leaf
constexpr int kFarSlots = 189;  // M2L slots per cell
constexpr int kTileTargets = 16;  // targets per M2L tile (the MFMA N dimension)
constexpr int kP2PBig = 64;    // leaves with more particles run P2P through LDS-staged sources
constexpr int kP2PItem = 128;  // targets per P2P work item of a big leaf (two per lane of a wave)
constexpr int kP2PMutualItem = 256;  // targets per work item of the mutual mixed P2P (four per lane)
constexpr int kNbrForward = 13;  // neighbour slots 0..12 (positions 0..12): the mutual P2P's forward half
constexpr int kSortedPad = 16; // padding doubles after the sorted particle SoA
constexpr int kWaveM2LMaxOrder = 6;  // highest order of the wave-level mixed M2L kernel

// Grow-only device allocation.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

struct KernelTimer {
    std::string name;
    hipEvent_t start, stop;
};

// FChebKernel's M2L operators (one global SVD; cheb_global.cpp): U and B
// nn x k column-major (un-weighted), C [343][k x k] column-major, S1 / S2 the
// singular values of K_col / K_row.
struct GlobalOperators {
    int k = 0;  // doubles per compressed expansion (the uniform scheme: 2 NF)
    std::vector<double> U, B, C, S1, S2;
    // the uniform scheme: kernel spectra [343][NF] complex (zero at the near
    // positions) and the twiddles w^j = (cos 2 pi j/N, sin 2 pi j/N), N = 2p-1
    std::vector<double> Khat, tw;
};
int build_global_operators(const ChebOperators& op, GlobalOperators& g, std::string& err);
// FChebDenseKernel's operators: rank p^3, identity bases, C_t = K_t
int build_dense_operators(const ChebOperators& op, GlobalOperators& g);
// FUnifKernel's M2L in Fourier space (Kernels/Uniform/FUnifM2LHandler.hpp:26-
// 120, 244-297): the kernel between the equispaced nodes of two cells depends
// on the node index difference only, so each far-field operator is a 3-level
// Toeplitz matrix, embedded in a 3D circulant of side N = 2p-1 and applied as
// a pointwise product of 3D DFTs.  Builds the kernel spectra (half spectrum:
// kx < p, NF = p N^2 entries per position, index kx + p (ky + N kz)) of the
// uniform basis op and the twiddle table; g.k = 2 NF.
int build_uniform_fft_operators(const ChebOperators& op, GlobalOperators& g);

struct Context {
    hfmm_params P{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false when the stream belongs to a batch
    // Near field beside the far field (off by default): with its own stream a
    // context can run the P2P on `side` while P2M..L2L run on `stream` (L2P
    // waits for both)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool overlap = false;
    std::string err;

    // ---- operators (host + device copies)
    ChebOperators ops;
    int LD = 0;              // leading dimension of one expansion (nn rounded up to 4)
    int kpad_max = 0;        // largest padded rank
    DevBuf d_Troots;         // [p][p]
    DevBuf d_child;          // [8][3][p][p]
    DevBuf d_pidx;           // [343] uint32
    DevBuf d_pvec;           // [343][nn] uint16
    DevBuf d_rank;           // [343] int32 (by canonical index)
    DevBuf d_opoff;          // [343] int64 offset of U_c (doubles) in d_ops
    DevBuf d_ops;            // packed U_c, V_c row-major [nn][kpad]
    DevBuf d_slotpos;        // [8][189] int16 relative position of each slot
    DevBuf d_slotord;        // [8][189][2] int16 (slot, position) ordered by rank tiles
    DevBuf d_slotseg;        // [8][5] int32 start of the 1..4 rank-tile segments (+ end)
    DevBuf d_posoff;         // [343] int64 offset of a position's mixed operators in d_opm
    DevBuf d_opm;            // mixed precision: position-specific operators, bf16 hi/mid/lo planes
    DevBuf d_opm2;           // the same with U's rank columns interleaved (wave-level kernel, p <= 6)
    int m2l_variant = -1;    // M2L kernel variant; -1: pick by timing at the next M2L
    float m2l_tune_ms[4] = {0.f, 0.f, 0.f, 0.f};  // times measured by the last autotune

    // ---- particles
    int64_t n = 0;
    const double* x = nullptr;  // input order (device)
    const double* y = nullptr;
    const double* z = nullptr;
    const double* q = nullptr;
    DevBuf d_in;  // owned copy when set from host: x, y, z, then nrhs charge vectors
    DevBuf d_qin; // charges set by hfmm_set_charges (host variant)
    // Particle sets (a forest): set s owns input particles setoff[s]..setoff[s+1]-1
    // and lives in its own copy of the box.  Its index sits in the key bits
    // above the set-local Morton key, i.e. set s is the cell s of level
    // set_levels of a virtual tree whose levels above set_levels+2 carry no
    // work.  One set: set_levels = 0 and the virtual tree is the tree.
    int32_t nsets = 1;
    int set_levels = 0;
    DevBuf d_setoff;  // int64 [nsets + 1] (unused when nsets == 1)
    int32_t nrhs = 1;  // charge vectors per particle (RHS-major, n each)
    // working levels [work_upper, work_lower) of the current evaluation, in
    // levels of the virtual tree (FAbstractAlgorithm upper/lowerWorkingLevel)
    int work_upper = 0, work_lower = 0;

    // ---- tree
    bool tree_built = false;
    bool evaluated = false;
    DevBuf d_keys, d_keys2, d_vals, d_vals2;  // sort buffers (uint64 / int32)
    DevBuf d_hist;                            // radix histogram / scan scratch
    DevBuf d_scan_part;                       // scan partials
    DevBuf d_flags, d_scan;                   // int32 per element
    DevBuf d_sorted;                          // sx, sy, sz, sq (4*n doubles)
    DevBuf d_perm;                            // int32 sorted -> input index
    DevBuf d_iperm;                           // int32 input -> sorted index
    DevBuf d_pleaf;                           // int32 leaf of each sorted particle
    DevBuf d_count;                           // small device counters (int64)
    int64_t ncell[kMaxLevels] = {0};
    int64_t lvl_off[kMaxLevels + 1] = {0};    // offsets in the cell arrays
    int64_t total_cells = 0;
    DevBuf d_cellkey;                         // uint64 [total_cells]
    DevBuf d_cellkey_tmp;                     // scratch per level
    DevBuf d_childstart;                      // int32 [total_cells + levels] (ncell[l]+1 per level)
    int64_t cs_off[kMaxLevels] = {0};
    DevBuf d_leafstart;                       // int32 [nleaf+1]
    DevBuf d_p2p;                             // int32 [nleaf][26]
    DevBuf d_p2pitems;                        // int2 (leaf, first target) work items of the big leaves
    int64_t p2p_items_cap = 0;                // upper bound of the item count (the count itself is on the device)
    // Mutual mixed P2P (each unordered pair of big leaves evaluated once): work
    // items of kP2PMutualItem targets, and the source-side sums that a leaf
    // receives from the items of its big backward neighbours (slots 13..25):
    // per leaf 13 slot offsets (-1: none) and the first float4 of its slots
    // (each slot holds one float4 per particle of the leaf)
    bool p2p_mutual = false;                  // tables built and the partial buffer allocated
    bool p2p_mutual_on = true;                // hfmm_set_p2p_mutual (the ordered-pair kernel when off)
    DevBuf d_p2pitems_m;                      // int2 (leaf, first target), costliest first
    DevBuf d_itemkeys, d_itemidx, d_items_tmp; // sorting scratch of the mutual work items
    int64_t p2p_items_m_cap = 0;
    DevBuf d_csoff;                           // int32 [nleaf][13]
    DevBuf d_csbase;                          // int64 [nleaf + 1]
    DevBuf d_cpart;                           // float4 source-side partial sums
    int64_t cpart_units = 0;                  // float4 entries in use
    DevBuf d_m2l;                             // int32 [cells of levels 2..h-1][189]
    int64_t m2l_off[kMaxLevels] = {0};        // offset (in cells) of a level in d_m2l
    DevBuf d_order;                           // int32 parity-sorted cell index per level (same offsets as d_m2l)
    DevBuf d_pcount;                          // int32 [levels][8]
    DevBuf d_dense;                           // int32 Morton -> cell index tables (levels <= 8)

    // ---- expansions and results
    DevBuf d_M, d_L;      // [total_cells][LD]
    DevBuf d_Mb;          // mixed precision: multipoles in FP32 [total_cells + 1][LD] (the M2L splits them)
    DevBuf d_loc;         // mixed precision: leaf-local fixed-point int4 (x, y, z, q bits) of the sorted particles
    DevBuf d_res_sorted;  // pot, grad(phi) x, y, z (sorted order; the charge is applied by the scatter),
                          // one such block per charge vector when the near field runs for all vectors at once
    DevBuf d_qall;        // the sorted charges of every vector, in chunks of R vectors laid out [n][R]
    int res_rhs = 0;      // the vector whose block of d_res_sorted the passes read and write
    bool p2p_small_done = false;  // the leaves of at most kP2PBig particles already hold their near field
    DevBuf d_res;         // pot, fx, fy, fz (input order)
    DevBuf d_field;       // grad(phi) x, y, z (input order): the force without the target charge
    // Cell state between evaluations: fresh (every evaluation starts from zero
    // expansions) or kept (the passes add into the expansions of earlier ones)
    bool keep_cells = false;
    bool cells_dirty = true;  // kept mode: zero M and L before the next evaluation
    // M2L scheme: HFMM_M2L_SYMMETRIC (FChebSymKernel) or HFMM_M2L_GLOBAL_SVD
    // (FChebKernel: compressed multipoles / locals of rank k beside M and L)
    int m2l_scheme = 0;
    // the 16 symmetric operators: compressed (host, c->ops.K) and uploaded
    // (d_ops, d_rank, ...); a context created with another scheme defers both
    bool sym_compressed = false, sym_uploaded = false;
    // interpolation basis of P2M, M2M, L2L and L2P: 0 = Chebyshev (the
    // symmetric, global-SVD and dense schemes), 1 = uniform Lagrange
    // (HFMM_M2L_UNIFORM: d_child then holds the Lagrange child tables)
    int basis = 0;
    DevBuf d_uroots;          // [p] the equispaced roots of the uniform basis
    DevBuf d_uK;              // the uniform scheme's kernel spectra [343][NF] complex
    DevBuf d_utw;             // the uniform scheme's twiddles [N] complex
    GlobalOperators glob;
    DevBuf d_gBt;             // B^T as [nn][k] (the compression reads a row per node)
    DevBuf d_gU;              // U as [k][nn]
    DevBuf d_gC;              // C [343][k x k]
    DevBuf d_gCr;             // C as [343][KP][KP] row-major, zero-padded (the MFMA M2L's operands)
    int gcr_kp = 0;           // KP of d_gCr (0: not built for the current operators)
    DevBuf d_Mt, d_Lt;        // [total_cells][k]
    DevBuf d_keep;            // kept mode: snapshot of the expansions a pass overwrites
    DevBuf d_kept;            // kept mode, nrhs > 1: M and L of every charge vector
    std::vector<double> block_xyz;  // positions of the last particle-block evaluation (tree reuse)

    // ---- profiling
    bool profiling = false;
    std::string prof_filter;  // ",name,name," of the kernels to time; empty: every kernel
    bool prof_open = false;   // prof_begin recorded a start event that prof_end closes
    std::vector<KernelTimer> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<hfmm_kernel_time> totals;

    // height of the virtual tree: the per-set height plus the set levels
    int height() const { return P.tree_height + set_levels; }
    // first level with M2L / M2M / L2L work (2 below each set root)
    int top() const { return 2 + set_levels; }
    int order() const { return P.order; }
};


// M2L and P2P in the mixed forms (three-part or two-part bfloat16 splits)
inline bool is_mixed(const Context* c) {
    return c->P.precision == HFMM_PRECISION_MIXED || c->P.precision == HFMM_PRECISION_MIXED_HILO;
}

// Kept cells: allocates the expansion buffers of the built tree and zeroes them
// when the tree is new or the cells were reset (hfmm_api.cpp).
int prepare_kept_cells(Context* c);

// Error helpers ---------------------------------------------------------------
int fail(Context* c, int code, const std::string& msg);
#define HFMM_HIP(ctx, call)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::hfmm::fail((ctx), HFMM_ERR_DEVICE,                                  \
                                std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)
#define HFMM_TRY(expr)            \
    do {                          \
        int s_ = (expr);          \
        if (s_ != HFMM_OK) return s_; \
    } while (0)

int ensure(Context* c, DevBuf& b, size_t bytes);
void release(DevBuf& b);

// Profiling: bracket a launch by events when enabled.
void prof_begin(Context* c, const char* name);
void prof_end(Context* c);
int check_launch(Context* c, const char* name);

// The 16 canonical compressed M2L operators (FP64, on the device) into c->ops.
int compress_operators_device(Context* c);
// A context on the given stream (nullptr: a stream of its own).
int create_context(const hfmm_params* params, hipStream_t shared_stream, hfmm_ctx** out);
// Operator cache file of the reference (fma_io.cpp).
int save_operators(Context* c, const char* path);
int read_operators(Context* c, const char* path, std::array<int, 343>& rank, std::array<std::vector<double>, 343>& K);
// The context behind a handle.
Context* context_of(hfmm_ctx* h);

// Tree construction (tree_device.hip)
int build_tree_device(Context* c);
// New charges (c->q, input order) on the built tree: refreshes the sorted charges.
int update_charges_device(Context* c, int rhs = 0);
// Occupied M2L slots of a level per (parity, slot) into counts[8][189] (device).
int m2l_slot_counts(Context* c, int level, unsigned long long* counts);
// Device counter holding the number of P2P work items of the big leaves.
inline int32_t* p2p_item_count(Context* c) { return reinterpret_cast<int32_t*>(c->d_count.as<int64_t>() + 12); }
// Device counters of the mutual mixed P2P: its work item count and the total of
// its source-side partial sums (float4 entries).
inline int32_t* p2p_mutual_item_count(Context* c) { return reinterpret_cast<int32_t*>(c->d_count.as<int64_t>() + 16); }
inline int64_t* p2p_mutual_units(Context* c) { return c->d_count.as<int64_t>() + 17; }
// The next work item of the mutual mixed P2P (claimed by the waves one at a
// time; cleared before each launch).
inline int32_t* p2p_mutual_next(Context* c) { return reinterpret_cast<int32_t*>(c->d_count.as<int64_t>() + 18); }
// Builds the mutual P2P's work items and slot tables on a built tree and sizes
// the partial buffer (mixed forms); leaves c->p2p_mutual false when the
// partial sums would exceed a quarter of the free device memory.
int build_p2p_mutual_tables(Context* c);

// FP64 passes (fmm_device.hip)
int run_p2m(Context* c);
int run_m2m(Context* c, int level);  // parents at `level`, children at level+1
int run_m2l(Context* c);  // every level 2..h-1 in one launch
int run_l2l(Context* c, int level);  // parents at `level`
int run_l2p(Context* c);
int run_p2p(Context* c);
int zero_results(Context* c);
// The sorted results of vector c->res_rhs.
inline double* sorted_results(Context* c) {
    return c->d_res_sorted.as<double>() + static_cast<size_t>(c->res_rhs) * 4 * static_cast<size_t>(c->n);
}
// Charge vectors of the all-vector near field in chunks of R = 8, 4, 2, 1
// (largest first): the chunk of vectors r0 .. r0 + R - 1 holds its sorted
// charges at d_qall + n r0 (elements of the form's type) as [n][R].
struct RhsChunk {
    int r0, R;
};
inline int rhs_chunks(int nrhs, RhsChunk* out) {
    int k = 0;
    for (int r0 = 0; r0 < nrhs;) {
        const int left = nrhs - r0;
        const int R = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
        out[k++] = RhsChunk{r0, R};
        r0 += R;
    }
    return k;
}
// The sorted charges of every vector into d_qall (FP32 for the mixed forms).
int gather_all_charges(Context* c, bool f32);
// Near field of the leaves of at most kP2PBig particles for every charge vector
// at once (the pair geometry once per pair), into the per-vector blocks of
// d_res_sorted; the big leaves then run per vector.
int run_p2p_small_all(Context* c);
int run_p2p_mixed_small_all(Context* c);
// results of charge vector `rhs` into its slot of the input-order outputs
int scatter_results(Context* c, int rhs = 0);
// clears device memory with a fill kernel (capturable, unlike hipMemsetAsync)
int zero_device(Context* c, void* ptr, size_t bytes);
// FChebKernel passes (fmm_global.hip): M~ = B^T M for the cells of level l,
// L += U L~ for the cells of level l, and the compressed M2L over the working
// levels (L~ += 2/w C_t M~)
int run_global_compress(Context* c, int l);
int run_global_decompress(Context* c, int l);
int run_global_m2l(Context* c);
// dst[i] = src[i] (add = false) or dst[i] += src[i] (add = true), i < count
int combine_device(Context* c, double* dst, const double* src, int64_t count, bool add);
int selftest_mfma_f64(Context* c, double* out);

// Mixed-precision passes (fmm_mixed.hip)
int prepare_mixed_operators(Context* c);  // position-specific bf16 operators
int run_split_multipoles(Context* c);     // d_M -> d_Mb
int run_m2l_mixed(Context* c);
int run_p2p_mixed(Context* c);
int m2l_variant_count(Context* c);
bool m2l_variant_available(Context* c, int v);
int selftest_mfma_bf16(Context* c, float* out);

// FP64 direct sum for sampled targets (direct.hip)
int direct_sum(Context* c, int64_t nt, const int64_t* targets, double* pot, double* fx, double* fy, double* fz);
// hfmm_direct's kernel with its staging in the static per-device array;
// init_direct_scratch reads that array's address on c's device (once per
// device, at context creation).
int direct_sum_noalloc(Context* c, int64_t nt, const int64_t* targets, double* pot, double* fx, double* fy, double* fz);
int init_direct_scratch(Context* c);

}  // namespace hfmm
