// hfmm_api.cpp -- the extern "C" entry points of include/scalfmm_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "hfmm_internal.hpp"

namespace hfmm {

int fail(Context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int ensure(Context* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.bytes >= bytes) return HFMM_OK;
    if (b.ptr) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(c, HFMM_ERR_DEVICE, std::string("sync before realloc: ") + hipGetErrorString(e));
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return fail(c, HFMM_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return HFMM_OK;
}

void release(DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

static hipEvent_t take_event(Context* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(Context* c, const char* name) {
    c->prof_open = false;
    if (!c->profiling) return;
    if (!c->prof_filter.empty() && c->prof_filter.find("," + std::string(name) + ",") == std::string::npos) return;
    KernelTimer t;
    t.name = name;
    t.start = take_event(c);
    t.stop = take_event(c);
    if (!t.start || !t.stop) return;
    (void)hipEventRecord(t.start, c->stream);
    c->pending.push_back(t);
    c->prof_open = true;
}

void prof_end(Context* c) {
    if (!c->profiling || !c->prof_open || c->pending.empty()) return;
    (void)hipEventRecord(c->pending.back().stop, c->stream);
    c->prof_open = false;
}

int check_launch(Context* c, const char* name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, HFMM_ERR_DEVICE, std::string("launch ") + name + ": " + hipGetErrorString(e));
    return HFMM_OK;
}

static int collect_times(Context* c) {
    if (c->pending.empty()) return HFMM_OK;
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            auto it = std::find_if(c->totals.begin(), c->totals.end(),
                                   [&](const hfmm_kernel_time& k) { return t.name == k.name; });
            if (it == c->totals.end()) {
                hfmm_kernel_time k{};
                std::strncpy(k.name, t.name.c_str(), sizeof(k.name) - 1);
                c->totals.push_back(k);
                it = c->totals.end() - 1;
            }
            it->total_ms += ms;
            it->launches += 1;
        }
        c->event_pool.push_back(t.start);
        c->event_pool.push_back(t.stop);
    }
    c->pending.clear();
    return HFMM_OK;
}

// Uploads the operator tables.  The compressed operators are packed per
// canonical index as row-major [nn+1][kpad] blocks (U then V): kpad = rank
// rounded up to 16 with zero columns, and a zero row nn that the padded node
// permutations point at, so the M2L kernel needs no bounds masks.
static int upload_operators(Context* c) {
    const ChebOperators& op = c->ops;
    const int p = op.p, nn = op.nn;
    c->LD = (nn + 15) & ~15;
    HFMM_TRY(ensure(c, c->d_Troots, sizeof(double) * p * p));
    HFMM_HIP(c, hipMemcpy(c->d_Troots.ptr, op.Troots.data(), sizeof(double) * p * p, hipMemcpyHostToDevice));
    HFMM_TRY(ensure(c, c->d_child, sizeof(double) * op.child.size()));
    HFMM_HIP(c, hipMemcpy(c->d_child.ptr, op.child.data(), sizeof(double) * op.child.size(), hipMemcpyHostToDevice));
    HFMM_TRY(ensure(c, c->d_pidx, sizeof(uint32_t) * 343));
    HFMM_HIP(c, hipMemcpy(c->d_pidx.ptr, op.pidx.data(), sizeof(uint32_t) * 343, hipMemcpyHostToDevice));
    // node permutations padded to LD entries per position (padding -> zero row nn)
    std::vector<uint16_t> pvp(static_cast<size_t>(343) * c->LD, static_cast<uint16_t>(nn));
    for (int pos = 0; pos < 343; ++pos)
        for (int n = 0; n < nn; ++n) pvp[static_cast<size_t>(pos) * c->LD + n] = op.pvec[static_cast<size_t>(pos) * nn + n];
    HFMM_TRY(ensure(c, c->d_pvec, sizeof(uint16_t) * pvp.size()));
    HFMM_HIP(c, hipMemcpy(c->d_pvec.ptr, pvp.data(), sizeof(uint16_t) * pvp.size(), hipMemcpyHostToDevice));
    std::vector<int16_t> sp(8 * kFarSlots);
    for (int e = 0; e < 8; ++e) {
        FarSlot s[kFarSlots];
        far_slots(e, s);
        for (int j = 0; j < kFarSlots; ++j) sp[e * kFarSlots + j] = s[j].pos;
    }
    HFMM_TRY(ensure(c, c->d_slotpos, sizeof(int16_t) * sp.size()));
    HFMM_HIP(c, hipMemcpy(c->d_slotpos.ptr, sp.data(), sizeof(int16_t) * sp.size(), hipMemcpyHostToDevice));
    // the rest serves the symmetric scheme's compressed operators only
    if (!c->sym_compressed) return HFMM_OK;
    std::vector<int32_t> rank(343, 0);
    std::vector<int64_t> off(343, 0);
    std::vector<double> packed;
    int kpmax = 16;
    for (int idx = 0; idx < 343; ++idx) {
        const int k = op.rank[idx];
        if (!k) continue;
        const int kpad = (k + 15) & ~15;
        kpmax = std::max(kpmax, kpad);
        rank[idx] = k;
        off[idx] = static_cast<int64_t>(packed.size());
        packed.resize(packed.size() + static_cast<size_t>(2) * (nn + 1) * kpad, 0.0);
        double* U = packed.data() + off[idx];
        double* V = U + static_cast<size_t>(nn + 1) * kpad;
        const std::vector<double>& K = op.K[idx];
        for (int r = 0; r < k; ++r)
            for (int n = 0; n < nn; ++n) {
                U[static_cast<size_t>(n) * kpad + r] = K[static_cast<size_t>(r) * nn + n];
                V[static_cast<size_t>(n) * kpad + r] = K[static_cast<size_t>(k) * nn + static_cast<size_t>(r) * nn + n];
            }
    }
    c->kpad_max = kpmax;
    HFMM_TRY(ensure(c, c->d_rank, sizeof(int32_t) * 343));
    HFMM_HIP(c, hipMemcpy(c->d_rank.ptr, rank.data(), sizeof(int32_t) * 343, hipMemcpyHostToDevice));
    HFMM_TRY(ensure(c, c->d_opoff, sizeof(int64_t) * 343));
    HFMM_HIP(c, hipMemcpy(c->d_opoff.ptr, off.data(), sizeof(int64_t) * 343, hipMemcpyHostToDevice));
    HFMM_TRY(ensure(c, c->d_ops, sizeof(double) * packed.size()));
    HFMM_HIP(c, hipMemcpy(c->d_ops.ptr, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice));
    // slots of every parity ordered by the rank tiles of their operator, so the
    // M2L kernels run one specialised loop per rank-tile count.  Inside a
    // rank-tile group the slots are ordered by the parity class of their
    // offset, then by the offset: a tile's targets are one cell of every other
    // parent (a lattice of stride 2), so consecutive slots whose offsets differ
    // by 2 cells read largely the same source expansions from the cache.
    std::vector<int16_t> so(8 * kFarSlots * 2);
    std::vector<int32_t> seg(8 * 5);
    for (int e = 0; e < 8; ++e) {
        int i = 0;
        for (int kt = 1; kt <= 4; ++kt) {
            seg[e * 5 + kt - 1] = i;
            std::vector<std::pair<int, int>> group;  // (order key, slot)
            for (int j = 0; j < kFarSlots; ++j) {
                const int pos = sp[e * kFarSlots + j];
                const int k = op.rank[op.pidx[pos]];
                if (((k + 15) >> 4) != kt) continue;
                const int dx = pos / 49, dy = (pos / 7) % 7, dz = pos % 7;  // offsets + 3
                const int cls = ((dx + 1) & 1) * 4 + ((dy + 1) & 1) * 2 + ((dz + 1) & 1);
                group.emplace_back(cls * 343 + pos, j);
            }
            std::sort(group.begin(), group.end());
            for (const auto& g : group) {
                so[(e * kFarSlots + i) * 2] = static_cast<int16_t>(g.second);
                so[(e * kFarSlots + i) * 2 + 1] = sp[e * kFarSlots + g.second];
                ++i;
            }
        }
        seg[e * 5 + 4] = i;
        if (i != kFarSlots) return fail(c, HFMM_ERR_ARG, "operator rank above 64 is not supported");
    }
    HFMM_TRY(ensure(c, c->d_slotord, sizeof(int16_t) * so.size()));
    HFMM_HIP(c, hipMemcpy(c->d_slotord.ptr, so.data(), sizeof(int16_t) * so.size(), hipMemcpyHostToDevice));
    HFMM_TRY(ensure(c, c->d_slotseg, sizeof(int32_t) * seg.size()));
    HFMM_HIP(c, hipMemcpy(c->d_slotseg.ptr, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice));
    return HFMM_OK;
}

static void destroy_side(Context* c) {
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    c->ev_fork = c->ev_join = nullptr;
    c->side = nullptr;
}

static void free_all(Context* c) {
    DevBuf* bufs[] = {&c->d_Troots, &c->d_child, &c->d_pidx, &c->d_pvec, &c->d_rank, &c->d_opoff, &c->d_ops,
                      &c->d_slotpos, &c->d_slotord, &c->d_slotseg, &c->d_posoff, &c->d_opm, &c->d_opm2, &c->d_in, &c->d_keys, &c->d_keys2, &c->d_vals, &c->d_vals2, &c->d_hist,
                      &c->d_scan_part, &c->d_qin, &c->d_flags, &c->d_scan, &c->d_sorted, &c->d_perm, &c->d_iperm, &c->d_pleaf,
                      &c->d_count, &c->d_cellkey, &c->d_cellkey_tmp, &c->d_childstart, &c->d_leafstart,
                      &c->d_p2p, &c->d_p2pitems, &c->d_m2l, &c->d_order, &c->d_pcount, &c->d_dense, &c->d_M, &c->d_L, &c->d_Mb, &c->d_loc, &c->d_res_sorted, &c->d_qall,
                      &c->d_res, &c->d_field, &c->d_setoff, &c->d_keep, &c->d_kept, &c->d_gBt, &c->d_gU, &c->d_gC, &c->d_gCr, &c->d_Mt, &c->d_Lt, &c->d_uroots, &c->d_uK, &c->d_utw,
                      &c->d_p2pitems_m, &c->d_itemkeys, &c->d_itemidx, &c->d_items_tmp, &c->d_csoff, &c->d_csbase, &c->d_cpart};
    for (DevBuf* b : bufs) release(*b);
}

}  // namespace hfmm

using hfmm::Context;

struct hfmm_ctx {
    Context c;
};

#define CTX_CHECK(ctx)                    \
    do {                                  \
        if (!(ctx)) return HFMM_ERR_ARG;  \
    } while (0)

namespace hfmm {
// Kept cells: the expansion buffers of the built tree, put in their initial
// state (zero) when the tree is new or the cells were reset.
int prepare_kept_cells(Context* c) {
    const size_t exp_bytes = sizeof(double) * static_cast<size_t>(c->total_cells) * c->LD;
    HFMM_TRY(ensure(c, c->d_M, exp_bytes + sizeof(double) * c->LD));
    HFMM_TRY(ensure(c, c->d_L, exp_bytes));
    HFMM_TRY(ensure(c, c->d_keep, exp_bytes));
    const bool many = c->nrhs > 1;
    if (many) HFMM_TRY(ensure(c, c->d_kept, 2 * exp_bytes * c->nrhs));
    if (c->cells_dirty) {  // the cells of a new tree, or after hfmm_reset_cells
        HFMM_TRY(zero_device(c, c->d_M.ptr, exp_bytes));
        HFMM_TRY(zero_device(c, c->d_L.ptr, exp_bytes));
        if (many) HFMM_TRY(zero_device(c, c->d_kept.ptr, 2 * exp_bytes * c->nrhs));
        c->cells_dirty = false;
    }
    return HFMM_OK;
}
}  // namespace hfmm

extern "C" {

const char* hfmm_version(void) { return "scalfmm_amd 0.1 (gfx950, FP64 Chebyshev 1/r)"; }

int hfmm_create(const hfmm_params* params, hfmm_ctx** out) { return hfmm::create_context(params, nullptr, out); }

}  // extern "C"

namespace hfmm {

Context* context_of(hfmm_ctx* h) { return &h->c; }

int create_context(const hfmm_params* params, hipStream_t shared_stream, hfmm_ctx** out) {
    if (!params || !out) return HFMM_ERR_ARG;
    *out = nullptr;
    if (params->order < 5 || params->order > 8) return HFMM_ERR_ARG;
    if (params->tree_height < 3 || params->tree_height > 20) return HFMM_ERR_ARG;
    if (!(params->box_width > 0) || !std::isfinite(params->box_width)) return HFMM_ERR_ARG;
    if (params->precision != HFMM_PRECISION_FP64 && params->precision != HFMM_PRECISION_MIXED &&
        params->precision != HFMM_PRECISION_MIXED_HILO)
        return HFMM_ERR_ARG;
    hfmm_ctx* h = new (std::nothrow) hfmm_ctx();
    if (!h) return HFMM_ERR_NOMEM;
    Context* c = &h->c;
    c->P = *params;
    c->device = params->device;
    c->own_stream = shared_stream == nullptr;
    if (hipSetDevice(c->device) != hipSuccess) {
        delete h;
        return HFMM_ERR_DEVICE;
    }
    if (shared_stream) c->stream = shared_stream;
    else if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete h;
        return HFMM_ERR_DEVICE;
    }
    // the side stream of the near field (default priority, as the main one)
    if (c->own_stream &&
        (hipStreamCreate(&c->side) != hipSuccess ||
         hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)) {
        hfmm::destroy_side(c);
        (void)hipStreamDestroy(c->stream);
        delete h;
        return HFMM_ERR_DEVICE;
    }
    std::string err;
    const int scheme = params->m2l_scheme;
    int st = hfmm::init_direct_scratch(c);
    if (st == HFMM_OK) st = hfmm::build_cheb_operators(params->order, params->epsilon, c->ops, err);
    if (st == HFMM_OK && scheme != HFMM_M2L_SYMMETRIC &&
        (scheme < HFMM_M2L_GLOBAL_SVD || scheme > HFMM_M2L_UNIFORM || params->precision != HFMM_PRECISION_FP64))
        st = HFMM_ERR_ARG;  // unknown scheme, or a non-symmetric one outside FP64
    // the symmetric scheme's operators now, the others' below
    if (st == HFMM_OK && scheme == HFMM_M2L_SYMMETRIC) st = hfmm::compress_operators_device(c);
    c->sym_compressed = c->sym_uploaded = scheme == HFMM_M2L_SYMMETRIC;
    if (st == HFMM_OK) st = hfmm::upload_operators(c);  // rank-0 operators until compressed
    if (st == HFMM_OK && hfmm::is_mixed(c)) st = hfmm::prepare_mixed_operators(c);
    if (st == HFMM_OK && scheme != HFMM_M2L_SYMMETRIC) st = hfmm_set_m2l_scheme(h, scheme);
    if (st != HFMM_OK) {
        hfmm::free_all(c);
        hfmm::destroy_side(c);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete h;
        return st;
    }
    *out = h;
    return HFMM_OK;
}

// Kept cells: runs `pass`, which overwrites `count` doubles at `dst`, so that
// it adds to them instead (snapshot, pass, add the snapshot back).  Fresh
// cells: runs the pass.
template <class Pass>
static int accumulating(Context* c, double* dst, int64_t count, Pass pass) {
    if (!c->keep_cells) return pass();
    double* snap = c->d_keep.as<double>();
    HFMM_TRY(combine_device(c, snap, dst, count, false));
    HFMM_TRY(pass());
    return combine_device(c, dst, snap, count, true);
}

// The passes of one charge vector over the working levels [work_upper,
// work_lower) with the loop bounds of FGroupSeqAlgorithm
// (GroupTree/Core/FGroupSeqAlgorithm.hpp:88, 137, 238): M2M parents from
// min(h-2, lower-1) down to upper, M2L levels lower-1 .. upper, L2L parents
// upper .. lower-2.  Results into d_res_sorted.  With fresh cells every
// evaluation starts from zero expansions; with kept cells every pass adds
// into the expansions left by earlier evaluations, as the reference's P2M,
// M2M, M2L and L2L add into the tree's cells (FChebInterpolator.hpp:627-700,
// 560-582; FChebSymKernel.hpp:193-206, 210-328, 331-350).
static int evaluate_passes(Context* c, uint32_t ops, size_t exp_bytes) {
    const int L = c->height() - 1;
    const int up = c->work_upper, lo = c->work_lower;
    const int64_t LD = c->LD;
    double* M = c->d_M.as<double>();
    double* Lx = c->d_L.as<double>();
    const bool keep = c->keep_cells;
    const bool mixed = hfmm::is_mixed(c);
    // The P2P reads only the sorted particles and the leaf lists and writes
    // only the results (and its own local coordinates): it runs on the side
    // stream from here, and the L2P below waits for it.  Not when profiling
    // (per-kernel times need one kernel at a time).
    const bool side = c->side && c->overlap && !c->profiling && (ops & HFMM_P2P);
    if (side) {
        HFMM_HIP(c, hipEventRecord(c->ev_fork, c->stream));
        HFMM_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
        std::swap(c->stream, c->side);
        const int st = mixed ? hfmm::run_p2p_mixed(c) : hfmm::run_p2p(c);
        std::swap(c->stream, c->side);
        HFMM_TRY(st);
        HFMM_HIP(c, hipEventRecord(c->ev_join, c->side));
    }
    // FChebKernel's scheme: compressed expansions beside M and L
    const bool glob = c->m2l_scheme != HFMM_M2L_SYMMETRIC;
    const size_t cexp_bytes = sizeof(double) * static_cast<size_t>(c->total_cells) * c->glob.k;
    // upward pass (fresh cells: multipoles that no pass of this evaluation
    // writes read as zero)
    if (!keep && (ops & HFMM_P2M) && !(ops & HFMM_M2M)) HFMM_TRY(hfmm::zero_device(c, c->d_M.ptr, exp_bytes));
    if (glob && !keep) HFMM_TRY(hfmm::zero_device(c, c->d_Mt.ptr, cexp_bytes));
    if (ops & HFMM_P2M) {
        HFMM_TRY(accumulating(c, M + c->lvl_off[L] * LD, c->ncell[L] * LD, [&] { return hfmm::run_p2m(c); }));
        if (glob) HFMM_TRY(hfmm::run_global_compress(c, L));
    } else if (!keep)
        HFMM_TRY(hfmm::zero_device(c, c->d_M.ptr, exp_bytes));
    if (ops & HFMM_M2M)
        for (int l = std::min(L - 1, lo - 1); l >= up; --l) {
            HFMM_TRY(accumulating(c, M + c->lvl_off[l] * LD, c->ncell[l] * LD, [&] { return hfmm::run_m2m(c, l); }));
            if (glob) HFMM_TRY(hfmm::run_global_compress(c, l));
        }
    // transfer pass: M2L writes every local of its levels; the others start at zero
    const bool all_levels = up == c->top() && lo == c->height();
    if (glob) {  // the compressed M2L adds into L~; L gets U L~ before L2L / L2P
        if (!keep) {
            HFMM_TRY(hfmm::zero_device(c, c->d_L.ptr, exp_bytes));
            HFMM_TRY(hfmm::zero_device(c, c->d_Lt.ptr, cexp_bytes));
        }
        if (ops & HFMM_M2L) HFMM_TRY(hfmm::run_global_m2l(c));
    } else if (ops & HFMM_M2L) {
        HFMM_TRY(accumulating(c, Lx, static_cast<int64_t>(exp_bytes / sizeof(double)), [&] {
            if (keep || !all_levels) HFMM_TRY(hfmm::zero_device(c, c->d_L.ptr, exp_bytes));
            if (!mixed) return hfmm::run_m2l(c);
            HFMM_TRY(hfmm::run_split_multipoles(c));
            return hfmm::run_m2l_mixed(c);
        }));
    } else if (!keep)
        HFMM_TRY(hfmm::zero_device(c, c->d_L.ptr, exp_bytes));
    // downward pass (L2L adds into the children)
    if (ops & HFMM_L2L)
        for (int l = up; l < lo - 1; ++l) {
            if (glob) HFMM_TRY(hfmm::run_global_decompress(c, l));
            HFMM_TRY(hfmm::run_l2l(c, l));
        }
    // near field initialises the results, the merge adds the far field
    if (side) HFMM_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    else if (ops & HFMM_P2P) HFMM_TRY(mixed ? hfmm::run_p2p_mixed(c) : hfmm::run_p2p(c));
    else HFMM_TRY(hfmm::zero_results(c));
    if (ops & HFMM_L2P) {
        if (glob) HFMM_TRY(hfmm::run_global_decompress(c, L));
        HFMM_TRY(hfmm::run_l2p(c));
    }
    return HFMM_OK;
}

}  // namespace hfmm

extern "C" {

int hfmm_destroy(hfmm_ctx* h) {
    CTX_CHECK(h);
    Context* c = &h->c;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (auto& t : c->pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    hfmm::free_all(c);
    hfmm::destroy_side(c);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete h;
    return HFMM_OK;
}

const char* hfmm_last_error(const hfmm_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int hfmm_set_particles(hfmm_ctx* h, int64_t n, const double* x, const double* y, const double* z, const double* q) {
    return hfmm_set_particle_sets(h, 1, nullptr, 1, n, x, y, z, q);
}

int hfmm_set_particle_sets(hfmm_ctx* h, int32_t nsets, const int64_t* set_offsets, int32_t nrhs, int64_t n,
                           const double* x, const double* y, const double* z, const double* q) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (n <= 0 || !x || !y || !z || !q) return hfmm::fail(c, HFMM_ERR_ARG, "set_particles: bad arguments");
    if (nsets < 1 || (nsets > 1 && !set_offsets)) return hfmm::fail(c, HFMM_ERR_ARG, "set_particles: bad sets");
    if (nrhs < 1 || nrhs > HFMM_MAX_RHS) return hfmm::fail(c, HFMM_ERR_ARG, "set_particles: bad nrhs");
    if (set_offsets) {
        if (set_offsets[0] != 0 || set_offsets[nsets] != n)
            return hfmm::fail(c, HFMM_ERR_ARG, "set_particles: set offsets must run from 0 to n");
        for (int32_t s = 0; s < nsets; ++s)
            if (set_offsets[s + 1] < set_offsets[s])
                return hfmm::fail(c, HFMM_ERR_ARG, "set_particles: set offsets must not decrease");
    }
    // set levels: the smallest S with 8^S >= nsets; the virtual tree must fit the keys
    int S = 0;
    while ((int64_t(1) << (3 * S)) < nsets) ++S;
    if (c->P.tree_height + S > hfmm::kMaxLevels)
        return hfmm::fail(c, HFMM_ERR_ARG, "set_particles: too many sets for this tree height");
    HFMM_HIP(c, hipSetDevice(c->device));
    const size_t nv = static_cast<size_t>(3 + nrhs) * n;
    HFMM_TRY(hfmm::ensure(c, c->d_in, sizeof(double) * nv));
    double* d = c->d_in.as<double>();
    HFMM_HIP(c, hipMemcpyAsync(d, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HFMM_HIP(c, hipMemcpyAsync(d + n, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HFMM_HIP(c, hipMemcpyAsync(d + 2 * n, z, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HFMM_HIP(c, hipMemcpyAsync(d + 3 * n, q, sizeof(double) * n * nrhs, hipMemcpyHostToDevice, c->stream));
    if (nsets > 1) {
        HFMM_TRY(hfmm::ensure(c, c->d_setoff, sizeof(int64_t) * (nsets + 1)));
        HFMM_HIP(c, hipMemcpyAsync(c->d_setoff.ptr, set_offsets, sizeof(int64_t) * (nsets + 1), hipMemcpyHostToDevice,
                                   c->stream));
    }
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    c->n = n;
    c->x = d;
    c->y = d + n;
    c->z = d + 2 * n;
    c->q = d + 3 * n;
    c->nsets = nsets;
    c->set_levels = S;
    c->nrhs = nrhs;
    c->tree_built = false;
    c->evaluated = false;
    return HFMM_OK;
}

int hfmm_set_charges(hfmm_ctx* h, const double* q) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!q) return hfmm::fail(c, HFMM_ERR_ARG, "set_charges: bad arguments");
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "set_charges before build_tree");
    HFMM_HIP(c, hipSetDevice(c->device));
    // nrhs charge vectors (RHS-major) into a buffer of the context
    const size_t nq = static_cast<size_t>(c->n) * c->nrhs;
    HFMM_TRY(hfmm::ensure(c, c->d_qin, sizeof(double) * nq));
    HFMM_HIP(c, hipMemcpyAsync(c->d_qin.ptr, q, sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
    c->q = c->d_qin.as<double>();
    c->evaluated = false;
    return hfmm::update_charges_device(c);
}

int hfmm_set_charges_device(hfmm_ctx* h, const double* q) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!q) return hfmm::fail(c, HFMM_ERR_ARG, "set_charges_device: bad arguments");
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "set_charges before build_tree");
    HFMM_HIP(c, hipSetDevice(c->device));
    c->q = q;
    c->evaluated = false;
    return hfmm::update_charges_device(c);
}

int hfmm_set_particles_device(hfmm_ctx* h, int64_t n, const double* x, const double* y, const double* z,
                              const double* q) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (n <= 0 || !x || !y || !z || !q) return hfmm::fail(c, HFMM_ERR_ARG, "set_particles_device: bad arguments");
    c->nsets = 1;
    c->set_levels = 0;
    c->nrhs = 1;
    c->n = n;
    c->x = x;
    c->y = y;
    c->z = z;
    c->q = q;
    c->tree_built = false;
    c->evaluated = false;
    return HFMM_OK;
}

int hfmm_build_tree(hfmm_ctx* h) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (c->n <= 0) return hfmm::fail(c, HFMM_ERR_STATE, "build_tree before set_particles");
    HFMM_HIP(c, hipSetDevice(c->device));
    c->tree_built = false;
    c->cells_dirty = true;  // a new tree: cells in their initial state
    return hfmm::build_tree_device(c);
}

int hfmm_evaluate(hfmm_ctx* h, uint32_t ops) {
    CTX_CHECK(h);
    return hfmm_evaluate_levels(h, ops, 2, h->c.P.tree_height);
}

int hfmm_evaluate_levels(hfmm_ctx* h, uint32_t ops, int32_t upper, int32_t lower) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "evaluate before build_tree");
    if (ops & ~HFMM_NEAR_AND_FAR) return hfmm::fail(c, HFMM_ERR_ARG, "unknown operation bits");
    // validateLevels (FCoreCommon.hpp:49-52), levels of one set's tree
    if (upper < 2 || upper > lower || lower > c->P.tree_height)
        return hfmm::fail(c, HFMM_ERR_ARG, "working levels must satisfy 2 <= upper <= lower <= height");
    c->work_upper = upper + c->set_levels;
    c->work_lower = lower + c->set_levels;
    HFMM_HIP(c, hipSetDevice(c->device));
    const size_t exp_bytes = sizeof(double) * static_cast<size_t>(c->total_cells) * c->LD;
    // one extra zero multipole after the last cell: the M2L kernel reads it for missing sources
    HFMM_TRY(hfmm::ensure(c, c->d_M, exp_bytes + sizeof(double) * c->LD));
    HFMM_TRY(hfmm::ensure(c, c->d_L, exp_bytes));
    // several charge vectors: the near field of the leaves of at most kP2PBig
    // particles runs once for all of them (the pair geometry once per pair),
    // and the sorted results hold one block per vector
    const bool near_all = c->nrhs > 1 && (ops & HFMM_P2P);
    HFMM_TRY(hfmm::ensure(c, c->d_res_sorted, sizeof(double) * 4 * c->n * (near_all ? c->nrhs : 1)));
    HFMM_TRY(hfmm::ensure(c, c->d_res, sizeof(double) * 4 * c->n * c->nrhs));
    HFMM_TRY(hfmm::ensure(c, c->d_field, sizeof(double) * 3 * c->n * c->nrhs));
    HFMM_TRY(hfmm::zero_device(c, c->d_M.as<double>() + static_cast<size_t>(c->total_cells) * c->LD,
                               sizeof(double) * c->LD));
    // kept cells of several charge vectors: the cells of vector r are stored at
    // d_kept + 2 r exp (M, then L) between evaluations
    const int64_t exp_words = static_cast<int64_t>(exp_bytes / sizeof(double));
    const bool kept_many = c->keep_cells && c->nrhs > 1;
    const bool glob = c->m2l_scheme != HFMM_M2L_SYMMETRIC;
    if (glob) {
        const size_t cexp = sizeof(double) * static_cast<size_t>(c->total_cells) * c->glob.k;
        HFMM_TRY(hfmm::ensure(c, c->d_Mt, cexp));
        HFMM_TRY(hfmm::ensure(c, c->d_Lt, cexp));
        if (c->keep_cells && c->nrhs > 1)
            return hfmm::fail(c, HFMM_ERR_STATE, "kept cells of the global M2L scheme hold one charge vector");
        if (c->keep_cells && c->cells_dirty) {
            HFMM_TRY(hfmm::zero_device(c, c->d_Mt.ptr, cexp));
            HFMM_TRY(hfmm::zero_device(c, c->d_Lt.ptr, cexp));
        }
    }
    if (c->keep_cells) HFMM_TRY(hfmm::prepare_kept_cells(c));
    struct PerVector {  // the single-vector state again on every way out
        Context* c;
        ~PerVector() {
            c->res_rhs = 0;
            c->p2p_small_done = false;
        }
    } per_vector{c};
    if (near_all)
        HFMM_TRY(hfmm::is_mixed(c) ? hfmm::run_p2p_mixed_small_all(c) : hfmm::run_p2p_small_all(c));
    // one pass sequence per charge vector (the tree and the lists are shared);
    // with one vector the sorted charges are already in place
    for (int r = 0; r < c->nrhs; ++r) {
        c->res_rhs = near_all ? r : 0;
        c->p2p_small_done = near_all;
        double* kept = kept_many ? c->d_kept.as<double>() + 2 * exp_words * r : nullptr;
        if (c->nrhs > 1) HFMM_TRY(hfmm::update_charges_device(c, r));
        if (kept) {
            HFMM_TRY(hfmm::combine_device(c, c->d_M.as<double>(), kept, exp_words, false));
            HFMM_TRY(hfmm::combine_device(c, c->d_L.as<double>(), kept + exp_words, exp_words, false));
        }
        HFMM_TRY(hfmm::evaluate_passes(c, ops, exp_bytes));
        if (kept) {
            HFMM_TRY(hfmm::combine_device(c, kept, c->d_M.as<double>(), exp_words, false));
            HFMM_TRY(hfmm::combine_device(c, kept + exp_words, c->d_L.as<double>(), exp_words, false));
        }
        HFMM_TRY(hfmm::scatter_results(c, r));
    }
    c->evaluated = true;
    return HFMM_OK;
}

int hfmm_synchronize(hfmm_ctx* h) {
    CTX_CHECK(h);
    Context* c = &h->c;
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    return HFMM_OK;
}

// Copies nrhs blocks of n values per output (RHS-major) from a device array
// laid out as [rhs][k][n].
static int copy_outputs(Context* c, const double* src, int nk, double* const* outs) {
    const size_t n = static_cast<size_t>(c->n);
    for (int r = 0; r < c->nrhs; ++r)
        for (int k = 0; k < nk; ++k)
            if (outs[k])
                HFMM_HIP(c, hipMemcpyAsync(outs[k] + r * n, src + (static_cast<size_t>(r) * nk + k) * n,
                                           sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    return HFMM_OK;
}

int hfmm_get_results(hfmm_ctx* h, double* pot, double* fx, double* fy, double* fz) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->evaluated) return hfmm::fail(c, HFMM_ERR_STATE, "get_results before evaluate");
    double* outs[4] = {pot, fx, fy, fz};
    return copy_outputs(c, c->d_res.as<double>(), 4, outs);
}

int hfmm_get_field(hfmm_ctx* h, double* ex, double* ey, double* ez) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->evaluated) return hfmm::fail(c, HFMM_ERR_STATE, "get_field before evaluate");
    double* outs[3] = {ex, ey, ez};
    return copy_outputs(c, c->d_field.as<double>(), 3, outs);
}

int hfmm_get_field_device(hfmm_ctx* h, const double** ex, const double** ey, const double** ez) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->evaluated) return hfmm::fail(c, HFMM_ERR_STATE, "get_field before evaluate");
    const double* f = c->d_field.as<double>();
    if (ex) *ex = f;
    if (ey) *ey = f + c->n;
    if (ez) *ez = f + 2 * c->n;
    return HFMM_OK;
}

int hfmm_get_results_device(hfmm_ctx* h, const double** pot, const double** fx, const double** fy, const double** fz) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->evaluated) return hfmm::fail(c, HFMM_ERR_STATE, "get_results before evaluate");
    const double* r = c->d_res.as<double>();
    if (pot) *pot = r;
    if (fx) *fx = r + c->n;
    if (fy) *fy = r + 2 * c->n;
    if (fz) *fz = r + 3 * c->n;
    return HFMM_OK;
}

int hfmm_get_tree_info(hfmm_ctx* h, hfmm_tree_info* info) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!info) return HFMM_ERR_ARG;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    std::memset(info, 0, sizeof(*info));
    info->n_particles = c->n;
    info->height = c->height();
    info->order = c->order();
    const int L = c->height() - 1;
    for (int l = 2; l <= L; ++l) info->n_cells[l] = c->ncell[l];
    // list sizes: count the non-empty slots on the host
    std::vector<int32_t> s(static_cast<size_t>(c->ncell[L]) * hfmm::kNbrSlots);
    HFMM_HIP(c, hipMemcpy(s.data(), c->d_p2p.ptr, sizeof(int32_t) * s.size(), hipMemcpyDeviceToHost));
    info->p2p_entries = std::count_if(s.begin(), s.end(), [](int32_t v) { return v >= 0; });
    for (int l = c->top(); l <= L; ++l) {
        int64_t counts[343];
        HFMM_TRY(hfmm_get_m2l_counts(h, l, counts));
        for (int pos = 0; pos < 343; ++pos) info->m2l_entries[l] += counts[pos];
    }
    return HFMM_OK;
}

int hfmm_get_sorted_keys(hfmm_ctx* h, int64_t* keys, int64_t* perm) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    if (keys) HFMM_HIP(c, hipMemcpy(keys, c->d_keys.ptr, sizeof(int64_t) * c->n, hipMemcpyDeviceToHost));
    if (perm) {
        std::vector<int32_t> p(c->n);
        HFMM_HIP(c, hipMemcpy(p.data(), c->d_perm.ptr, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < c->n; ++i) perm[i] = p[i];
    }
    return HFMM_OK;
}

int hfmm_get_cell_keys(hfmm_ctx* h, int32_t level, int64_t* keys) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    if (level < 2 || level >= c->height() || !keys) return hfmm::fail(c, HFMM_ERR_ARG, "bad level");
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    HFMM_HIP(c, hipMemcpy(keys, c->d_cellkey.as<uint64_t>() + c->lvl_off[level], sizeof(int64_t) * c->ncell[level],
                          hipMemcpyDeviceToHost));
    return HFMM_OK;
}

int hfmm_get_leaf_ranges(hfmm_ctx* h, int64_t* start, int64_t* count) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    const int64_t nl = c->ncell[c->height() - 1];
    std::vector<int32_t> s(nl + 1);
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    HFMM_HIP(c, hipMemcpy(s.data(), c->d_leafstart.ptr, sizeof(int32_t) * (nl + 1), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nl; ++i) {
        if (start) start[i] = s[i];
        if (count) count[i] = s[i + 1] - s[i];
    }
    return HFMM_OK;
}

int hfmm_get_p2p_list(hfmm_ctx* h, int64_t* offsets, int64_t* nbr, int32_t* pos) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    const int64_t nl = c->ncell[c->height() - 1];
    std::vector<int32_t> s(static_cast<size_t>(nl) * hfmm::kNbrSlots);
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    HFMM_HIP(c, hipMemcpy(s.data(), c->d_p2p.ptr, sizeof(int32_t) * s.size(), hipMemcpyDeviceToHost));
    int64_t e = 0;
    if (offsets) offsets[0] = 0;
    for (int64_t i = 0; i < nl; ++i) {
        for (int j = 0; j < hfmm::kNbrSlots; ++j) {
            const int32_t v = s[static_cast<size_t>(i) * hfmm::kNbrSlots + j];
            if (v < 0) continue;
            if (nbr) nbr[e] = v;
            if (pos) pos[e] = j < 13 ? j : j + 1;  // slots skip the centre position 13
            ++e;
        }
        if (offsets) offsets[i + 1] = e;
    }
    return HFMM_OK;
}

int hfmm_get_m2l_list(hfmm_ctx* h, int32_t level, int64_t* offsets, int64_t* src, int32_t* pos) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    if (level < 2 || level >= c->height()) return hfmm::fail(c, HFMM_ERR_ARG, "bad level");
    const int64_t nc = c->ncell[level];
    std::vector<int32_t> s(static_cast<size_t>(nc) * hfmm::kFarSlots);
    std::vector<int64_t> keys(nc);
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    HFMM_HIP(c, hipMemcpy(s.data(), c->d_m2l.as<int32_t>() + static_cast<size_t>(c->m2l_off[level]) * hfmm::kFarSlots,
                          sizeof(int32_t) * s.size(), hipMemcpyDeviceToHost));
    HFMM_HIP(c, hipMemcpy(keys.data(), c->d_cellkey.as<uint64_t>() + c->lvl_off[level], sizeof(int64_t) * nc,
                          hipMemcpyDeviceToHost));
    hfmm::FarSlot table[8][hfmm::kFarSlots];
    for (int e = 0; e < 8; ++e) hfmm::far_slots(e, table[e]);
    int64_t e = 0;
    if (offsets) offsets[0] = 0;
    for (int64_t i = 0; i < nc; ++i) {
        const int par = static_cast<int>(keys[i] & 7);
        for (int j = 0; j < hfmm::kFarSlots; ++j) {
            const int32_t v = s[static_cast<size_t>(i) * hfmm::kFarSlots + j];
            if (v < 0) continue;
            if (src) src[e] = v;
            if (pos) pos[e] = table[par][j].pos;
            ++e;
        }
        if (offsets) offsets[i + 1] = e;
    }
    return HFMM_OK;
}

int hfmm_get_m2l_counts(hfmm_ctx* h, int32_t level, int64_t* counts) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->tree_built) return hfmm::fail(c, HFMM_ERR_STATE, "tree not built");
    if (level < 2 || level >= c->height() || !counts) return hfmm::fail(c, HFMM_ERR_ARG, "bad level");
    std::fill(counts, counts + 343, int64_t(0));
    if (level < c->top()) return HFMM_OK;
    HFMM_HIP(c, hipSetDevice(c->device));
    hfmm::DevBuf b;
    HFMM_TRY(hfmm::ensure(c, b, sizeof(unsigned long long) * 8 * hfmm::kFarSlots));
    std::vector<unsigned long long> hc(8 * hfmm::kFarSlots);
    int st = hfmm::m2l_slot_counts(c, level, b.as<unsigned long long>());
    if (st == HFMM_OK && (hipMemcpyAsync(hc.data(), b.ptr, sizeof(unsigned long long) * hc.size(),
                                         hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                          hipStreamSynchronize(c->stream) != hipSuccess))
        st = hfmm::fail(c, HFMM_ERR_DEVICE, "m2l counts copy failed");
    hfmm::release(b);
    if (st != HFMM_OK) return st;
    for (int e = 0; e < 8; ++e) {
        hfmm::FarSlot table[hfmm::kFarSlots];
        hfmm::far_slots(e, table);
        for (int j = 0; j < hfmm::kFarSlots; ++j) counts[table[j].pos] += static_cast<int64_t>(hc[e * hfmm::kFarSlots + j]);
    }
    return HFMM_OK;
}

// The 16 symmetric operators on the host, compressed on first need by a
// context created with another M2L scheme.
static int ensure_symmetric_compressed(Context* c) {
    if (c->sym_compressed) return HFMM_OK;
    HFMM_HIP(c, hipSetDevice(c->device));
    HFMM_TRY(hfmm::compress_operators_device(c));
    c->sym_compressed = true;
    return HFMM_OK;
}

int hfmm_save_operators(hfmm_ctx* h, const char* path) {
    CTX_CHECK(h);
    if (!path) return hfmm::fail(&h->c, HFMM_ERR_ARG, "save_operators: no path");
    HFMM_TRY(ensure_symmetric_compressed(&h->c));
    return hfmm::save_operators(&h->c, path);
}

int hfmm_load_operators(hfmm_ctx* h, const char* path) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!path) return hfmm::fail(c, HFMM_ERR_ARG, "load_operators: no path");
    std::array<int, 343> rank;
    std::array<std::vector<double>, 343> K;
    HFMM_TRY(hfmm::read_operators(c, path, rank, K));
    for (int idx = 0; idx < 343; ++idx)
        if (rank[idx] > 64) return hfmm::fail(c, HFMM_ERR_ARG, "operator rank above 64 is not supported");
    HFMM_HIP(c, hipSetDevice(c->device));
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    c->ops.rank = rank;
    c->ops.K = std::move(K);
    c->sym_compressed = true;
    c->sym_uploaded = false;
    if (c->m2l_scheme == HFMM_M2L_SYMMETRIC) {  // otherwise uploaded on a switch to it
        HFMM_TRY(hfmm::upload_operators(c));
        c->sym_uploaded = true;
    }
    if (hfmm::is_mixed(c)) HFMM_TRY(hfmm::prepare_mixed_operators(c));
    c->m2l_variant = -1;  // other operators: tune again
    c->evaluated = false;
    return HFMM_OK;
}

int hfmm_get_operator(hfmm_ctx* h, int32_t idx, int32_t* rank, double* K) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (idx < 0 || idx >= 343) return hfmm::fail(c, HFMM_ERR_ARG, "bad operator index");
    HFMM_TRY(ensure_symmetric_compressed(c));
    const int k = c->ops.rank[idx];
    if (rank) *rank = k;
    if (K && k) std::memcpy(K, c->ops.K[idx].data(), sizeof(double) * c->ops.K[idx].size());
    return HFMM_OK;
}

int hfmm_get_expansions(hfmm_ctx* h, int32_t level, double* M, double* L) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (!c->evaluated) return hfmm::fail(c, HFMM_ERR_STATE, "get_expansions before evaluate");
    if (level < 2 || level >= c->height()) return hfmm::fail(c, HFMM_ERR_ARG, "bad level");
    const int64_t nc = c->ncell[level];
    const int nn = c->ops.nn;
    std::vector<double> tmp(static_cast<size_t>(nc) * c->LD);
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    for (int w = 0; w < 2; ++w) {
        double* out = w == 0 ? M : L;
        if (!out) continue;
        if (c->keep_cells && c->cells_dirty) {  // reset, not yet evaluated: the initial state
            std::memset(out, 0, sizeof(double) * static_cast<size_t>(nc) * nn);
            continue;
        }
        const hfmm::DevBuf& b = w == 0 ? c->d_M : c->d_L;
        HFMM_HIP(c, hipMemcpy(tmp.data(), b.as<double>() + static_cast<size_t>(c->lvl_off[level]) * c->LD,
                              sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nc; ++i)
            std::memcpy(out + static_cast<size_t>(i) * nn, tmp.data() + static_cast<size_t>(i) * c->LD,
                        sizeof(double) * nn);
    }
    return HFMM_OK;
}

int hfmm_set_m2l_scheme(hfmm_ctx* h, int32_t scheme) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (scheme != HFMM_M2L_SYMMETRIC && scheme != HFMM_M2L_GLOBAL_SVD && scheme != HFMM_M2L_DENSE &&
        scheme != HFMM_M2L_UNIFORM)
        return hfmm::fail(c, HFMM_ERR_ARG, "unknown M2L scheme");
    HFMM_HIP(c, hipSetDevice(c->device));
    if (scheme != HFMM_M2L_SYMMETRIC) {
        if (c->P.precision != HFMM_PRECISION_FP64)
            return hfmm::fail(c, HFMM_ERR_ARG, "the global-SVD, dense and uniform M2L schemes run in FP64 only");
        if (c->glob.k == 0 || scheme != c->m2l_scheme) {
            std::string err;
            hfmm::GlobalOperators g;
            hfmm::ChebOperators u;  // the uniform basis: equispaced roots, Lagrange child tables
            int st = HFMM_OK;
            if (scheme == HFMM_M2L_UNIFORM) {
                u.p = c->ops.p;
                u.nn = c->ops.nn;
                u.child.assign(static_cast<size_t>(8) * 3 * u.p * u.p, 0.);
                if (hfmm::make_uniform_tables(u) != HFMM_OK)
                    return hfmm::fail(c, HFMM_ERR_ARG, "no equispaced roots for this order");
                st = hfmm::build_uniform_fft_operators(u, g);
            } else {
                st = scheme == HFMM_M2L_DENSE ? hfmm::build_dense_operators(c->ops, g)
                                              : hfmm::build_global_operators(c->ops, g, err);
            }
            if (st != HFMM_OK) return hfmm::fail(c, st, err);
            if (scheme == HFMM_M2L_UNIFORM) {
                HFMM_TRY(hfmm::ensure(c, c->d_uK, sizeof(double) * g.Khat.size()));
                HFMM_TRY(hfmm::ensure(c, c->d_utw, sizeof(double) * g.tw.size()));
                HFMM_HIP(c, hipMemcpy(c->d_uK.ptr, g.Khat.data(), sizeof(double) * g.Khat.size(), hipMemcpyHostToDevice));
                HFMM_HIP(c, hipMemcpy(c->d_utw.ptr, g.tw.data(), sizeof(double) * g.tw.size(), hipMemcpyHostToDevice));
            } else {
                if (g.k > 512) return hfmm::fail(c, HFMM_ERR_ARG, "global M2L rank above 512");
                const int nn = c->ops.nn, k = g.k;
                std::vector<double> bt(static_cast<size_t>(nn) * k);
                for (int a = 0; a < k; ++a)
                    for (int n = 0; n < nn; ++n) bt[static_cast<size_t>(n) * k + a] = g.B[static_cast<size_t>(a) * nn + n];
                HFMM_TRY(hfmm::ensure(c, c->d_gBt, sizeof(double) * bt.size()));
                HFMM_TRY(hfmm::ensure(c, c->d_gU, sizeof(double) * g.U.size()));
                HFMM_TRY(hfmm::ensure(c, c->d_gC, sizeof(double) * g.C.size()));
                HFMM_HIP(c, hipMemcpy(c->d_gBt.ptr, bt.data(), sizeof(double) * bt.size(), hipMemcpyHostToDevice));
                HFMM_HIP(c, hipMemcpy(c->d_gU.ptr, g.U.data(), sizeof(double) * g.U.size(), hipMemcpyHostToDevice));
                HFMM_HIP(c, hipMemcpy(c->d_gC.ptr, g.C.data(), sizeof(double) * g.C.size(), hipMemcpyHostToDevice));
            }
            c->glob = std::move(g);
            c->gcr_kp = 0;  // the MFMA M2L's padded operators belong to the old set
            // the interpolation tables of the basis the scheme works in
            const hfmm::ChebOperators& tab = scheme == HFMM_M2L_UNIFORM ? u : c->ops;
            HFMM_HIP(c, hipMemcpy(c->d_child.ptr, tab.child.data(), sizeof(double) * tab.child.size(),
                                  hipMemcpyHostToDevice));
            if (scheme == HFMM_M2L_UNIFORM) {
                HFMM_TRY(hfmm::ensure(c, c->d_uroots, sizeof(double) * u.p));
                HFMM_HIP(c, hipMemcpy(c->d_uroots.ptr, u.roots.data(), sizeof(double) * u.p, hipMemcpyHostToDevice));
            }
            c->basis = scheme == HFMM_M2L_UNIFORM ? 1 : 0;
        }
    } else {
        HFMM_TRY(ensure_symmetric_compressed(c));
        if (!c->sym_uploaded) {  // also restores the Chebyshev tables
            HFMM_HIP(c, hipStreamSynchronize(c->stream));
            HFMM_TRY(hfmm::upload_operators(c));
            c->sym_uploaded = true;
        } else if (c->basis != 0) {
            HFMM_HIP(c, hipMemcpy(c->d_child.ptr, c->ops.child.data(), sizeof(double) * c->ops.child.size(),
                                  hipMemcpyHostToDevice));
        }
        c->basis = 0;
    }
    if (scheme != c->m2l_scheme) c->cells_dirty = true;
    c->m2l_scheme = scheme;
    c->evaluated = false;
    return HFMM_OK;
}

int hfmm_get_global_operators(hfmm_ctx* h, int32_t* rank, double* U, double* B, double* C, double* S1, double* S2) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (c->glob.k == 0 || c->m2l_scheme == HFMM_M2L_UNIFORM)
        return hfmm::fail(c, HFMM_ERR_STATE, "no global-SVD or dense M2L operators are built");
    if (rank) *rank = c->glob.k;
    if (U) std::memcpy(U, c->glob.U.data(), sizeof(double) * c->glob.U.size());
    if (B) std::memcpy(B, c->glob.B.data(), sizeof(double) * c->glob.B.size());
    if (C) std::memcpy(C, c->glob.C.data(), sizeof(double) * c->glob.C.size());
    if (S1) std::memcpy(S1, c->glob.S1.data(), sizeof(double) * c->glob.S1.size());
    if (S2) std::memcpy(S2, c->glob.S2.data(), sizeof(double) * c->glob.S2.size());
    return HFMM_OK;
}

int hfmm_set_stream_overlap(hfmm_ctx* h, int32_t enable) {
    CTX_CHECK(h);
    h->c.overlap = enable != 0;
    return HFMM_OK;
}

int hfmm_set_p2p_mutual(hfmm_ctx* h, int32_t enable) {
    CTX_CHECK(h);
    h->c.p2p_mutual_on = enable != 0;
    return HFMM_OK;
}

int hfmm_set_cell_mode(hfmm_ctx* h, int32_t mode) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (mode != HFMM_CELLS_FRESH && mode != HFMM_CELLS_KEPT) return hfmm::fail(c, HFMM_ERR_ARG, "unknown cell mode");
    const bool keep = mode == HFMM_CELLS_KEPT;
    if (keep != c->keep_cells) c->cells_dirty = true;  // kept cells start from their initial state
    c->keep_cells = keep;
    return HFMM_OK;
}

int hfmm_reset_cells(hfmm_ctx* h) {
    CTX_CHECK(h);
    h->c.cells_dirty = true;
    return HFMM_OK;
}

int hfmm_set_m2l_variant(hfmm_ctx* h, int32_t variant) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (variant != -1 && !hfmm::m2l_variant_available(c, variant))
        return hfmm::fail(c, HFMM_ERR_ARG, "M2L variant not available for this order and precision");
    c->m2l_variant = variant;
    return HFMM_OK;
}

int hfmm_get_m2l_variant(hfmm_ctx* h, int32_t* variant, float* tune_ms) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (variant) *variant = c->m2l_variant;
    if (tune_ms)
        for (int v = 0; v < 4; ++v) tune_ms[v] = c->m2l_tune_ms[v];
    return HFMM_OK;
}

int hfmm_set_profiling(hfmm_ctx* h, int32_t enable) {
    CTX_CHECK(h);
    h->c.profiling = enable != 0;
    return HFMM_OK;
}

int hfmm_set_profiling_filter(hfmm_ctx* h, const char* names) {
    CTX_CHECK(h);
    Context* c = &h->c;
    c->prof_filter.clear();
    if (names && *names) c->prof_filter = "," + std::string(names) + ",";
    return HFMM_OK;
}

int hfmm_get_kernel_times(hfmm_ctx* h, hfmm_kernel_time* out, int32_t max, int32_t* count) {
    CTX_CHECK(h);
    Context* c = &h->c;
    HFMM_TRY(hfmm::collect_times(c));
    const int32_t m = static_cast<int32_t>(c->totals.size());
    if (count) *count = m;
    for (int32_t i = 0; i < std::min(m, max); ++i)
        if (out) out[i] = c->totals[i];
    return HFMM_OK;
}

int hfmm_reset_kernel_times(hfmm_ctx* h) {
    CTX_CHECK(h);
    Context* c = &h->c;
    HFMM_TRY(hfmm::collect_times(c));
    c->totals.clear();
    return HFMM_OK;
}

int hfmm_selftest_mfma_f64(hfmm_ctx* h, double* out) {
    CTX_CHECK(h);
    if (!out) return HFMM_ERR_ARG;
    return hfmm::selftest_mfma_f64(&h->c, out);
}

int hfmm_direct(hfmm_ctx* h, int64_t ntargets, const int64_t* targets, double* pot, double* fx, double* fy,
                double* fz) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (c->n <= 0) return hfmm::fail(c, HFMM_ERR_STATE, "direct before set_particles");
    if (c->nsets > 1) return hfmm::fail(c, HFMM_ERR_ARG, "direct: the context holds several particle sets");
    if (ntargets <= 0 || !targets) return hfmm::fail(c, HFMM_ERR_ARG, "direct: bad targets");
    for (int64_t i = 0; i < ntargets; ++i)
        if (targets[i] < 0 || targets[i] >= c->n) return hfmm::fail(c, HFMM_ERR_ARG, "direct: target out of range");
    HFMM_HIP(c, hipSetDevice(c->device));
    return hfmm::direct_sum(c, ntargets, targets, pot, fx, fy, fz);
}

int hfmm_direct_noalloc(hfmm_ctx* h, int64_t ntargets, const int64_t* targets, double* pot, double* fx, double* fy,
                        double* fz) {
    CTX_CHECK(h);
    Context* c = &h->c;
    if (c->n <= 0) return hfmm::fail(c, HFMM_ERR_STATE, "direct before set_particles");
    if (c->nsets > 1) return hfmm::fail(c, HFMM_ERR_ARG, "direct: the context holds several particle sets");
    if (ntargets <= 0 || !targets) return hfmm::fail(c, HFMM_ERR_ARG, "direct: bad targets");
    for (int64_t i = 0; i < ntargets; ++i)
        if (targets[i] < 0 || targets[i] >= c->n) return hfmm::fail(c, HFMM_ERR_ARG, "direct: target out of range");
    HFMM_HIP(c, hipSetDevice(c->device));
    return hfmm::direct_sum_noalloc(c, ntargets, targets, pot, fx, fy, fz);
}

int hfmm_selftest_mfma_bf16(hfmm_ctx* h, float* out) {
    CTX_CHECK(h);
    if (!out) return HFMM_ERR_ARG;
    return hfmm::selftest_mfma_bf16(&h->c, out);
}

}  // extern "C"
