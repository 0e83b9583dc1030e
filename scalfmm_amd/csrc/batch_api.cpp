// batch_api.cpp -- hfmm_batch_*: many independent particle sets of uneven
// sizes, each with several charge vectors, evaluated by one HIP graph launch.
//
// Sets are grouped by tree height; each group is one context holding its sets
// as a forest (hfmm_set_particle_sets), so every pass of a group is one launch
// over all of its sets.  The groups share one stream; the pass sequences of
// every group and every charge vector are captured into one graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "hfmm_internal.hpp"

namespace {

struct Group {
    int32_t height = 0;
    std::vector<int32_t> sets;  // set indices, ascending
    std::vector<int64_t> off;   // group-local offsets of its sets (sets.size() + 1)
    hfmm_ctx* ctx = nullptr;
};

}  // namespace

struct hfmm_batch {
    hfmm_params P{};
    int32_t nsets = 0;
    int32_t nrhs = 1;
    int64_t n = 0;
    std::vector<int64_t> off;     // nsets + 1
    std::vector<int32_t> height;  // per set
    hipStream_t stream = nullptr;
    std::vector<Group> groups;    // groups with at least one particle
    bool loaded = false, built = false, evaluated = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_ops = 0;
    std::string err;
};

namespace {

int bfail(hfmm_batch* b, int code, const std::string& msg) {
    b->err = msg;
    return code;
}

void drop_graph(hfmm_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Enqueues the passes of every group (every charge vector).
int enqueue_all(hfmm_batch* b, uint32_t ops) {
    for (Group& g : b->groups) {
        const int st = hfmm_evaluate(g.ctx, ops);
        if (st != HFMM_OK)
            return bfail(b, st, "height " + std::to_string(g.height) + ": " + hfmm_last_error(g.ctx));
    }
    return HFMM_OK;
}

}  // namespace

extern "C" {

int32_t hfmm_batch_default_height(int64_t n) {
    int32_t h = 3;
    while (h < 20 && n > (int64_t(32) << (3 * (h - 1)))) ++h;
    return h;
}

int hfmm_batch_create(const hfmm_params* params, int32_t nsets, const int64_t* set_offsets,
                      const int32_t* set_heights, int32_t nrhs, hfmm_batch** out) {
    if (!params || !out || !set_offsets || nsets < 1) return HFMM_ERR_ARG;
    *out = nullptr;
    if (nrhs < 1 || nrhs > HFMM_MAX_RHS) return HFMM_ERR_ARG;
    if (set_offsets[0] != 0) return HFMM_ERR_ARG;
    for (int32_t s = 0; s < nsets; ++s) {
        if (set_offsets[s + 1] < set_offsets[s]) return HFMM_ERR_ARG;
        if (set_heights && (set_heights[s] < 3 || set_heights[s] > 20)) return HFMM_ERR_ARG;
    }
    if (set_offsets[nsets] <= 0) return HFMM_ERR_ARG;
    if (params->m2l_scheme != HFMM_M2L_SYMMETRIC) return HFMM_ERR_ARG;  // batches run the symmetric scheme
    hfmm_batch* b = new (std::nothrow) hfmm_batch();
    if (!b) return HFMM_ERR_NOMEM;
    b->P = *params;
    b->nsets = nsets;
    b->nrhs = nrhs;
    b->n = set_offsets[nsets];
    b->off.assign(set_offsets, set_offsets + nsets + 1);
    b->height.resize(nsets);
    for (int32_t s = 0; s < nsets; ++s)
        b->height[s] = set_heights ? set_heights[s] : hfmm_batch_default_height(set_offsets[s + 1] - set_offsets[s]);
    if (hipSetDevice(params->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        return HFMM_ERR_DEVICE;
    }
    // one group per height with particles, in increasing height
    std::map<int32_t, Group> by_h;
    for (int32_t s = 0; s < nsets; ++s) {
        const int64_t m = b->off[s + 1] - b->off[s];
        if (m == 0) continue;
        Group& g = by_h[b->height[s]];
        g.height = b->height[s];
        if (g.off.empty()) g.off.push_back(0);
        g.sets.push_back(s);
        g.off.push_back(g.off.back() + m);
    }
    int st = HFMM_OK;
    for (auto& kv : by_h) {
        Group g = kv.second;
        hfmm_params p = b->P;
        p.tree_height = g.height;
        st = hfmm::create_context(&p, b->stream, &g.ctx);
        if (st != HFMM_OK) break;
        b->groups.push_back(g);
    }
    if (st != HFMM_OK) {
        hfmm_batch_destroy(b);
        return st;
    }
    *out = b;
    return HFMM_OK;
}

int hfmm_batch_destroy(hfmm_batch* b) {
    if (!b) return HFMM_ERR_ARG;
    (void)hipSetDevice(b->P.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (Group& g : b->groups)
        if (g.ctx) hfmm_destroy(g.ctx);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return HFMM_OK;
}

const char* hfmm_batch_last_error(const hfmm_batch* b) { return b ? b->err.c_str() : "null batch"; }

int hfmm_batch_set_particles(hfmm_batch* b, const double* x, const double* y, const double* z, const double* q) {
    if (!b) return HFMM_ERR_ARG;
    if (!x || !y || !z || !q) return bfail(b, HFMM_ERR_ARG, "set_particles: bad arguments");
    drop_graph(b);
    b->loaded = b->built = b->evaluated = false;
    const size_t N = static_cast<size_t>(b->n);
    for (Group& g : b->groups) {
        // gather the group's sets into contiguous host arrays
        const int64_t m = g.off.back();
        std::vector<double> gx(m), gy(m), gz(m), gq(static_cast<size_t>(m) * b->nrhs);
        for (size_t i = 0; i < g.sets.size(); ++i) {
            const int32_t s = g.sets[i];
            const size_t src = static_cast<size_t>(b->off[s]), cnt = static_cast<size_t>(g.off[i + 1] - g.off[i]);
            const size_t dst = static_cast<size_t>(g.off[i]);
            std::memcpy(gx.data() + dst, x + src, sizeof(double) * cnt);
            std::memcpy(gy.data() + dst, y + src, sizeof(double) * cnt);
            std::memcpy(gz.data() + dst, z + src, sizeof(double) * cnt);
            for (int32_t r = 0; r < b->nrhs; ++r)
                std::memcpy(gq.data() + r * static_cast<size_t>(m) + dst, q + r * N + src, sizeof(double) * cnt);
        }
        const int st = hfmm_set_particle_sets(g.ctx, static_cast<int32_t>(g.sets.size()), g.off.data(), b->nrhs, m,
                                              gx.data(), gy.data(), gz.data(), gq.data());
        if (st != HFMM_OK)
            return bfail(b, st, "height " + std::to_string(g.height) + ": " + hfmm_last_error(g.ctx));
    }
    b->loaded = true;
    return HFMM_OK;
}

int hfmm_batch_build(hfmm_batch* b) {
    if (!b) return HFMM_ERR_ARG;
    if (!b->loaded) return bfail(b, HFMM_ERR_STATE, "build before set_particles");
    drop_graph(b);
    b->built = b->evaluated = false;
    for (Group& g : b->groups) {
        const int st = hfmm_build_tree(g.ctx);
        if (st != HFMM_OK)
            return bfail(b, st, "height " + std::to_string(g.height) + ": " + hfmm_last_error(g.ctx));
    }
    b->built = true;
    return HFMM_OK;
}

int hfmm_batch_evaluate(hfmm_batch* b, uint32_t ops) {
    if (!b) return HFMM_ERR_ARG;
    if (!b->built) return bfail(b, HFMM_ERR_STATE, "evaluate before build");
    if (ops & ~HFMM_NEAR_AND_FAR) return bfail(b, HFMM_ERR_ARG, "unknown operation bits");
    if (hipSetDevice(b->P.device) != hipSuccess) return bfail(b, HFMM_ERR_DEVICE, "hipSetDevice failed");
    // profiling brackets launches with events: such runs stay eager
    bool profiling = false;
    for (const Group& g : b->groups) profiling = profiling || hfmm::context_of(g.ctx)->profiling;
    if (profiling) {
        const int st = enqueue_all(b, ops);
        if (st == HFMM_OK) b->evaluated = true;
        return st;
    }
    if (b->exec && b->graph_ops == ops) {
        const hipError_t e = hipGraphLaunch(b->exec, b->stream);
        if (e != hipSuccess) return bfail(b, HFMM_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        b->evaluated = true;
        return HFMM_OK;
    }
    drop_graph(b);
    // eager run: sizes every device buffer (no allocation may happen while
    // capturing) and produces this call's results
    int st = enqueue_all(b, ops);
    if (st != HFMM_OK) return st;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(b, HFMM_ERR_DEVICE, std::string("eager run: ") + hipGetErrorString(e));
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return bfail(b, HFMM_ERR_DEVICE, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    st = enqueue_all(b, ops);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->stream, &g);
    if (st != HFMM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return bfail(b, HFMM_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return bfail(b, HFMM_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    b->graph_ops = ops;
    b->evaluated = true;
    return HFMM_OK;
}

int hfmm_batch_synchronize(hfmm_batch* b) {
    if (!b) return HFMM_ERR_ARG;
    const hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(b, HFMM_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return HFMM_OK;
}

int hfmm_batch_get_results(hfmm_batch* b, double* pot, double* fx, double* fy, double* fz) {
    if (!b) return HFMM_ERR_ARG;
    if (!b->evaluated) return bfail(b, HFMM_ERR_STATE, "get_results before evaluate");
    double* outs[4] = {pot, fx, fy, fz};
    const size_t N = static_cast<size_t>(b->n);
    for (Group& g : b->groups) {
        const size_t m = static_cast<size_t>(g.off.back());
        std::vector<double> tmp[4];
        double* p[4];
        for (int k = 0; k < 4; ++k) {
            p[k] = nullptr;
            if (outs[k]) {
                tmp[k].resize(m * b->nrhs);
                p[k] = tmp[k].data();
            }
        }
        const int st = hfmm_get_results(g.ctx, p[0], p[1], p[2], p[3]);
        if (st != HFMM_OK)
            return bfail(b, st, "height " + std::to_string(g.height) + ": " + hfmm_last_error(g.ctx));
        for (int k = 0; k < 4; ++k) {
            if (!outs[k]) continue;
            for (int32_t r = 0; r < b->nrhs; ++r)
                for (size_t i = 0; i < g.sets.size(); ++i) {
                    const int32_t s = g.sets[i];
                    const size_t cnt = static_cast<size_t>(g.off[i + 1] - g.off[i]);
                    std::memcpy(outs[k] + r * N + static_cast<size_t>(b->off[s]),
                                p[k] + r * m + static_cast<size_t>(g.off[i]), sizeof(double) * cnt);
                }
        }
    }
    return HFMM_OK;
}

int hfmm_batch_tree(hfmm_batch* b, int32_t i, hfmm_ctx** tree, int32_t* height, int32_t* nsets) {
    if (!b || !tree) return HFMM_ERR_ARG;
    if (i < 0 || i >= static_cast<int32_t>(b->groups.size())) return bfail(b, HFMM_ERR_ARG, "no such tree");
    *tree = b->groups[i].ctx;
    if (height) *height = b->groups[i].height;
    if (nsets) *nsets = static_cast<int32_t>(b->groups[i].sets.size());
    return HFMM_OK;
}

int hfmm_batch_info(const hfmm_batch* b, int32_t* ntrees, int32_t* graph_ready) {
    if (!b) return HFMM_ERR_ARG;
    if (ntrees) *ntrees = static_cast<int32_t>(b->groups.size());
    if (graph_ready) *graph_ready = b->exec ? 1 : 0;
    return HFMM_OK;
}

}  // extern "C"
