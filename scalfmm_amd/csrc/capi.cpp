// capi.cpp -- the reference's public C API (include/scalfmm_capi.h) for its
// built-in Chebyshev kernel (order 7, 1/r, FP64) over the engine of
// scalfmm_hip.h.  Follows Addons/CKernelApi/Src/FInterEngine.hpp and
// FScalFMMEngine.hpp of the reference for the semantics of every call:
// particles addressed by their index within their insert call, results that
// accumulate over executions, sources and targets as one set except with the
// source_target algorithm.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/scalfmm_capi.h"
#include "hfmm_internal.hpp"

namespace {

constexpr int kCApiOrder = 7;  // FChebSymKernel<..., 7> (FScalfmmApiInit.cpp:42, 85)

struct Particles {
    std::vector<double> x, y, z, q, pot, fx, fy, fz;
    size_t size() const { return x.size(); }
    void reset(int n) {
        for (auto* v : {&x, &y, &z, &q, &pot, &fx, &fy, &fz}) v->assign(static_cast<size_t>(n), 0.0);
    }
};

struct CApi {
    scalfmm_algorithm algo = sequential;
    bool built = false;
    int height = 0;
    double width = 0, center[3] = {0, 0, 0};
    Particles both, src, tgt;
    hfmm_ctx* ctx = nullptr;
    double timers[7] = {0, 0, 0, 0, 0, 0, 0};
};

void say(const char* what) { std::fprintf(stderr, "scalfmm C API: %s\n", what); }

CApi* H(scalfmm_handle h) { return static_cast<CApi*>(h); }

// The particle set a call addresses: one set unless sources and targets are separate.
Particles* set_of(CApi* c, PartType t) {
    if (c->algo != source_target) return &c->both;
    if (t == SOURCE) return &c->src;
    if (t == TARGET) return &c->tgt;
    say("BOTH is not a particle set with the source_target algorithm");
    return nullptr;
}

bool fits(const Particles* p, int n, const char* what) {
    if (!p) return false;
    if (n < 0 || static_cast<size_t>(n) > p->size()) {
        say((std::string(what) + ": more values than particles").c_str());
        return false;
    }
    if (static_cast<size_t>(n) < p->size()) say((std::string("Not all parts has been read/set: ") + what).c_str());
    return true;
}

// idx[i] -> position of the particle with that index (-1 when absent).
bool index_ok(const Particles* p, int idx) { return idx >= 0 && static_cast<size_t>(idx) < p->size(); }

void insert(CApi* c, int n, const double* X, const double* Y, const double* Z, int stride, PartType type) {
    if (!c || n < 0) return;
    Particles* p = set_of(c, type);
    if (!p) return;
    p->reset(n);
    for (int i = 0; i < n; ++i) {
        p->x[i] = X[static_cast<size_t>(i) * stride];
        p->y[i] = Y[static_cast<size_t>(i) * stride];
        p->z[i] = Z[static_cast<size_t>(i) * stride];
    }
}

// Runs the engine on the particle sets and adds its results into them.
void execute(CApi* c, uint32_t ops) {
    if (!c->built || !c->ctx) return say("execute before build_tree");
    const bool st = c->algo == source_target;
    const size_t ns = st ? c->src.size() : c->both.size(), nt = st ? c->tgt.size() : 0;
    const size_t n = ns + nt;
    if (n == 0) return say("execute without particles");
    std::vector<double> x(n), y(n), z(n), q(n, 0.0);
    const Particles& a = st ? c->src : c->both;
    for (size_t i = 0; i < ns; ++i) x[i] = a.x[i], y[i] = a.y[i], z[i] = a.z[i], q[i] = a.q[i];
    for (size_t i = 0; i < nt; ++i)  // targets carry no charge into the sums
        x[ns + i] = c->tgt.x[i], y[ns + i] = c->tgt.y[i], z[ns + i] = c->tgt.z[i];
    hfmm_ctx* h = c->ctx;
    std::vector<double> pot(n), fx(n), fy(n), fz(n);
    hfmm_reset_kernel_times(h);
    if (hfmm_set_particles(h, static_cast<int64_t>(n), x.data(), y.data(), z.data(), q.data()) ||
        hfmm_build_tree(h) || hfmm_evaluate(h, ops) ||
        (st ? hfmm_get_results(h, pot.data(), nullptr, nullptr, nullptr) || hfmm_get_field(h, fx.data(), fy.data(), fz.data())
            : hfmm_get_results(h, pot.data(), fx.data(), fy.data(), fz.data())))
        return say(hfmm_last_error(h));
    if (st) {
        Particles& t = c->tgt;
        for (size_t i = 0; i < nt; ++i) {
            t.pot[i] += pot[ns + i];
            t.fx[i] += t.q[i] * fx[ns + i];  // force = q grad(phi) of the target
            t.fy[i] += t.q[i] * fy[ns + i];
            t.fz[i] += t.q[i] * fz[ns + i];
        }
    } else {
        Particles& b = c->both;
        for (size_t i = 0; i < ns; ++i) b.pot[i] += pot[i], b.fx[i] += fx[i], b.fy[i] += fy[i], b.fz[i] += fz[i];
    }
    // timers (seconds): P2M, M2M, M2L, L2L, L2P, P2P, near field
    hfmm_kernel_time kt[64];
    int32_t cnt = 0;
    std::memset(c->timers, 0, sizeof(c->timers));
    if (hfmm_get_kernel_times(h, kt, 64, &cnt) == HFMM_OK) {
        for (int i = 0; i < cnt && i < 64; ++i) {
            const std::string nm(kt[i].name);
            const double s = kt[i].total_ms * 1e-3;
            if (nm == "p2m") c->timers[0] += s;
            else if (nm == "m2m") c->timers[1] += s;
            else if (nm.rfind("m2l", 0) == 0 && nm.find("list") == std::string::npos) c->timers[2] += s;
            else if (nm == "l2l") c->timers[3] += s;
            else if (nm == "l2p") c->timers[4] += s;
            else if (nm.rfind("p2p", 0) == 0 && nm.find("list") == std::string::npos &&
                     nm.find("items") == std::string::npos)
                c->timers[5] += s;
        }
        c->timers[6] = c->timers[5];
    }
}

// Copies a field of a set out (index order) or in; npart variants address the
// particles idx[0..n).
void get_field(CApi* c, const std::vector<double> Particles::*f, int n, double* out, PartType type, const char* what) {
    Particles* p = set_of(c, type);
    if (!fits(p, n, what)) return;
    for (int i = 0; i < n; ++i) out[i] = (p->*f)[i];
}
void set_field(CApi* c, std::vector<double> Particles::*f, int n, const double* in, PartType type, const char* what) {
    Particles* p = set_of(c, type);
    if (!fits(p, n, what)) return;
    for (int i = 0; i < n; ++i) (p->*f)[i] = in[i];
}
void get_field_npart(CApi* c, const std::vector<double> Particles::*f, int n, const int* idx, double* out,
                     PartType type) {
    Particles* p = set_of(c, type);
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) out[i] = (p->*f)[idx[i]];
        else say("npart: no particle with this index");
}
void set_field_npart(CApi* c, std::vector<double> Particles::*f, int n, const int* idx, const double* in,
                     PartType type) {
    Particles* p = set_of(c, type);
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) (p->*f)[idx[i]] = in[i];
        else say("npart: no particle with this index");
}

bool no_source_forces(PartType type, const CApi* c) {
    if (type == SOURCE && c->algo == source_target) {
        say("No meaning to retrieve source forces ...");
        return true;
    }
    return false;
}

}  // namespace

extern "C" {

scalfmm_handle scalfmm_init(scalfmm_kernel_type KernelType, scalfmm_algorithm algo) {
    if (KernelType != chebyshev) {
        say("Kernel type unsupported (this library serves the chebyshev kernel)");
        return nullptr;
    }
    if (algo != sequential && algo != multi_thread && algo != source_target) {
        say("algorithm unsupported (sequential, multi_thread or source_target)");
        return nullptr;
    }
    CApi* c = new (std::nothrow) CApi();
    if (c) c->algo = algo;
    return c;
}

void scalfmm_build_tree(scalfmm_handle handle, int TreeHeight, double BoxWidth, double* BoxCenter,
                        Scalfmm_Cell_Descriptor, Scalfmm_Leaf_Descriptor) {
    CApi* c = H(handle);
    if (!c || !BoxCenter) return;
    if (c->ctx) hfmm_destroy(c->ctx);
    c->ctx = nullptr;
    hfmm_params p{};
    p.tree_height = TreeHeight;
    p.order = kCApiOrder;
    p.box_width = BoxWidth;
    for (int d = 0; d < 3; ++d) p.box_center[d] = BoxCenter[d];
    p.epsilon = 0.0;
    p.precision = HFMM_PRECISION_FP64;
    p.device = 0;
    if (hfmm_create(&p, &c->ctx) != HFMM_OK) {
        c->ctx = nullptr;
        return say("cannot create the engine (bad tree height or box, or no device)");
    }
    hfmm_set_profiling(c->ctx, 1);
    c->height = TreeHeight;
    c->width = BoxWidth;
    for (int d = 0; d < 3; ++d) c->center[d] = BoxCenter[d];
    c->built = true;
}

void scalfmm_tree_insert_particles(scalfmm_handle Handle, int NbPositions, double* arrayX, double* arrayY,
                                   double* arrayZ, PartType type) {
    if (NbPositions > 0 && (!arrayX || !arrayY || !arrayZ)) return say("insert: null arrays");
    insert(H(Handle), NbPositions, arrayX, arrayY, arrayZ, 1, type);
}

void scalfmm_tree_insert_particles_xyz(scalfmm_handle Handle, int NbPositions, double* XYZ, PartType type) {
    if (NbPositions > 0 && !XYZ) return say("insert: null array");
    insert(H(Handle), NbPositions, XYZ, XYZ + 1, XYZ + 2, 3, type);
}

void scalfmm_set_physical_values(scalfmm_handle Handle, int n, double* v, PartType type) {
    if (H(Handle)) set_field(H(Handle), &Particles::q, n, v, type, "physical values");
}
void scalfmm_get_physical_values(scalfmm_handle Handle, int n, double* v, PartType type) {
    if (H(Handle)) get_field(H(Handle), &Particles::q, n, v, type, "physical values");
}
void scalfmm_set_physical_values_npart(scalfmm_handle Handle, int n, int* idx, double* v, PartType type) {
    if (H(Handle)) set_field_npart(H(Handle), &Particles::q, n, idx, v, type);
}
void scalfmm_get_physical_values_npart(scalfmm_handle Handle, int n, int* idx, double* v, PartType type) {
    if (H(Handle)) get_field_npart(H(Handle), &Particles::q, n, idx, v, type);
}

void scalfmm_get_forces_xyz(scalfmm_handle Handle, int n, double* f, PartType type) {
    CApi* c = H(Handle);
    if (!c || no_source_forces(type, c)) return;
    Particles* p = set_of(c, type);
    if (!fits(p, n, "forces")) return;
    for (int i = 0; i < n; ++i) f[3 * i] = p->fx[i], f[3 * i + 1] = p->fy[i], f[3 * i + 2] = p->fz[i];
}
void scalfmm_get_forces_xyz_npart(scalfmm_handle Handle, int n, int* idx, double* f, PartType type) {
    CApi* c = H(Handle);
    if (!c || no_source_forces(type, c)) return;
    Particles* p = set_of(c, type);
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) f[3 * i] = p->fx[idx[i]], f[3 * i + 1] = p->fy[idx[i]], f[3 * i + 2] = p->fz[idx[i]];
}
void scalfmm_get_forces(scalfmm_handle Handle, int n, double* fX, double* fY, double* fZ, PartType type) {
    CApi* c = H(Handle);
    if (!c || no_source_forces(type, c)) return;
    get_field(c, &Particles::fx, n, fX, type, "forces");
    get_field(c, &Particles::fy, n, fY, type, "forces");
    get_field(c, &Particles::fz, n, fZ, type, "forces");
}
void scalfmm_get_forces_npart(scalfmm_handle Handle, int n, int* idx, double* fX, double* fY, double* fZ,
                              PartType type) {
    CApi* c = H(Handle);
    if (!c || no_source_forces(type, c)) return;
    get_field_npart(c, &Particles::fx, n, idx, fX, type);
    get_field_npart(c, &Particles::fy, n, idx, fY, type);
    get_field_npart(c, &Particles::fz, n, idx, fZ, type);
}
void scalfmm_set_forces_xyz(scalfmm_handle Handle, int n, double* f, PartType type) {
    CApi* c = H(Handle);
    if (!c) return;
    Particles* p = set_of(c, type);
    if (!fits(p, n, "forces")) return;
    for (int i = 0; i < n; ++i) p->fx[i] = f[3 * i], p->fy[i] = f[3 * i + 1], p->fz[i] = f[3 * i + 2];
}
void scalfmm_set_forces_xyz_npart(scalfmm_handle Handle, int n, int* idx, double* f, PartType type) {
    CApi* c = H(Handle);
    if (!c) return;
    Particles* p = set_of(c, type);
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) p->fx[idx[i]] = f[3 * i], p->fy[idx[i]] = f[3 * i + 1], p->fz[idx[i]] = f[3 * i + 2];
}
void scalfmm_set_forces(scalfmm_handle Handle, int n, double* fX, double* fY, double* fZ, PartType type) {
    CApi* c = H(Handle);
    if (!c) return;
    set_field(c, &Particles::fx, n, fX, type, "forces");
    set_field(c, &Particles::fy, n, fY, type, "forces");
    set_field(c, &Particles::fz, n, fZ, type, "forces");
}
void scalfmm_set_forces_npart(scalfmm_handle Handle, int n, int* idx, double* fX, double* fY, double* fZ,
                              PartType type) {
    CApi* c = H(Handle);
    if (!c) return;
    set_field_npart(c, &Particles::fx, n, idx, fX, type);
    set_field_npart(c, &Particles::fy, n, idx, fY, type);
    set_field_npart(c, &Particles::fz, n, idx, fZ, type);
}

void scalfmm_get_potentials(scalfmm_handle Handle, int n, double* v, PartType type) {
    if (H(Handle)) get_field(H(Handle), &Particles::pot, n, v, type, "potentials");
}
void scalfmm_set_potentials(scalfmm_handle Handle, int n, double* v, PartType type) {
    if (H(Handle)) set_field(H(Handle), &Particles::pot, n, v, type, "potentials");
}
void scalfmm_get_potentials_npart(scalfmm_handle Handle, int n, int* idx, double* v, PartType type) {
    if (H(Handle)) get_field_npart(H(Handle), &Particles::pot, n, idx, v, type);
}
void scalfmm_set_potentials_npart(scalfmm_handle Handle, int n, int* idx, double* v, PartType type) {
    if (H(Handle)) set_field_npart(H(Handle), &Particles::pot, n, idx, v, type);
}

void scalfmm_add_to_positions_xyz(scalfmm_handle Handle, int n, double* d, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!fits(p, n, "positions")) return;
    for (int i = 0; i < n; ++i) p->x[i] += d[3 * i], p->y[i] += d[3 * i + 1], p->z[i] += d[3 * i + 2];
}
void scalfmm_add_to_positions(scalfmm_handle Handle, int n, double* X, double* Y, double* Z, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!fits(p, n, "positions")) return;
    for (int i = 0; i < n; ++i) p->x[i] += X[i], p->y[i] += Y[i], p->z[i] += Z[i];
}
void scalfmm_add_to_positions_xyz_npart(scalfmm_handle Handle, int n, int* idx, double* d, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) p->x[idx[i]] += d[3 * i], p->y[idx[i]] += d[3 * i + 1], p->z[idx[i]] += d[3 * i + 2];
}
void scalfmm_add_to_positions_npart(scalfmm_handle Handle, int n, int* idx, double* X, double* Y, double* Z,
                                    PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) p->x[idx[i]] += X[i], p->y[idx[i]] += Y[i], p->z[idx[i]] += Z[i];
}
void scalfmm_set_positions_xyz(scalfmm_handle Handle, int n, double* d, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!fits(p, n, "positions")) return;
    for (int i = 0; i < n; ++i) p->x[i] = d[3 * i], p->y[i] = d[3 * i + 1], p->z[i] = d[3 * i + 2];
}
void scalfmm_set_positions(scalfmm_handle Handle, int n, double* X, double* Y, double* Z, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!fits(p, n, "positions")) return;
    for (int i = 0; i < n; ++i) p->x[i] = X[i], p->y[i] = Y[i], p->z[i] = Z[i];
}
void scalfmm_set_positions_xyz_npart(scalfmm_handle Handle, int n, int* idx, double* d, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) p->x[idx[i]] = d[3 * i], p->y[idx[i]] = d[3 * i + 1], p->z[idx[i]] = d[3 * i + 2];
}
void scalfmm_set_positions_npart(scalfmm_handle Handle, int n, int* idx, double* X, double* Y, double* Z,
                                 PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) p->x[idx[i]] = X[i], p->y[idx[i]] = Y[i], p->z[idx[i]] = Z[i];
}
void scalfmm_update_tree(scalfmm_handle) {}

void scalfmm_get_positions_xyz(scalfmm_handle Handle, int n, double* d, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!fits(p, n, "positions")) return;
    for (int i = 0; i < n; ++i) d[3 * i] = p->x[i], d[3 * i + 1] = p->y[i], d[3 * i + 2] = p->z[i];
}
void scalfmm_get_positions(scalfmm_handle Handle, int n, double* X, double* Y, double* Z, PartType type) {
    CApi* c = H(Handle);
    if (!c) return;
    get_field(c, &Particles::x, n, X, type, "positions");
    get_field(c, &Particles::y, n, Y, type, "positions");
    get_field(c, &Particles::z, n, Z, type, "positions");
}
void scalfmm_get_positions_xyz_npart(scalfmm_handle Handle, int n, int* idx, double* d, PartType type) {
    CApi* c = H(Handle);
    Particles* p = c ? set_of(c, type) : nullptr;
    if (!p) return;
    for (int i = 0; i < n; ++i)
        if (index_ok(p, idx[i])) d[3 * i] = p->x[idx[i]], d[3 * i + 1] = p->y[idx[i]], d[3 * i + 2] = p->z[idx[i]];
}
void scalfmm_get_positions_npart(scalfmm_handle Handle, int n, int* idx, double* X, double* Y, double* Z,
                                 PartType type) {
    CApi* c = H(Handle);
    if (!c) return;
    get_field_npart(c, &Particles::x, n, idx, X, type);
    get_field_npart(c, &Particles::y, n, idx, Y, type);
    get_field_npart(c, &Particles::z, n, idx, Z, type);
}

void scalfmm_algorithm_config(scalfmm_handle Handle, scalfmm_algorithm config) {
    CApi* c = H(Handle);
    if (!c) return;
    if (config != sequential && config != multi_thread && config != source_target)
        return say("algorithm unsupported (sequential, multi_thread or source_target)");
    const bool was_st = c->algo == source_target, is_st = config == source_target;
    if (was_st != is_st && (c->both.size() || c->src.size() || c->tgt.size()))
        return say("switching to or from source_target needs the particles to be inserted again");
    c->algo = config;
}

void scalfmm_execute_fmm(scalfmm_handle Handle) {
    if (H(Handle)) execute(H(Handle), HFMM_NEAR_AND_FAR);
}
void scalfmm_execute_fmm_far_field(scalfmm_handle Handle) {
    if (H(Handle)) execute(H(Handle), HFMM_FAR_FIELD);
}

void scalfmm_dealloc_handle(scalfmm_handle handle, Callback_free_cell) {
    CApi* c = H(handle);
    if (!c) return;
    if (c->ctx) hfmm_destroy(c->ctx);
    delete c;
}

int scalfmm_get_nb_timers(scalfmm_handle) { return 9; }  // FAlgorithmTimers::nbTimers

void scalfmm_get_timers(scalfmm_handle handle, double* Timers) {
    CApi* c = H(handle);
    if (!c || !Timers) return;
    for (int i = 0; i < 7; ++i) Timers[i] = c->timers[i];
}

void scalfmm_set_upper_limit(scalfmm_handle, int) {
    say("This feature is not available with Chebyshev Kernel");
}

}  // extern "C"
