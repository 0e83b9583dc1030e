// torch_potential.cpp -- the FMM potential of the engine as a libtorch C++
// autograd function over device tensors (Python module
// scalfmm_amd._lib._torch_fmm).
//
//   phi_i = sum_{j != i} q_j / |x_i - x_j|
//
// For a loss L(phi) with g = dL/dphi the backward returns
//   dL/dq_j = psi_j = sum_{i != j} g_i / |x_i - x_j|   (the 1/r operator is
//             symmetric: the adjoint FMM is the forward FMM with charges g on
//             the forward's tree)
//   dL/dx_j = g_j grad(phi)(x_j) + q_j grad(psi)(x_j)
// The forward keeps grad(phi) (3N values, hfmm_get_field_device) and the
// inputs; the backward recomputes every expansion instead of storing them.
// Tensors stay on the device: the engine reads x, y, z and q from torch's
// memory (hfmm_set_particles_device, hfmm_set_charges_device) and its results
// are copied into new tensors on the device.  The reference has no
// counterpart (its kernels have no derivative with respect to the charges or
// the positions).
//
// The engine runs on a stream of its own: every entry waits for the device
// (torch's pending work on the inputs) and for the engine before torch reads
// its results.  Engine and torch must share one HIP runtime, which holds
// when torch is loaded before libscalfmm_hip.so (both carry the SONAME
// libamdhip64.so.7, so the dynamic loader binds the engine to torch's copy);
// the first call checks it and refuses device tensors otherwise.
#include <dlfcn.h>
#include <torch/cuda.h>
#include <torch/extension.h>

#include <climits>
#include <cstdlib>
#include <string>

#include "../../include/scalfmm_hip.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void check(hfmm_ctx* e, int rc, const char* what) {
    TORCH_CHECK(rc == HFMM_OK, "scalfmm engine: ", what, ": ", hfmm_last_error(e));
}

// Path of the library that defines `sym` as seen from the library holding `anchor`.
std::string binding(const void* anchor, const char* sym) {
    Dl_info self{};
    if (!dladdr(anchor, &self) || !self.dli_fname) return "";
    void* h = dlopen(self.dli_fname, RTLD_NOW | RTLD_NOLOAD);
    if (!h) return "";
    void* s = dlsym(h, sym);
    dlclose(h);
    Dl_info where{};
    if (!s || !dladdr(s, &where) || !where.dli_fname) return "";
    char buf[PATH_MAX];
    return realpath(where.dli_fname, buf) ? std::string(buf) : std::string(where.dli_fname);
}

void check_runtime() {
    static int ok = -1;
    static std::string msg;
    if (ok < 0) {
        const std::string engine = binding(reinterpret_cast<const void*>(&hfmm_create), "hipMalloc");
        std::string torch_rt;
        if (void* th = dlopen("libtorch_hip.so", RTLD_NOW | RTLD_NOLOAD)) {
            void* s = dlsym(th, "hipMalloc");
            Dl_info where{};
            if (s && dladdr(s, &where) && where.dli_fname) {
                char buf[PATH_MAX];
                torch_rt = realpath(where.dli_fname, buf) ? std::string(buf) : std::string(where.dli_fname);
            }
            dlclose(th);
        }
        ok = !engine.empty() && engine == torch_rt;
        msg = "the engine's HIP runtime (" + engine + ") is not torch's (" + torch_rt +
              "): import torch before loading libscalfmm_hip.so";
    }
    TORCH_CHECK(ok == 1, msg);
}

torch::Tensor device_copy(const double* p, int64_t n, const torch::Device& dev) {
    const auto opts = torch::TensorOptions().dtype(torch::kFloat64).device(dev);
    return torch::from_blob(const_cast<double*>(p), {n}, opts).clone();
}

// Potential and field of the engine's last evaluation as new device tensors.
void results(hfmm_ctx* e, int64_t n, const torch::Device& dev, torch::Tensor& pot, torch::Tensor& field) {
    const double *p = nullptr, *ex = nullptr, *ey = nullptr, *ez = nullptr;
    check(e, hfmm_get_results_device(e, &p, nullptr, nullptr, nullptr), "get_results_device");
    check(e, hfmm_get_field_device(e, &ex, &ey, &ez), "get_field_device");
    pot = device_copy(p, n, dev);
    field = torch::stack({device_copy(ex, n, dev), device_copy(ey, n, dev), device_copy(ez, n, dev)}, 1);
    torch::cuda::synchronize(dev.index());  // the copies are done before the engine reuses its buffers
}

// One FMM of charges q at the positions xs (rows x, y, z), both device
// tensors that stay alive while the engine refers to them.
void evaluate(hfmm_ctx* e, const torch::Tensor& xs, const torch::Tensor& q, int64_t n, int device) {
    torch::cuda::synchronize(device);  // torch's producers of xs and q
    check(e, hfmm_set_particles_device(e, n, xs[0].data_ptr<double>(), xs[1].data_ptr<double>(),
                                       xs[2].data_ptr<double>(), q.data_ptr<double>()),
          "set_particles_device");
    check(e, hfmm_build_tree(e), "build_tree");
    check(e, hfmm_evaluate(e, HFMM_NEAR_AND_FAR), "evaluate");
    check(e, hfmm_synchronize(e), "synchronize");
}

class FmmPotential : public torch::autograd::Function<FmmPotential> {
public:
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor x, torch::Tensor q, int64_t handle) {
        check_runtime();
        TORCH_CHECK(x.is_cuda() && q.is_cuda() && x.device() == q.device(), "x and q on one GPU");
        TORCH_CHECK(x.dim() == 2 && x.size(1) == 3, "x of shape (N, 3)");
        TORCH_CHECK(q.dim() == 1 && q.size(0) == x.size(0), "q of shape (N,)");
        hfmm_ctx* e = reinterpret_cast<hfmm_ctx*>(handle);
        TORCH_CHECK(e != nullptr, "null engine");
        const int64_t n = x.size(0);
        // SoA copies in FP64 owned by the autograd context: the engine reads
        // them during this forward and again in the backward
        const torch::Tensor xs = x.detach().to(torch::kFloat64).t().contiguous();
        const torch::Tensor qs = q.detach().to(torch::kFloat64).contiguous();
        const int dev = static_cast<int>(x.device().index());
        evaluate(e, xs, qs, n, dev);
        torch::Tensor pot, field;
        results(e, n, x.device(), pot, field);
        ctx->saved_data["xs"] = xs;
        ctx->saved_data["qs"] = qs;
        ctx->saved_data["field"] = field;
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["xtype"] = static_cast<int64_t>(x.scalar_type());
        ctx->saved_data["qtype"] = static_cast<int64_t>(q.scalar_type());
        return pot.to(x.scalar_type());
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        hfmm_ctx* e = reinterpret_cast<hfmm_ctx*>(ctx->saved_data["handle"].toInt());
        const torch::Tensor xs = ctx->saved_data["xs"].toTensor();
        const torch::Tensor qs = ctx->saved_data["qs"].toTensor();
        const torch::Tensor field = ctx->saved_data["field"].toTensor();
        const int64_t n = qs.size(0);
        const int dev = static_cast<int>(qs.device().index());
        const torch::Tensor g = grads[0].detach().to(torch::kFloat64).contiguous();
        // the adjoint: the forward's positions with charges g (the tree is
        // rebuilt, the engine may have run another problem since)
        evaluate(e, xs, g, n, dev);
        torch::Tensor psi, field_g;
        results(e, n, qs.device(), psi, field_g);
        // the engine is left with the forward's charges on the forward's
        // tree; both arrays belong to this autograd context
        check(e, hfmm_set_charges_device(e, qs.data_ptr<double>()), "set_charges_device");
        const torch::Tensor dx = g.unsqueeze(1) * field + qs.unsqueeze(1) * field_g;
        const auto xt = static_cast<torch::ScalarType>(ctx->saved_data["xtype"].toInt());
        const auto qt = static_cast<torch::ScalarType>(ctx->saved_data["qtype"].toInt());
        return {dx.to(xt), psi.to(qt), torch::Tensor()};
    }
};

torch::Tensor potential(torch::Tensor x, torch::Tensor q, int64_t handle) {
    return FmmPotential::apply(x, q, handle);
}

bool same_runtime() {
    try {
        check_runtime();
        return true;
    } catch (const c10::Error&) {
        return false;
    }
}

}  // namespace

PYBIND11_MODULE(_torch_fmm, m) {
    m.doc() = "FMM potential as a torch::autograd::Function over device tensors";
    m.def("potential", &potential, "phi = potential(x (N, 3), q (N,), engine handle)");
    m.def("same_runtime", &same_runtime, "whether the engine and torch share one HIP runtime");
}
