// direct.hip -- FP64 direct summation on the device for sampled targets: the
// O(N^2) reference the FMM results are measured against, as the reference's
// test programs compute it pair by pair with FP2P::NonMutualParticles
// (Tests/Kernels/DirectAlgorithm.cpp:203-215, Src/Kernels/P2P/FP2P.hpp:80-101).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

constexpr int kDirectBlock = 256;
constexpr int kDirectTargets = 8;  // targets per workgroup (sources are read once for all of them)

// A workgroup per kDirectTargets targets: the threads stride over all
// sources, then a tree reduction in LDS per target.  Force convention
// F = q_t * sum_j -q_j d / r^3, the FMM's.
__global__ __launch_bounds__(kDirectBlock) void k_direct(int64_t n, const double* __restrict__ x,
                                                        const double* __restrict__ y, const double* __restrict__ z,
                                                        const double* __restrict__ q,
                                                        const int64_t* __restrict__ targets, int64_t nt,
                                                        double* __restrict__ out) {
    __shared__ double red[4][kDirectBlock];
    int64_t t[kDirectTargets];
    double tx[kDirectTargets], ty[kDirectTargets], tz[kDirectTargets], acc[kDirectTargets][4];
    for (int u = 0; u < kDirectTargets; ++u) {
        const int64_t g = static_cast<int64_t>(blockIdx.x) * kDirectTargets + u;
        t[u] = g < nt ? targets[g] : -1;
        const int64_t tt = t[u] >= 0 ? t[u] : 0;
        tx[u] = x[tt];
        ty[u] = y[tt];
        tz[u] = z[tt];
        for (int k = 0; k < 4; ++k) acc[u][k] = 0.;
    }
    for (int64_t j = threadIdx.x; j < n; j += kDirectBlock) {
        const double xj = x[j], yj = y[j], zj = z[j], qj = q[j];
        for (int u = 0; u < kDirectTargets; ++u) {
            if (j == t[u]) continue;
            const double dx = tx[u] - xj, dy = ty[u] - yj, dz = tz[u] - zj;
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double qi = qj * inv;
            const double qi3 = qi * inv * inv;
            acc[u][0] += qi;
            acc[u][1] -= qi3 * dx;
            acc[u][2] -= qi3 * dy;
            acc[u][3] -= qi3 * dz;
        }
    }
    for (int u = 0; u < kDirectTargets; ++u) {
        for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[u][k];
        __syncthreads();
        for (int s = kDirectBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s)
                for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0 && t[u] >= 0) {
            const int64_t g = static_cast<int64_t>(blockIdx.x) * kDirectTargets + u;
            const double qt = q[t[u]];
            out[g] = red[0][0];
            out[nt + g] = red[1][0] * qt;
            out[2 * nt + g] = red[2][0] * qt;
            out[3 * nt + g] = red[3][0] * qt;
        }
        __syncthreads();
    }
}

// Staging of hfmm_direct_noalloc: the target indices (kDirectStaticTargets
// int64) followed by the results (4 kDirectStaticTargets doubles), 40 bytes per
// target, 2.5 MiB in all.  Every GPU the library runs on holds its own copy
// until the process ends; host code uses the per-device address that
// init_direct_scratch reads, never the array's name.
constexpr int64_t kDirectStaticTargets = HFMM_DIRECT_NOALLOC_MAX_TARGETS;
__device__ __align__(256) unsigned char scratch[kDirectStaticTargets * 5 * sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device

// Launch of k_direct over c's particles for nt targets whose indices are at
// device address dt, results (four arrays of nt) at dout.
int launch_direct(Context* c, int64_t nt, const int64_t* dt, double* dout) {
    const int64_t blocks = (nt + kDirectTargets - 1) / kDirectTargets;
    hipLaunchKernelGGL(k_direct, dim3(static_cast<unsigned>(blocks)), dim3(kDirectBlock), 0, c->stream, c->n, c->x,
                       c->y, c->z, c->q, dt, nt, dout);
    return check_launch(c, "direct");
}

// Copies the four result arrays at dout (nt each) into the caller's host
// arrays (NULL skips one) and waits for them.
int copy_direct_results(Context* c, int64_t nt, const double* dout, double* pot, double* fx, double* fy, double* fz) {
    double* outs[4] = {pot, fx, fy, fz};
    for (int k = 0; k < 4; ++k)
        if (outs[k] && hipMemcpyAsync(outs[k], dout + k * nt, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream) !=
                           hipSuccess)
            return fail(c, HFMM_ERR_DEVICE, "direct: result copy failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, HFMM_ERR_DEVICE, "direct: synchronize failed");
    return HFMM_OK;
}

}  // namespace

int init_direct_scratch(Context* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, HFMM_ERR_ARG, "direct staging: device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return HFMM_OK;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, HFMM_ERR_DEVICE, "direct staging: hipSetDevice failed");
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(c, HFMM_ERR_DEVICE, std::string("direct staging: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch_ptr[c->device] = ptr;
    return HFMM_OK;
}

int direct_sum_noalloc(Context* c, int64_t nt, const int64_t* targets, double* pot, double* fx, double* fy,
                       double* fz) {
    if (nt > kDirectStaticTargets)
        return fail(c, HFMM_ERR_ARG,
                    "direct_noalloc: " + std::to_string(nt) + " targets exceed the static staging's " +
                        std::to_string(kDirectStaticTargets));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxDevices) base = g_scratch_ptr[c->device];
    }
    if (!base) return fail(c, HFMM_ERR_STATE, "direct_noalloc: the static staging of this device is not set up");
    int64_t* dt = static_cast<int64_t*>(base);
    double* dout = reinterpret_cast<double*>(dt + kDirectStaticTargets);
    // every call rewrites the staging it reads: the target indices here, each
    // result entry by the kernel (one per target), so nothing needs clearing
    if (hipMemcpyAsync(dt, targets, sizeof(int64_t) * nt, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return fail(c, HFMM_ERR_DEVICE, "direct_noalloc: target copy failed");
    HFMM_TRY(launch_direct(c, nt, dt, dout));
    return copy_direct_results(c, nt, dout, pot, fx, fy, fz);
}

int direct_sum(Context* c, int64_t nt, const int64_t* targets, double* pot, double* fx, double* fy, double* fz) {
    DevBuf dt, dout;
    int st = ensure(c, dt, sizeof(int64_t) * nt);
    if (st == HFMM_OK) st = ensure(c, dout, sizeof(double) * 4 * nt);
    if (st == HFMM_OK && hipMemcpyAsync(dt.ptr, targets, sizeof(int64_t) * nt, hipMemcpyHostToDevice, c->stream) !=
                             hipSuccess)
        st = fail(c, HFMM_ERR_DEVICE, "direct: target copy failed");
    if (st == HFMM_OK) st = launch_direct(c, nt, dt.as<int64_t>(), dout.as<double>());
    if (st == HFMM_OK) st = copy_direct_results(c, nt, dout.as<double>(), pot, fx, fy, fz);
    release(dt);
    release(dout);
    return st;
}

}  // namespace hfmm
