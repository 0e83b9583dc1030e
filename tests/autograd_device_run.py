"""Checks of the libtorch C++ autograd function on device tensors, run as a
process of its own by test_gpu_autograd.py: torch is imported before the
engine library, so both use torch's HIP runtime (a process that loaded the
engine first holds two runtimes, which cannot share device pointers).
Prints one JSON line of measured errors and times."""
import json
import os
import sys
import time

import torch  # first: the engine binds to torch's HIP runtime

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    from oracle import pyoracle
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED, Engine
    from scalfmm_amd.autograd import fmm_potential
    from test_gpu_autograd import _direct_grad, _gauss
    dev = torch.device("cuda")
    out = {}
    # small: against FP64 direct-sum gradients and the host-tensor path
    n, h, p = 4000, 4, 6
    X, q = _gauss(n, 5)
    ddq, ddx, dphi = _direct_grad(pyoracle, X, q)
    for mode, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, prec)
        xt = torch.tensor(X, device=dev, requires_grad=True)
        qt = torch.tensor(q, device=dev, requires_grad=True)
        phi = fmm_potential(xt, qt, e)
        (0.5 * (phi * phi).sum()).backward()
        dq, dx = qt.grad.cpu().numpy(), xt.grad.cpu().numpy()
        xh = torch.tensor(X, requires_grad=True)
        qh = torch.tensor(q, requires_grad=True)
        phih = fmm_potential(xh, qh, e)  # the host-tensor Function on the same engine
        (0.5 * (phih * phih).sum()).backward()
        out[mode] = {"phi_vs_direct": rel(phi.detach().cpu().numpy(), dphi), "dq_vs_direct": rel(dq, ddq),
                     "dx_vs_direct": rel(dx, ddx), "phi_vs_host": rel(phi.detach().cpu().numpy(), phih.detach().numpy()),
                     "dq_vs_host": rel(dq, qh.grad.numpy()), "dx_vs_host": rel(dx, xh.grad.numpy()),
                     "grad_dtype": str(xt.grad.dtype), "grad_device": str(xt.grad.device)}
        e.close()
    # float32 inputs: gradients come back in float32
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p)
    x32 = torch.tensor(X, device=dev, dtype=torch.float32, requires_grad=True)
    q32 = torch.tensor(q, device=dev, dtype=torch.float32, requires_grad=True)
    (0.5 * (fmm_potential(x32, q32, e) ** 2).sum()).backward()
    out["float32_grad_dtypes"] = [str(x32.grad.dtype), str(q32.grad.dtype)]
    e.close()
    # the differentiable configuration: 2M, height 7, order 6, mixed
    n, h, p = 2_000_000, 7, 6
    X, q = _gauss(n, 7)
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, PRECISION_MIXED)
    xt = torch.tensor(X, device=dev, requires_grad=True)
    qt = torch.tensor(q, device=dev, requires_grad=True)
    for _ in range(2):  # warm-up then timed
        xt.grad = qt.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phi = fmm_potential(xt, qt, e)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        (0.5 * (phi * phi).sum()).backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    dq, dx = qt.grad.cpu().numpy(), xt.grad.cpu().numpy()
    g = phi.detach().cpu().numpy()
    e.close()
    t = np.sort(np.random.default_rng(2).choice(n, 64, replace=False))
    ref = Engine(3, 1.0, (0.5, 0.5, 0.5), 5)
    ref.set_particles(X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(), q)
    phi_t, fxq, fyq, fzq = ref.direct(t)
    ref.set_particles(X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(), g)
    psi_t, fxg, fyg, fzg = ref.direct(t)
    ref.close()
    field_q = np.stack([fxq, fyq, fzq], axis=1) / q[t, None]
    field_g = np.stack([fxg, fyg, fzg], axis=1) / g[t, None]
    ddx = g[t, None] * field_q + q[t, None] * field_g
    out["full"] = {"n": n, "phi_vs_direct": rel(g[t], phi_t), "dq_vs_direct": rel(dq[t], psi_t),
                   "dx_vs_direct": rel(dx[t], ddx), "forward_s": t1 - t0, "backward_s": t2 - t1}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
