"""The differentiable potential (scalfmm_amd.autograd) against FP64 direct sums.

L = sum_i phi_i^2 / 2.  Checks, for the FP64 and the mixed engine:
  - dL/dq and dL/dx against the gradient assembled from FP64 direct sums
    (oracle): relative L2 <= 2e-6 and 2e-5 at 4,000 particles, order 6 (the
    measured 1.7e-7 and 2.4e-6 with a tenfold margin);
  - FP64 central finite differences (h = 1e-4) of the direct-sum loss on
    sampled coordinates and charges: within 4e-6 of the largest gradient entry
    (measured 3.9e-7);
  - at 200k particles (height 6, order 6) and at 2M (height 7, order 6), dL/dx
    and dL/dq on 64 sampled particles against the device FP64 direct sum:
    <= 1e-5 and 1e-4 (measured 1.5e-6 and 9.7e-6 at 200k).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _engine(h, p, precision):
    from scalfmm_amd import Engine
    return Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, precision)


def _gauss(n, seed):
    from scalfmm_amd.distributions import Drand48, uniform_cube
    x, y, z, _ = uniform_cube(n, seed)
    g = Drand48(seed + 1000)  # Box-Muller on drand48 pairs
    u = g.draw(2 * n).reshape(n, 2)
    q = np.sqrt(-2.0 * np.log(1.0 - u[:, 0])) * np.cos(2 * np.pi * u[:, 1])
    return np.stack([x, y, z], axis=1), q


def _direct_grad(oracle, X, q):
    """dL/dq and dL/dx of L = sum phi^2 / 2 from FP64 direct sums."""
    x, y, z = X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy()
    phi, fx, fy, fz = oracle.direct(x, y, z, q)
    field_q = np.stack([fx, fy, fz], axis=1) / q[:, None]
    g = phi
    psi, gx, gy, gz = oracle.direct(x, y, z, g)
    field_g = np.stack([gx, gy, gz], axis=1) / g[:, None]
    return psi, g[:, None] * field_q + q[:, None] * field_g, phi


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_gradient_vs_direct_and_fd(oracle, mode):
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED
    from scalfmm_amd.autograd import fmm_potential
    n, h, p = 4000, 4, 6
    X, q = _gauss(n, 5)
    e = _engine(h, p, PRECISION_FP64 if mode == "fp64" else PRECISION_MIXED)
    xt = torch.from_numpy(X.copy()).requires_grad_(True)
    qt = torch.from_numpy(q.copy()).requires_grad_(True)
    phi = fmm_potential(xt, qt, e)
    loss = 0.5 * (phi * phi).sum()
    loss.backward()
    dq, dx = qt.grad.numpy(), xt.grad.numpy()
    ddq, ddx, dphi = _direct_grad(oracle, X, q)
    eq, ex = _rel(dq, ddq), _rel(dx, ddx)
    _say(f"{mode}: potential {_rel(phi.detach().numpy(), dphi):.2e}, dL/dq {eq:.2e}, dL/dx {ex:.2e} vs direct")
    assert eq < 2e-6 and ex < 2e-5

    # central differences of the direct-sum loss, FP64, h = 1e-4
    def L(Xv, qv):
        return 0.5 * float(np.sum(oracle.direct(Xv[:, 0].copy(), Xv[:, 1].copy(), Xv[:, 2].copy(), qv)[0] ** 2))
    rng = np.random.default_rng(1)
    hstep = 1e-4
    errs = []
    for _ in range(8):
        i, d = int(rng.integers(n)), int(rng.integers(3))
        Xp, Xm = X.copy(), X.copy()
        Xp[i, d] += hstep
        Xm[i, d] -= hstep
        fd = (L(Xp, q) - L(Xm, q)) / (2 * hstep)
        errs.append(abs(dx[i, d] - fd) / np.abs(ddx).max())
        j = int(rng.integers(n))
        qp, qm = q.copy(), q.copy()
        qp[j] += hstep
        qm[j] -= hstep
        fdq = (L(X, qp) - L(X, qm)) / (2 * hstep)
        errs.append(abs(dq[j] - fdq) / np.abs(ddq).max())
    _say(f"  finite differences: max error {max(errs):.2e} (relative to the largest gradient entry)")
    assert max(errs) < 4e-6


@pytest.mark.parametrize("mode,n,h", [("fp64", 200_000, 6), ("mixed", 200_000, 6), ("mixed", 2_000_000, 7)])
def test_gradient_sampled(mode, n, h):
    """Charges N(0,1), order 6: 64 sampled particles against the device FP64
    direct sum (2M particles at height 7 is the differentiable configuration)."""
    import time
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED
    from scalfmm_amd.autograd import fmm_potential
    p = 6
    X, q = _gauss(n, 7)
    e = _engine(h, p, PRECISION_FP64 if mode == "fp64" else PRECISION_MIXED)
    xt = torch.from_numpy(X.copy()).requires_grad_(True)
    qt = torch.from_numpy(q.copy()).requires_grad_(True)
    t0 = time.perf_counter()
    phi = fmm_potential(xt, qt, e)
    t1 = time.perf_counter()
    (0.5 * (phi * phi).sum()).backward()
    t2 = time.perf_counter()
    dq, dx = qt.grad.numpy(), xt.grad.numpy()
    t = np.sort(np.random.default_rng(2).choice(n, 64, replace=False))
    ref = _engine(3, 5, PRECISION_FP64)
    ref.set_particles(X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(), q)
    phi_t, fxq, fyq, fzq = ref.direct(t)
    g = phi.detach().numpy()  # dL/dphi = phi, the FMM's, as in the backward pass
    ref.set_particles(X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(), g)
    psi_t, fxg, fyg, fzg = ref.direct(t)
    field_q = np.stack([fxq, fyq, fzq], axis=1) / q[t, None]
    field_g = np.stack([fxg, fyg, fzg], axis=1) / g[t, None]
    ddx = g[t, None] * field_q + q[t, None] * field_g
    eq, ex = _rel(dq[t], psi_t), _rel(dx[t], ddx)
    _say(f"{mode} n={n}: phi {_rel(g[t], phi_t):.2e}, dL/dq {eq:.2e}, dL/dx {ex:.2e} on 64 particles vs direct; "
         f"forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s (host round trips included)")
    assert eq < 1e-5 and ex < 1e-4


def test_gradient_fd_full_size():
    """The differentiable configuration at full size (2M particles, charges
    N(0,1), height 7, order 6, mixed form): dL/dx on 64 sampled coordinates
    against FP64 central finite differences (h = 1e-4) of the direct-sum loss
    L = sum_i phi_i^2 / 2, and dL/dq on 16 sampled charges against the same.

    Moving x_j by +-h changes phi_i (i != j) by d_i = q_j (1/|x_i - x_j'| -
    1/|x_i - x_j|) and recomputes phi_j, so L(+) - L(-) is summed from these
    differences directly (no cancellation of two large losses); phi is the
    FP64 direct sum at every particle (device kernel).  L is linear-quadratic
    in q: its central difference in q_j is sum_{i != j} phi_i / r_ij."""
    import time
    from scalfmm_amd import PRECISION_MIXED
    from scalfmm_amd.autograd import fmm_potential
    n, h, p = 2_000_000, 7, 6
    X, q = _gauss(n, 7)
    e = _engine(h, p, PRECISION_MIXED)
    xt = torch.from_numpy(X.copy()).requires_grad_(True)
    qt = torch.from_numpy(q.copy()).requires_grad_(True)
    phi = fmm_potential(xt, qt, e)
    (0.5 * (phi * phi).sum()).backward()
    dq, dx = qt.grad.numpy(), xt.grad.numpy()
    e.close()
    ref = _engine(3, 5, 0)
    ref.set_particles(X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(), q)
    t0 = time.perf_counter()
    phi_d = ref.direct(np.arange(n))[0]  # FP64 direct potential at every particle
    _say(f"  FP64 direct sum at {n} targets: {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(11)
    hs = 1e-4
    fd, got = [], []
    for _ in range(64):
        j, d = int(rng.integers(n)), int(rng.integers(3))
        diff = X - X[j]
        r0 = np.sqrt(np.einsum("ij,ij->i", diff, diff))
        r0[j] = np.inf
        parts = []
        for sgn in (1.0, -1.0):
            xj = X[j].copy()
            xj[d] += sgn * hs
            dd = X - xj
            r = np.sqrt(np.einsum("ij,ij->i", dd, dd))
            r[j] = np.inf
            delta = q[j] * (1.0 / r - 1.0 / r0)
            phij = np.sum(q / r)
            parts.append((delta, phij))
        (dp, pp), (dm, pm) = parts
        dL = np.sum(phi_d * (dp - dm) + 0.5 * (dp * dp - dm * dm)) + 0.5 * (pp * pp - pm * pm)
        fd.append(dL / (2 * hs))
        got.append(dx[j, d])
    fd, got = np.array(fd), np.array(got)
    ex = float(np.max(np.abs(got - fd)) / np.max(np.abs(fd)))
    fq, gq = [], []
    for _ in range(16):
        j = int(rng.integers(n))
        diff = X - X[j]
        r = np.sqrt(np.einsum("ij,ij->i", diff, diff))
        r[j] = np.inf
        fq.append(np.sum(phi_d / r))
        gq.append(dq[j])
    eq = float(np.max(np.abs(np.array(gq) - np.array(fq))) / np.max(np.abs(fq)))
    _say(f"  finite differences at n={n}: dL/dx {ex:.2e}, dL/dq {eq:.2e} (max error / max entry)")
    assert ex < 1e-4 and eq < 1e-5


def test_cpp_function_device_tensors():
    """The libtorch C++ autograd function (csrc/torch_potential.cpp) on device
    tensors, in a process that imports torch before the engine
    (tests/autograd_device_run.py): at 4,000 particles the same bars against
    the FP64 direct-sum gradient as above, and the host-tensor Function on the
    same engine within 1e-13 (FP64) / 1e-6 (mixed, both FP32 near fields);
    float32 inputs get float32 gradients; at the differentiable configuration
    (2M, height 7, order 6, mixed) the sampled bars of test_gradient_sampled."""
    import json
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-u", os.path.join(root, "tests", "autograd_device_run.py")],
                       capture_output=True, text=True, timeout=600)
    _say(r.stdout[-2000:] + r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for mode in ("fp64", "mixed"):
        m = out[mode]
        assert m["dq_vs_direct"] < 2e-6 and m["dx_vs_direct"] < 2e-5, m
        host_tol = 1e-13 if mode == "fp64" else 1e-6
        assert m["phi_vs_host"] < host_tol and m["dq_vs_host"] < host_tol and m["dx_vs_host"] < host_tol, m
        assert m["grad_device"].startswith("cuda") and m["grad_dtype"] == "torch.float64"
    assert out["float32_grad_dtypes"] == ["torch.float32", "torch.float32"]
    f = out["full"]
    assert f["dq_vs_direct"] < 1e-5 and f["dx_vs_direct"] < 1e-4, f
