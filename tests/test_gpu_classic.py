"""The classic-tree caller (scalfmm_amd.octree) on the GPU, as the reference's
unit test drives the Chebyshev kernel: UTests/utestChebyshev.cpp through
UTests/FUKernelTester.hpp:56-170 (DirectDouble.bfma, FOctree of 4 levels,
order 6, FFmmAlgorithm, particles inserted with their index, results read back
per leaf through the indexes and compared with the file's direct results).

Bars: the reference's own gates (relative L2 and RMS of the potential below
9e-3, of the forces below 9e-2, relative energy error below 9e-3); the FP64
values within 1e-12 of the oracle; leaves in Morton order with the oracle's
keys and counts, particles in insertion order inside a leaf.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load():
    from scalfmm_amd.fma import FFmaGenericLoader
    return FFmaGenericLoader(os.path.join(ROOT, "tests", "golden", "DirectDouble.bfma"))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("algo_name", ["FFmmAlgorithm", "FFmmAlgorithmThread"])
def test_kernel_tester_gates(oracle, algo_name):
    from scalfmm_amd import FChebSymKernel
    from scalfmm_amd import octree
    loader = _load()
    d = loader.data
    n = loader.getNumberOfParticles()
    tree = octree.FOctree(4, 2, loader.getBoxWidth(), loader.getCenterOfBox())
    for i in range(5):  # FOctree::insert one by one, then the rest in a block
        tree.insert(d[i, :3], i, d[i, 3])
    tree.insert_many(d[5:, 0], d[5:, 1], d[5:, 2], d[5:, 3], np.arange(5, n))
    kernel = FChebSymKernel(4, loader.getBoxWidth(), loader.getCenterOfBox(), 6)
    getattr(octree, algo_name)(tree, kernel).execute()
    pot = np.zeros(n)
    forces = np.zeros((n, 3))
    energy = 0.0
    seen = []

    def visit(leaf):
        nonlocal energy
        c = leaf.getTargets()
        idx = c.getIndexes()
        pot[idx] = c.getPotentials()
        forces[idx, 0], forces[idx, 1], forces[idx, 2] = c.getForcesX(), c.getForcesY(), c.getForcesZ()
        energy += float(np.dot(c.getPotentials(), c.getPhysicalValues()))
        seen.append((leaf.getMortonIndex(), c.getNbParticles(), idx.copy()))
    tree.forEachLeaf(visit)
    energy_d = float(np.dot(d[:, 4], d[:, 3]))
    acc = [oracle.accuracy(d[:, 4], pot)] + [oracle.accuracy(d[:, 5 + k], forces[:, k]) for k in range(3)]
    print(f"  {algo_name}: pot rel L2 {acc[0]['rel_l2']:.3e}, forces "
          f"{max(a['rel_l2'] for a in acc[1:]):.3e}, energy {abs(energy - energy_d) / energy_d:.3e}")
    assert acc[0]["rel_l2"] < 9e-3 and acc[0]["rms"] < 9e-3
    for a in acc[1:]:
        assert a["rel_l2"] < 9e-2 and a["rms"] < 9e-2
    assert sum(a["rel_l2"] ** 2 for a in acc[1:]) < 9e-2
    assert abs(energy - energy_d) / energy_d < 9e-3
    # the FP64 values are the oracle's
    ref = oracle.fmm(d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(), d[:, 3].copy(), loader.getCenterOfBox(),
                     loader.getBoxWidth(), 4, 6)
    assert _rel(pot, ref[0]) <= 1e-12
    for k in range(3):
        assert _rel(forces[:, k], ref[1 + k]) <= 1e-12
    # leaves: Morton order, the oracle's keys and counts, insertion order inside each leaf
    t = oracle.Tree(d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(), loader.getCenterOfBox(),
                    loader.getBoxWidth(), 4)
    keys = [k for k, _, _ in seen]
    assert keys == sorted(keys)
    assert np.array_equal(np.array(keys), t.cell_keys(3))
    assert sum(c for _, c, _ in seen) == n
    for _, _, idx in seen:
        assert np.all(np.diff(idx) > 0)


def test_results_accumulate():
    """The containers accumulate across executions (as the reference's kernels add into them)."""
    from scalfmm_amd import FChebSymKernel
    from scalfmm_amd.octree import FFmmAlgorithm, FOctree
    loader = _load()
    d = loader.data
    tree = FOctree(4, 2, loader.getBoxWidth(), loader.getCenterOfBox())
    tree.insert_many(d[:, 0], d[:, 1], d[:, 2], d[:, 3])
    algo = FFmmAlgorithm(tree, FChebSymKernel(4, loader.getBoxWidth(), loader.getCenterOfBox(), 5))
    algo.execute()
    once = {}
    tree.forEachLeaf(lambda leaf: once.update(zip(leaf.getTargets().getIndexes(), leaf.getTargets().getPotentials())))
    algo.execute()
    twice = {}
    tree.forEachLeaf(lambda leaf: twice.update(zip(leaf.getTargets().getIndexes(), leaf.getTargets().getPotentials())))
    assert all(twice[i] == 2 * once[i] for i in once)
