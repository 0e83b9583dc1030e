"""The classic-tree caller of the reference over the HIP engine.

The reference's examples and unit tests drive the Chebyshev kernel through
the classic octree (paths relative to its source tree):

  FOctree(height, subHeight, boxWidth, boxCenter)        Containers/FOctree.hpp:95-1448
  tree.insert(position, index, physicalValue)            Containers/FOctree.hpp:174-185
  FP2PParticleContainerIndexed / FSimpleLeaf             Kernels/P2P/FP2PParticleContainerIndexed.hpp,
                                                         Components/FSimpleLeaf.hpp
  FFmmAlgorithm(&tree, &kernels).execute(ops)            Core/FFmmAlgorithm.hpp:66-110
  FFmmAlgorithmThread(&tree, &kernels).execute(ops)      Core/FFmmAlgorithmThread.hpp:118-495
  tree.forEachLeaf([](LeafClass* leaf){ ... })           Containers/FOctree.hpp

Particles are kept in insertion order inside a leaf and the leaves are
visited in Morton order, as in the reference; results accumulate (+=) into
the leaf containers, which start at zero on insertion.  The tree structure,
the lists and every pass are the engine's (built on the device); this module
only moves particles in and results out.
"""
from __future__ import annotations

import numpy as np

from .fmm import NEAR_AND_FAR, Engine, FChebSymKernel


class FP2PParticleContainerIndexed:
    """Particles of one leaf: positions, physical values, potentials, forces and
    the index each particle was inserted with."""

    def __init__(self, x, y, z, q, pot, fx, fy, fz, idx):
        self._pos = (x, y, z)
        self._q, self._pot, self._f, self._idx = q, pot, (fx, fy, fz), idx

    def getNbParticles(self):
        return len(self._q)

    def getPositions(self):
        return self._pos

    def getPhysicalValues(self):
        return self._q

    def getPotentials(self):
        return self._pot

    def getForcesX(self):
        return self._f[0]

    def getForcesY(self):
        return self._f[1]

    def getForcesZ(self):
        return self._f[2]

    def getIndexes(self):
        return self._idx


class FSimpleLeaf:
    """A leaf whose sources and targets are the same container."""

    def __init__(self, key, container):
        self.key = int(key)
        self._c = container

    def getTargets(self):
        return self._c

    def getSrc(self):
        return self._c

    def getMortonIndex(self):
        return self.key


class FOctree:
    def __init__(self, height, sub_height, box_width, box_center):
        self.height = int(height)
        self.sub_height = int(sub_height)  # the reference's sub-octree storage detail; no effect here
        self.box_width = float(box_width)
        self.box_center = tuple(float(c) for c in box_center)
        self._chunks = []
        self._arrays = None
        self._layout = None

    # ---- particles ------------------------------------------------------------
    def insert(self, position, index, physical_value):
        """One particle (FOctree::insert with FP2PParticleContainerIndexed's
        push(position, index, physicalValue))."""
        p = np.asarray(position, dtype=np.float64).reshape(3)
        self.insert_many(p[0:1], p[1:2], p[2:3], [physical_value], [index])

    def insert_many(self, x, y, z, q, indexes=None):
        """Many particles at once, in order (indexes default to the running count)."""
        x, y, z, q = (np.asarray(a, dtype=np.float64).reshape(-1) for a in (x, y, z, q))
        start = self.getNbParticles()
        idx = np.arange(start, start + len(x)) if indexes is None else np.asarray(indexes, dtype=np.int64)
        if not (len(x) == len(y) == len(z) == len(q) == len(idx)):
            raise ValueError("positions, values and indexes must have the same length")
        self._chunks.append((x, y, z, q, idx))
        self._arrays = None
        self._layout = None

    def getNbParticles(self):
        if self._arrays is not None:
            return len(self._arrays[3])
        return int(sum(len(c[3]) for c in self._chunks))

    def _flat(self):
        if self._arrays is None:
            cols = list(zip(*self._chunks)) if self._chunks else [[np.zeros(0)]] * 4 + [[np.zeros(0, np.int64)]]
            x, y, z, q, idx = (np.concatenate(c) for c in cols)
            n = len(q)
            self._arrays = [x, y, z, q, idx, np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)]
        return self._arrays

    def getHeight(self):
        return self.height

    def getBoxWidth(self):
        return self.box_width

    def getBoxCenter(self):
        return self.box_center

    # ---- leaves -----------------------------------------------------------------
    def _set_layout(self, engine):
        """Leaf grouping from the engine's device tree: sorted permutation (stable,
        so insertion order inside a leaf), leaf ranges and leaf keys."""
        keys, perm = engine.sorted_keys()
        start, count = engine.leaf_ranges()
        self._layout = (perm, start, count, keys[start])

    def forEachLeaf(self, fn):
        if self._layout is None:
            e = Engine(self.height, self.box_width, self.box_center, 5)  # tree only
            x, y, z, q = self._flat()[:4]
            e.set_particles(x, y, z, q)
            e.build_tree()
            self._set_layout(e)
            e.close()
        x, y, z, q, idx, pot, fx, fy, fz = self._flat()
        perm, start, count, lkeys = self._layout
        for s, c, k in zip(start, count, lkeys):
            sel = perm[s:s + c]
            cont = FP2PParticleContainerIndexed(x[sel], y[sel], z[sel], q[sel], pot[sel], fx[sel], fy[sel], fz[sel],
                                                idx[sel])
            fn(FSimpleLeaf(k, cont))


class FFmmAlgorithm:
    """execute(ops) over a classic octree with an FChebSymKernel."""

    def __init__(self, tree: FOctree, kernel: FChebSymKernel):
        if (tree.height, tree.box_width, tree.box_center) != (kernel.tree_height, kernel.box_width,
                                                              kernel.box_center):
            raise ValueError("tree and kernel disagree on height / box")
        self.tree = tree
        self.kernel = kernel

    def execute(self, operations=NEAR_AND_FAR):
        e = self.kernel.engine
        arr = self.tree._flat()
        e.set_particles(*arr[:4])
        e.build_tree()
        e.evaluate(operations)
        self.tree._set_layout(e)
        for acc, val in zip(arr[5:], e.results()):
            acc += val  # the containers accumulate, as the kernels' P2P / L2P do
        return self.tree


class FFmmAlgorithmThread(FFmmAlgorithm):
    """The threaded classic algorithm: the same passes (all of them run on the GPU)."""
