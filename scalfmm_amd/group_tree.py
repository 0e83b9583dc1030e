"""The reference's group tree as a caller of the HIP engine, with its particle
blocks in the reference's byte layout (paths relative to its source tree):

  FGroupTree(height, boxWidth, boxCenter, nbElementsPerBlock, particles)
      GroupTree/Core/FGroupTree.hpp:178-268 (the leaf level: particles sorted
      by Morton index, leaves grouped nbElementsPerBlock per block)
  FGroupOfParticles                       GroupTree/Core/FGroupOfParticles.hpp:19-430
  FP2PGroupParticleContainer<double, 1, 1, NVALS>
                                          GroupTree/Core/FP2PGroupParticleContainer.hpp:8-57
  FGroupSeqAlgorithm(&tree, &kernel).execute(ops[, upper, lower])
      GroupTree/Core/FGroupSeqAlgorithm.hpp:61-83, Core/FCoreCommon.hpp:65-94

Each particle block is a byte buffer laid out as the reference allocates it
(a 64-byte header, 24-byte leaf headers, then x, y, z and the NVALS physical
values as arrays padded to 64-byte packs) plus a separate attributes buffer
(potentials and forces).  execute() hands the raw buffers to the engine's
block entry (hfmm_evaluate_particle_blocks), which reads the particles out of
the blocks and adds the results into the attributes buffers.  The cell blocks
of the reference's tree are not built: multipoles and locals live in the
engine, which keeps them between executions on the same particles as the
reference's tree keeps its cells (every pass adds into them), until
resetCells() (the reference's forEachCell(resetToInitialState)).
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .fmm import NEAR_AND_FAR, FChebSymKernel, FmmError

HEADER_BYTES = 64           # sizeof(BlockHeader), alignas(8)
LEAF_HEADER_BYTES = 24      # sizeof(LeafHeader): mindex, nbParticles, offSet
ALIGN_BYTES = 64            # FP2PDefaultAlignement (Utils/FGlobal.hpp:110)
ALIGN_PARTICLES = ALIGN_BYTES // 8

_HEADER = np.dtype([("startingIndex", "<i8"), ("endingIndex", "<i8"), ("numberOfLeavesInBlock", "<i4"),
                    ("pad", "<i4"), ("nbParticlesAllocatedInGroup", "<i8"), ("offsetPosition", "<u8"),
                    ("positionsLeadingDim", "<u8"), ("attributeLeadingDim", "<u8"), ("nbParticlesInGroup", "<i8")])
_LEAF = np.dtype([("mindex", "<i8"), ("nbParticles", "<i8"), ("offSet", "<u8")])
assert _HEADER.itemsize == HEADER_BYTES and _LEAF.itemsize == LEAF_HEADER_BYTES


def round_to_upper_particles(n):
    """RoundToUpperParticles (FGroupOfParticles.hpp:53-62): always adds at
    least one particle, n + (8 - n % 8)."""
    return n + (ALIGN_PARTICLES - (n % ALIGN_PARTICLES))


def morton_keys(x, y, z, box_width, box_center, tree_height):
    """Leaf Morton index of every particle as the group tree computes it:
    FCoordinateComputer::GetCoordinateFromPositionAndCorner
    (Containers/FCoordinateComputer.hpp:15-54, truncation, the upper face in
    the last leaf) and FTreeCoordinate::getMortonIndex (FTreeCoordinate.hpp:91-116,
    bits interleaved x, y, z from the top)."""
    w = float(box_width)
    leafw = w / float(1 << (tree_height - 1))
    coords = []
    for p, c in zip((x, y, z), box_center):
        rel = np.asarray(p, dtype=np.float64) - (float(c) - w / 2)
        if np.any(rel < 0) or np.any(rel > w):
            raise FmmError(_abi.ERR_ARG, "a particle lies outside the box")
        i = (rel / leafw).astype(np.int64)
        i[rel == w] = (1 << (tree_height - 1)) - 1
        coords.append(i)
    key = np.zeros(len(coords[0]), dtype=np.int64)
    for b in range(tree_height - 1):
        for shift, ci in zip((2, 1, 0), coords):
            key |= ((ci >> b) & 1) << (3 * b + shift)
    return key


class FP2PGroupParticleContainer:
    """Views of one leaf's arrays inside a block (getPositions,
    getPhysicalValues, getPotentials, getForces*), all of NVALS vectors."""

    def __init__(self, group, leaf):
        self._g, self._leaf = group, leaf
        self._o = int(group.leaf_headers["offSet"][leaf])
        self._n = int(group.leaf_headers["nbParticles"][leaf])

    def getNbParticles(self):
        return self._n

    def getPositions(self):
        return tuple(self._g._pos[d][self._o:self._o + self._n] for d in range(3))

    def getPhysicalValues(self, idx_vals=0):
        return self._g._pos[3 + idx_vals][self._o:self._o + self._n]

    def _attr(self, a, v):
        return self._g.attributes[a * self._g.nvals + v][self._o:self._o + self._n]

    def getPotentials(self, idx_vals=0):
        return self._attr(0, idx_vals)

    def getForcesX(self, idx_vals=0):
        return self._attr(1, idx_vals)

    def getForcesY(self, idx_vals=0):
        return self._attr(2, idx_vals)

    def getForcesZ(self, idx_vals=0):
        return self._attr(3, idx_vals)


class FGroupOfParticles:
    """One particle block (FGroupOfParticles.hpp:130-200 allocates it)."""

    def __init__(self, starting_index, ending_index, number_of_leaves, nb_particles, nvals=1):
        nalloc = round_to_upper_particles(nb_particles + (ALIGN_PARTICLES - 1) * number_of_leaves)
        off_pos = -(-(HEADER_BYTES + LEAF_HEADER_BYTES * number_of_leaves) // ALIGN_BYTES) * ALIGN_BYTES
        # the reference sizes the block as header + 8 + 24 bytes per leaf + the arrays
        # and places the arrays at the next 64-byte boundary after the leaf headers
        size = max(HEADER_BYTES + (8 + LEAF_HEADER_BYTES) * number_of_leaves + nalloc * (3 + nvals) * 8,
                   off_pos + nalloc * (3 + nvals) * 8)
        self.nvals = int(nvals)
        self.raw = np.zeros(size, dtype=np.uint8)
        self.header = self.raw[:HEADER_BYTES].view(_HEADER)
        self.header["startingIndex"] = starting_index
        self.header["endingIndex"] = ending_index
        self.header["numberOfLeavesInBlock"] = number_of_leaves
        self.header["nbParticlesAllocatedInGroup"] = nalloc
        self.header["offsetPosition"] = off_pos
        self.header["positionsLeadingDim"] = 8 * nalloc
        self.header["attributeLeadingDim"] = 8 * nalloc
        self.header["nbParticlesInGroup"] = nb_particles
        self.leaf_headers = self.raw[HEADER_BYTES:HEADER_BYTES + LEAF_HEADER_BYTES * number_of_leaves].view(_LEAF)
        self.leaf_headers["mindex"] = -1
        arrays = self.raw[off_pos:off_pos + nalloc * (3 + nvals) * 8].view(np.float64)
        self._pos = [arrays[k * nalloc:(k + 1) * nalloc] for k in range(3 + nvals)]
        self.attributes_raw = np.zeros(4 * nvals * nalloc, dtype=np.float64)
        self.attributes = [self.attributes_raw[k * nalloc:(k + 1) * nalloc] for k in range(4 * nvals)]

    def newLeaf(self, mindex, leaf, nb_particles, offset):
        """FGroupOfParticles::newLeaf (:300-312): returns the next leaf's offset."""
        self.leaf_headers["nbParticles"][leaf] = nb_particles
        self.leaf_headers["offSet"][leaf] = offset
        self.leaf_headers["mindex"][leaf] = mindex
        return round_to_upper_particles(offset + nb_particles)

    def getRawBuffer(self):
        return self.raw

    def getBufferSizeInByte(self):
        return self.raw.nbytes

    def getRawAttributesBuffer(self):
        return self.attributes_raw

    def getNumberOfLeavesInBlock(self):
        return int(self.header["numberOfLeavesInBlock"][0])

    def getNbParticlesInGroup(self):
        return int(self.header["nbParticlesInGroup"][0])

    def getStartingIndex(self):
        return int(self.header["startingIndex"][0])

    def getEndingIndex(self):
        return int(self.header["endingIndex"][0])

    def getLeafMortonIndex(self, leaf):
        return int(self.leaf_headers["mindex"][leaf])

    def getLeaf(self, leaf):
        return FP2PGroupParticleContainer(self, leaf)

    def forEachLeaf(self, fn):
        for i in range(self.getNumberOfLeavesInBlock()):
            fn(self.getLeaf(i))


class FGroupTree:
    """The leaf level of the reference's group tree over a particle set:
    particles sorted by leaf Morton index (stable: insertion order inside a
    leaf), then nbElementsPerBlock leaves per particle block."""

    def __init__(self, tree_height, box_width, box_center, nb_elements_per_block, x, y, z, q):
        if nb_elements_per_block < 1:
            raise ValueError("nb_elements_per_block must be >= 1")
        self.tree_height = int(tree_height)
        self.box_width = float(box_width)
        self.box_center = tuple(float(c) for c in box_center)
        self.nb_elements_per_block = int(nb_elements_per_block)
        self._reset_pending = False
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (x, y, z))
        q = np.asarray(q, dtype=np.float64)
        q = q.reshape(1, -1) if q.ndim == 1 else q
        self.nvals = q.shape[0]
        if not (len(x) == len(y) == len(z) == q.shape[1]):
            raise ValueError("positions and physical values must have the same length")
        keys = morton_keys(x, y, z, self.box_width, self.box_center, self.tree_height)
        order = np.argsort(keys, kind="stable")
        self.original_index = order  # position in the blocks -> index in the input
        skeys = keys[order]
        leaf_keys, leaf_start, leaf_count = np.unique(skeys, return_index=True, return_counts=True)
        self.groups = []
        for b0 in range(0, len(leaf_keys), self.nb_elements_per_block):
            lk = leaf_keys[b0:b0 + self.nb_elements_per_block]
            lc = leaf_count[b0:b0 + self.nb_elements_per_block]
            ls = leaf_start[b0:b0 + self.nb_elements_per_block]
            g = FGroupOfParticles(int(lk[0]), int(lk[-1]) + 1, len(lk), int(lc.sum()), self.nvals)
            off = 0
            for i, (k, s, c) in enumerate(zip(lk, ls, lc)):
                src = order[s:s + c]
                for d, a in enumerate((x, y, z)):
                    g._pos[d][off:off + c] = a[src]
                for v in range(self.nvals):
                    g._pos[3 + v][off:off + c] = q[v, src]
                off = g.newLeaf(int(k), i, int(c), off)
            self.groups.append(g)

    def getHeight(self):
        return self.tree_height

    def getBoxWidth(self):
        return self.box_width

    def getBoxCenter(self):
        return self.box_center

    def resetCells(self):
        """tree.forEachCell([](cell){ cell->resetToInitialState(); }): the next
        execution starts from zero multipoles and locals."""
        self._reset_pending = True

    def getNbParticleGroup(self):
        return len(self.groups)

    def getParticleGroup(self, i):
        return self.groups[i]

    def forEachLeaf(self, fn):
        for g in self.groups:
            g.forEachLeaf(fn)

    def results(self, idx_vals=0):
        """Potentials and forces gathered back to the input order."""
        out = [np.zeros(len(self.original_index)) for _ in range(4)]
        pos = 0
        for g in self.groups:
            for leaf in range(g.getNumberOfLeavesInBlock()):
                o = int(g.leaf_headers["offSet"][leaf])
                c = int(g.leaf_headers["nbParticles"][leaf])
                dst = self.original_index[pos:pos + c]
                for a in range(4):
                    out[a][dst] = g.attributes[a * self.nvals + idx_vals][o:o + c]
                pos += c
        return tuple(out)


class FGroupSeqAlgorithm:
    """execute(ops) / execute(ops, upperWorkingLevel, lowerWorkingLevel) of the
    group algorithms over the tree's particle blocks."""

    def __init__(self, tree: FGroupTree, kernel: FChebSymKernel):
        if (tree.tree_height, tree.box_width, tree.box_center) != (kernel.tree_height, kernel.box_width,
                                                                   kernel.box_center):
            raise ValueError("tree and kernel disagree on height / box")
        self.tree = tree
        self.kernel = kernel
        kernel.engine.set_cell_mode(_abi.CELLS_KEPT)

    def execute(self, operations=NEAR_AND_FAR, upper_level=2, lower_level=None):
        t = self.tree
        if t._reset_pending:
            self.kernel.engine.reset_cells()
            t._reset_pending = False
        self.kernel.engine.evaluate_particle_blocks([g.raw for g in t.groups], [g.attributes_raw for g in t.groups],
                                                    t.nvals, operations, upper_level, lower_level)
        return t


FGroupTaskAlgorithm = FGroupSeqAlgorithm
