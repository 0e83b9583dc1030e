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

  FGroupOfCells with FChebCellPOD cells  GroupTree/Core/FGroupOfCells.hpp:18-370,
                                          GroupTree/Chebyshev/FChebCellPOD.hpp:11-112,
                                          GroupTree/Core/FGroupTree.hpp:288-356 (upper levels)

Each particle block is a byte buffer laid out as the reference allocates it
(a 64-byte header, 24-byte leaf headers, then x, y, z and the NVALS physical
values as arrays padded to 64-byte packs) plus a separate attributes buffer
(potentials and forces).  execute() hands the raw buffers to the engine's
block entry (hfmm_evaluate_particle_blocks), which reads the particles out of
the blocks and adds the results into the attributes buffers.

With an order, the tree also holds the reference's cell blocks of levels
height-1 .. 1: per block a raw buffer (24-byte header, Morton indexes, one
24-byte FBasicCellPOD per cell) and the multipole and local buffers
(nvals x 2 p^3 doubles per cell, FChebCellPOD's layout).  execute() copies the
blocks' expansions of levels 2 .. height-1 into the engine's kept cells
(hfmm_set_cell_blocks), evaluates, and copies them back
(hfmm_get_cell_blocks), so the blocks hold what the reference's cells hold
after its passes; a caller may read or seed them between executions.
Without an order the engine alone keeps the cells between executions on the
same particles, as the reference's tree keeps its cells (every pass adds into
them), until resetCells() (the reference's forEachCell(resetToInitialState)).
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


_CELL_HEADER = np.dtype([("startingIndex", "<i8"), ("endingIndex", "<i8"), ("numberOfCellsInBlock", "<i4"),
                         ("pad", "<i4")])
_CELL_POD = np.dtype([("mortonIndex", "<i8"), ("coordinates", "<i4", (3,)), ("level", "<i4")])
assert _CELL_HEADER.itemsize == 24 and _CELL_POD.itemsize == 24


def morton_to_coordinate(key, level):
    """FTreeCoordinate::setPositionFromMorton (FTreeCoordinate.hpp:126-142)."""
    c = [0, 0, 0]
    for b in range(level):
        for d, shift in enumerate((2, 1, 0)):
            c[d] |= ((int(key) >> (3 * b + shift)) & 1) << b
    return c


class FGroupOfCells:
    """One cell block (FGroupOfCells.hpp:125-168 allocates it) with the
    Chebyshev cells' multipole and local buffers."""

    def __init__(self, starting_index, ending_index, keys, level, order, nvals=1):
        n = len(keys)
        self.raw = np.zeros(_CELL_HEADER.itemsize + 8 * n + _CELL_POD.itemsize * n, dtype=np.uint8)
        self.header = self.raw[:24].view(_CELL_HEADER)
        self.header["startingIndex"] = starting_index
        self.header["endingIndex"] = ending_index
        self.header["numberOfCellsInBlock"] = n
        self.cell_indexes = self.raw[24:24 + 8 * n].view(np.int64)
        self.cell_indexes[:] = keys
        self.cells = self.raw[24 + 8 * n:].view(_CELL_POD)
        self.cells["mortonIndex"] = keys
        # the particle constructor sets the coordinates, not the level (FGroupTree.hpp:257-263, 341-347)
        self.cells["coordinates"] = [morton_to_coordinate(k, level) for k in keys]
        self.vector_size = 2 * order ** 3  # FChebCellPOD::VectorSize
        self.nvals = int(nvals)
        self.multipoles = np.zeros(n * self.nvals * self.vector_size)
        self.locals = np.zeros(n * self.nvals * self.vector_size)

    def getRawBuffer(self):
        return self.raw

    def getRawMultipoleBuffer(self):
        return self.multipoles

    def getRawLocalBuffer(self):
        return self.locals

    def getNumberOfCellsInBlock(self):
        return int(self.header["numberOfCellsInBlock"][0])

    def getStartingIndex(self):
        return int(self.header["startingIndex"][0])

    def getEndingIndex(self):
        return int(self.header["endingIndex"][0])

    def getCellMortonIndex(self, i):
        return int(self.cell_indexes[i])

    def getMultipole(self, i, rhs=0):
        """FChebCellPOD::getMultipole(rhs) of cell i (a view)."""
        o = (i * self.nvals + rhs) * self.vector_size
        return self.multipoles[o:o + self.vector_size]

    def getLocal(self, i, rhs=0):
        o = (i * self.nvals + rhs) * self.vector_size
        return self.locals[o:o + self.vector_size]


def _cell_blocks(keys, level, nb_elements_per_block, order, nvals):
    return [FGroupOfCells(int(k[0]), int(k[-1]) + 1, k, level, order, nvals)
            for k in (keys[b:b + nb_elements_per_block] for b in range(0, len(keys), nb_elements_per_block))]


class FGroupTree:
    """The leaf level of the reference's group tree over a particle set:
    particles sorted by leaf Morton index (stable: insertion order inside a
    leaf), then nbElementsPerBlock leaves per particle block."""

    def __init__(self, tree_height, box_width, box_center, nb_elements_per_block, x, y, z, q, order=None):
        if nb_elements_per_block < 1:
            raise ValueError("nb_elements_per_block must be >= 1")
        self.tree_height = int(tree_height)
        self.box_width = float(box_width)
        self.box_center = tuple(float(c) for c in box_center)
        self.nb_elements_per_block = int(nb_elements_per_block)
        self.order = None if order is None else int(order)  # of the Chebyshev cells
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
        # cell blocks: the leaves as the particle blocks group them, then per
        # level the distinct parent keys, nbElementsPerBlock per block
        # (FGroupTree.hpp:288-356), levels height-1 down to 1
        self.cell_groups = {}
        if self.order is not None:
            level_keys = leaf_keys.astype(np.int64)
            for level in range(self.tree_height - 1, 0, -1):
                self.cell_groups[level] = _cell_blocks(level_keys, level, self.nb_elements_per_block,
                                                       self.order, self.nvals)
                level_keys = np.unique(level_keys >> 3)

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
        for blocks in self.cell_groups.values():
            for g in blocks:
                g.multipoles[:] = 0.0
                g.locals[:] = 0.0

    def getNbCellGroupAtLevel(self, level):
        return len(self.cell_groups.get(level, ()))

    def getCellGroup(self, level, i):
        return self.cell_groups[level][i]

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
        e = self.kernel.engine
        if t._reset_pending:
            e.reset_cells()
            t._reset_pending = False
        blocks = [g.raw for g in t.groups]
        attrs = [g.attributes_raw for g in t.groups]
        levels = [lv for lv in sorted(t.cell_groups) if 2 <= lv < t.tree_height]
        if levels:
            if t.order != self.kernel.order:
                raise ValueError("tree and kernel disagree on the order")
            # the engine's tree over the blocks' particles (no pass), then the
            # blocks' cells into its kept cells (the symmetric scheme; the
            # others keep their cells in the engine)
            e.evaluate_particle_blocks(blocks, attrs, t.nvals, 0)
            for lv in levels if e.params.m2l_scheme == _abi.M2L_SYMMETRIC else ():
                cg = t.cell_groups[lv]
                e.set_cell_blocks(lv, [g.raw for g in cg], [g.multipoles for g in cg], [g.locals for g in cg],
                                  t.nvals)
        e.evaluate_particle_blocks(blocks, attrs, t.nvals, operations, upper_level, lower_level)
        for lv in levels:
            cg = t.cell_groups[lv]
            e.get_cell_blocks(lv, [g.raw for g in cg], [g.multipoles for g in cg], [g.locals for g in cg], t.nvals)
        return t


FGroupTaskAlgorithm = FGroupSeqAlgorithm
