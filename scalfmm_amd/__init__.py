"""MI355X-native Chebyshev 1/r FMM engine (HIP/CDNA4) behind the reference's
FChebSymKernel / FGroupTree / FGroupTaskAlgorithm interface."""
from ._abi import PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED_HILO, HipLibraryMissing, lib  # noqa: F401
from .fmm import (  # noqa: F401
    Engine, FChebDenseKernel, FChebKernel, FChebSymKernel, FUnifKernel, FFmmFarField, FFmmL2L, FFmmL2P, FFmmM2L, FFmmM2M, FFmmNearAndFarFields,
    FFmmNearField, FFmmP2M, FFmmP2P, FGroupTaskAlgorithm, FGroupTree, FmmError,
)
from .batch import BatchEngine, default_height  # noqa: F401
from .octree import FFmmAlgorithm, FFmmAlgorithmThread, FOctree  # noqa: F401
