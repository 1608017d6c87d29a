"""qobhip: QuantumOpticsBase.jl's mul! hot path on one MI355X.

Host-side mirror of the reference API over the C ABI of libqob_hip.so
(include/qob_hip.h).  See core.py for the type and function list.
"""
from ._lib import (QOB_C32, QOB_C64, QOB_MODE_FUSED, QOB_MODE_PLAIN, ArgumentError, DimensionMismatch,
                   IncompatibleBases, MethodError, QobError)
from .core import (AbstractOperator, Basis, Bra, CompositeBasis, Context, DenseOperator, DeviceArray, EyeOperator,
                   FockBasis, GenericBasis, Ket, LazyProduct, LazySum, LazyTensor, NLevelBasis, Operator, SparseOperator,
                   SpinBasis, axpy_, context, dagger, dot, expect, identityoperator, mul_, norm, normalize_,
                   synchronize, tensor, tr, variance)
from .spin import sigmam, sigmap, sigmax, sigmay, sigmaz
from .fock import create, destroy, number
from .batch import Batch, Graph
from .timedependent import TimeDependentSum
from .superop import SuperOperator, identitysuperoperator, liouvillian, spost, spre, sprepost

__all__ = [n for n in dir() if not n.startswith("_")]
