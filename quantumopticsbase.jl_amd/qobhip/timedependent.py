"""TimeDependentSum (src/time_dependent_operator.jl:97-300): a LazySum whose
coefficients are numbers or functions of time, with an internal clock.

mul! on it is mul! on its static LazySum at the current time (:274-277);
set_time! evaluates the coefficients and writes them into the LazySum
(update_static_coefficients!, :286-297).  Here the evaluated coefficients go
to the device and the LazySum's plans are rewritten there
(qob_lazysum_set_factors_d): no plan is rebuilt and nothing waits for the
device.  ``coefficient_table`` evaluates the coefficients over a time grid
once, so a captured graph of steps can take row k at step k.
"""
from __future__ import annotations

import numbers
from typing import Callable, Sequence, Union

import numpy as np

from ._lib import ArgumentError
from .core import AbstractOperator, DeviceArray, LazySum

Coef = Union[numbers.Number, Callable[[float], complex]]


def eval_coefficient(c: Coef, t: float) -> complex:
    """eval_coefficient (:281-282): call a function, keep a number."""
    return complex(c) if isinstance(c, numbers.Number) else complex(c(t))


class TimeDependentSum(AbstractOperator):
    """TimeDependentSum(coefficients, operators; init_time) or
    TimeDependentSum(lazysum; init_time) or TimeDependentSum(pairs...)."""

    def __init__(self, coefficients, operators=None, init_time: float = 0.0, *, _static=None):
        if isinstance(coefficients, LazySum) and operators is None:
            static = coefficients
            coefficients = list(static.factors)
        elif isinstance(coefficients, (list, tuple)) and operators is None:
            pairs = list(coefficients)
            coefficients = [c for c, _ in pairs]
            static = LazySum([0.0] * len(pairs), tuple(o for _, o in pairs))
        else:
            ops = tuple(operators)
            if len(ops) != len(coefficients):
                raise ArgumentError("TimeDependentSum: coefficients and operators have different lengths")
            static = _static or LazySum([0.0] * len(ops), ops)
        self.coefficients = list(coefficients)
        self.static_op = static
        self.basis_l, self.basis_r = static.basis_l, static.basis_r
        self.current_time = None
        self._dbuf = None
        self.set_time_(init_time)

    # reference API ---------------------------------------------------
    def static_operator(self) -> LazySum:
        return self.static_op

    def eval_coefficients(self, t: float):
        return [eval_coefficient(c, t) for c in self.coefficients]

    def is_const(self) -> bool:
        return all(isinstance(c, numbers.Number) for c in self.coefficients)

    def set_time_(self, t: float) -> "TimeDependentSum":
        """set_time! (:154-162): evaluate the coefficients at t and make them
        current in the static LazySum (on the device when supported)."""
        if self.current_time == t:
            return self
        self.current_time = t
        vals = self.eval_coefficients(t)
        self.static_op.factors = vals
        if self.static_op._h is None:
            self.static_op.set_factors(vals)  # handle created lazily from factors
            return self
        n = len(vals)
        if n == 0:
            return self
        if self._dbuf is None:
            self._dbuf = DeviceArray(n)
        host = np.asarray(vals, dtype=np.complex128)
        from ._lib import call
        from .core import context
        import ctypes as C
        call("qob_memcpy_htod", context().h, self._dbuf.ptr, host.ctypes.data_as(C.c_void_p), 16 * n, None)
        try:
            self.static_op.set_factors_device(self._dbuf)
        except Exception:
            self.static_op.set_factors(vals)  # terms the device refresh does not cover
        return self

    def __call__(self, t: float) -> "TimeDependentSum":
        """op(t) = set_time!(copy(op), t) (:55)."""
        return TimeDependentSum(list(self.coefficients), self.static_op.operators, t)

    def dagger(self) -> "TimeDependentSum":
        """dagger (:197-203): conjugated coefficients, adjoint operators."""
        from .core import dagger as _dagger

        def conj_coeff(c):
            if isinstance(c, numbers.Number):
                return complex(c).conjugate()
            return lambda t, c=c: complex(c(t)).conjugate()
        return TimeDependentSum([conj_coeff(c) for c in self.coefficients],
                                tuple(_dagger(o) for o in self.static_op.operators), self.current_time)

    # device-side helpers -----------------------------------------------
    def coefficient_table(self, times: Sequence[float]) -> DeviceArray:
        """Coefficients evaluated at every time of a grid, one row of
        len(operators) complex doubles per time, on the device."""
        tab = np.array([self.eval_coefficients(t) for t in times], dtype=np.complex128)
        return DeviceArray.from_numpy(tab.reshape(-1))

    def set_time_row(self, table: DeviceArray, k: int, stream=None) -> None:
        """Make row k of a coefficient table current (stream-ordered,
        capturable).  ``current_time`` is not tracked for such updates."""
        import ctypes as C
        n = len(self.coefficients)
        if k < 0 or (k + 1) * n > table.n:
            raise ArgumentError("coefficient table row out of range")
        ptr = C.c_void_p(table.ptr.value + 16 * n * k)
        self.static_op.set_factors_device(ptr, stream)
        self.current_time = None

    def _handle(self):
        return self.static_op._handle()
