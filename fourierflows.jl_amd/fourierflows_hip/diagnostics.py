"""Diagnostics (src/diagnostics.jl): ``Diagnostic``, ``increment``,
``update``, ``extend``, and the built-in energy / enstrophy / tracer
variance functions computed by libffhip's compensated FP64 reductions.

``stepforward(prob, diags, nsteps)`` increments the diagnostics after every
step as the reference does (src/timesteppers.jl:28-35); with a built-in
operator the steps between two diagnostic updates run as one library call,
so a diagnostic every ``freq`` steps costs one host synchronisation per
update instead of one per step.
"""

from __future__ import annotations

import math

import numpy as np


class Diagnostic:
    """``Diagnostic(calc, prob; freq=1, nsteps=100, ndata=ceil((nsteps+1)/freq))``
    (src/diagnostics.jl:8-50): ``calc(prob)`` saved every ``freq`` steps."""

    def __init__(self, calc, prob, freq=1, nsteps=100, ndata=None):
        self.calc, self.prob, self.freq = calc, prob, int(freq)
        ndata = math.ceil((nsteps + 1) / freq) if ndata is None else int(ndata)
        first = calc(prob)
        self.N = ndata
        self.data = [None] * ndata
        self.t = np.zeros(ndata)
        self.steps = np.zeros(ndata, dtype=np.int64)
        self.data[0] = first
        self.t[0] = prob.clock.t
        self.steps[0] = prob.clock.step
        self.i = 1  # number of entries saved (the reference's 1-based index of the last one)

    def __call__(self):
        return self.calc(self.prob)

    def __getitem__(self, key):
        """``d[:t]``, ``d[:data]``, ``d[:steps]`` (src/diagnostics.jl:107-108), or data[idx]."""
        if isinstance(key, str):
            v = getattr(self, key)
            return v[: self.i]
        return self.data[key]

    def __repr__(self):
        """Same fields as the reference's ``show`` (src/diagnostics.jl:117-126)."""
        name = getattr(self.calc, "__name__", repr(self.calc))
        n = len(self.steps)
        return (f"Diagnostic\n  ├─── calc: {name}\n  ├─── prob: {type(self.prob).__name__}\n"
                f"  ├─── data: {n}-element list\n  ├────── t: {n}-element float64 array\n"
                f"  ├── steps: {n}-element int64 array\n  ├─── freq: {self.freq}\n  └────── i: {self.i}")


def extend(diag, n=None):
    """``extend!(diag[, n])`` (src/diagnostics.jl:52-69): grow the storage by n (default N)."""
    n = diag.N if n is None else int(n)
    diag.data.extend([None] * n)
    diag.t = np.concatenate([diag.t, np.zeros(n)])
    diag.steps = np.concatenate([diag.steps, np.zeros(n, dtype=np.int64)])


def update(diag, i):
    """``update!(diag, i)`` (src/diagnostics.jl:71-84), i 1-based."""
    if i > len(diag.steps):
        extend(diag)
    diag.data[i - 1] = diag()
    diag.t[i - 1] = diag.prob.clock.t
    diag.steps[i - 1] = diag.prob.clock.step
    diag.i = i


def _as_list(diags):
    return list(diags) if isinstance(diags, (list, tuple)) else [diags]


def increment(diags):
    """``increment!(diag or diags)`` (src/diagnostics.jl:86-104)."""
    for d in _as_list(diags):
        if d.prob.clock.step % d.freq == 0:
            update(d, d.i + 1)


# built-in diagnostics (compensated FP64 Parseval sums in libffhip)
def energy(prob):
    """2D vorticity: 1/2 <|u|^2> over the domain."""
    return prob.eqn.calcN.diagnostics(prob.sol)[0]


def enstrophy(prob):
    """2D vorticity: 1/2 <zeta^2>."""
    return prob.eqn.calcN.diagnostics(prob.sol)[1]


def tracer_variance(prob):
    """3D tracer: 1/2 <c^2>."""
    return prob.eqn.calcN.diagnostics(prob.sol)[0]


def tracer_mean(prob):
    """3D tracer: <c>."""
    return prob.eqn.calcN.diagnostics(prob.sol)[1]
