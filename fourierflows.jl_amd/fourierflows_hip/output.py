"""On-disk output (src/output.jl): ``Output``, ``saveproblem``,
``saveoutput``, ``savediagnostic``, ``withoutjld2``/``uniquepath``.

FourierFlows writes JLD2 (HDF5) files.  Neither JLD2 nor an HDF5 library is
available to this package, so the same group layout -- ``snapshots/t/<step>``,
``snapshots/<field>/<step>``, ``grid/Lx``, ``eqn/L``, ``params/<name>``,
``clock/dt``, ``diagnostics/<name>/{steps,t,data}`` -- is written as an npz
archive: a zip of ``.npy`` members named by the group path, appended in place.
``numpy.load(path)`` reads it (keys are the group paths) and ``load(path)``
returns nested dicts indexed like the reference's files
(``file["snapshots"]["c"]["0"]``).

``saveoutput(out)`` copies device fields to the host and writes them before
it returns, as the reference does.  ``saveoutput(out, background=True)``
instead starts the device-to-host copies on the current stream into pinned
host buffers and leaves the file write to a writer thread, so the caller can
launch the next steps at once; ``out.flush()`` waits for pending writes.

Every write to one path holds that path's lock, and a synchronous write
(``saveoutput``, ``saveproblem``, ``savediagnostic``) first waits for the
background snapshots already queued for its path, so the archive receives
whole members in call order.  Writers still pending when the interpreter
exits are flushed and closed then.
"""

from __future__ import annotations

import atexit
import io
import os
import queue
import threading
import types
import weakref
import zipfile

import numpy as np


def withoutjld2(path):
    """src/output.jl:32: the path without a trailing ``.jld2`` (or ``.npz``)."""
    for ext in (".jld2", ".npz"):
        if len(path) > len(ext) and path.endswith(ext):
            return path[: -len(ext)]
    return path


def uniquepath(path):
    """src/output.jl:40-53: ``path`` with ``_n`` appended (n = 1, 2, ...) until no
    such file exists."""
    ext = ".npz" if path.endswith(".npz") else ".jld2"
    n = 1
    if os.path.isfile(path):
        path = withoutjld2(path) + f"_{n}{ext}"
    while os.path.isfile(path):
        n += 1
        path = withoutjld2(path)[: -len(f"_{n - 1}")] + f"_{n}{ext}"
    return path


def _host(value):
    """A host numpy value for a device tensor, array or scalar."""
    if hasattr(value, "detach") and hasattr(value, "cpu"):
        return value.detach().cpu().numpy()
    return np.asarray(value)


def _npy_bytes(arr):
    buf = io.BytesIO()
    np.lib.format.write_array(buf, np.asarray(arr), allow_pickle=False)
    return buf.getvalue()


_GUARD = threading.Lock()
_LOCKS = {}    # absolute path -> lock held by every write to that file
_WRITERS = {}  # absolute path -> weak references to Outputs with a background writer


def _lock_for(path):
    key = os.path.abspath(path)
    with _GUARD:
        lk = _LOCKS.get(key)
        if lk is None:
            lk = _LOCKS[key] = threading.Lock()
    return lk


def _pending_writers(path):
    key = os.path.abspath(path)
    with _GUARD:
        refs = _WRITERS.get(key, [])
        outs = [r() for r in refs]
        _WRITERS[key] = [r for r, o in zip(refs, outs) if o is not None]
    return [o for o in outs if o is not None]


def _flush_pending(path):
    """Wait for the background snapshots queued for ``path`` (in call order)."""
    for o in _pending_writers(path):
        o.flush()


@atexit.register
def _close_all():
    with _GUARD:
        outs = [r() for refs in _WRITERS.values() for r in refs]
    for o in outs:
        if o is not None:
            try:
                o.close()
            except Exception:  # the process is ending; nothing left to report to
                pass


def _write(path, items):
    """Append (key, host value) pairs; a key already in the file is an error, as
    a second write to one JLD2 dataset is."""
    with _lock_for(path):
        mode = "a" if os.path.isfile(path) else "w"
        with zipfile.ZipFile(path, mode, compression=zipfile.ZIP_STORED) as z:
            have = set(z.namelist())
            for key, value in items:
                name = key + ".npy"
                if name in have:
                    raise KeyError(f"{key!r} is already in {path}")
                z.writestr(name, _npy_bytes(value))
                have.add(name)


def load(path):
    """The archive as nested dicts: ``load(p)["snapshots"]["c"]["0"]``."""
    out = {}
    with np.load(path, allow_pickle=False) as f:
        for key in f.files:
            node = out
            parts = key.split("/")
            for part in parts[:-1]:
                node = node.setdefault(part, {})
            v = f[key]
            node[parts[-1]] = v.item() if v.shape == () else v
    return out


class Output:
    """``Output(prob, path, (name, func)...)`` (src/output.jl:8-30): fields
    ``func(prob)`` saved to ``path`` (made unique, ending ``.npz``)."""

    def __init__(self, prob, path, *fieldtuples):
        if len(fieldtuples) == 1 and isinstance(fieldtuples[0], dict):
            fields = dict(fieldtuples[0])
        else:
            fields = {str(name): func for name, func in fieldtuples}
        self.prob = prob
        self.path = uniquepath(withoutjld2(path) + ".npz")
        self.fields = fields
        self._queue = None
        self._thread = None
        self._error = None

    def __getitem__(self, key):
        """``getindex(out, key)`` (src/output.jl:55): the field computed now."""
        return self.fields[key](self.prob)

    def __repr__(self):
        return (f"Output\n  ├──── prob: {type(self.prob).__name__}\n  ├──── path: {self.path}\n"
                f"  └── fields: {sorted(self.fields)}")

    # background writer
    def _worker(self):
        while True:
            job = self._queue.get()
            if job is None:
                self._queue.task_done()
                return
            event, items = job
            try:
                if event is not None:
                    event.synchronize()
                _write(self.path, [(k, v.numpy() if hasattr(v, "numpy") else v) for k, v in items])
            except Exception as e:  # reported by flush()
                self._error = e
            self._queue.task_done()

    def _submit(self, event, items):
        if self._thread is None:
            self._queue = queue.Queue()
            self._thread = threading.Thread(target=self._worker, daemon=True)
            self._thread.start()
            with _GUARD:
                _WRITERS.setdefault(os.path.abspath(self.path), []).append(weakref.ref(self))
        self._queue.put((event, items))

    def flush(self):
        """Wait until every background write has reached the file."""
        if self._queue is not None:
            self._queue.join()
        if self._error is not None:
            e, self._error = self._error, None
            raise e

    def close(self):
        self.flush()
        if self._thread is not None:
            self._queue.put(None)
            self._thread.join()
            self._thread = None


def saveoutput(out, background=False):
    """``saveoutput(out)`` (src/output.jl:62-73): ``snapshots/t/<step>`` and
    ``snapshots/<field>/<step>`` for every field of ``out``."""
    clock = out.prob.clock
    step = int(clock.step)
    items = [(f"snapshots/t/{step}", np.float64(clock.t))]
    if not background:
        for name in out.fields:
            items.append((f"snapshots/{name}/{step}", _host(out[name])))
        _flush_pending(out.path)
        _write(out.path, items)
        return None
    import torch

    event = None
    for name in out.fields:
        v = out[name]
        if isinstance(v, torch.Tensor) and v.is_cuda:
            host = torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
            host.copy_(v, non_blocking=True)
            if event is None:
                event = torch.cuda.Event()
            items.append((f"snapshots/{name}/{step}", host))
        else:
            items.append((f"snapshots/{name}/{step}", _host(v)))
    if event is not None:
        event.record()
    out._submit(event, items)
    return None


def _grid_items(grid):
    from .domains import OneDGrid, ThreeDGrid, TwoDGrid

    if isinstance(grid, OneDGrid):
        names = ["nx", "Lx", "x"]
    elif isinstance(grid, TwoDGrid):
        names = ["nx", "ny", "Lx", "Ly", "x", "y"]
    elif isinstance(grid, ThreeDGrid):
        names = ["nx", "ny", "nz", "Lx", "Ly", "Lz", "x", "y", "z"]
    else:
        names = []
    return [(f"grid/{n}", _host(getattr(grid, n))) for n in names]


def _param_items(params):
    """src/output.jl:101-114: every field of params except functions and FFT plans."""
    if params is None:
        return []
    from .domains import Plan

    fields = vars(params) if not isinstance(params, dict) else params
    items = []
    for name, value in fields.items():
        if callable(value) or isinstance(value, (Plan, types.ModuleType)):
            continue
        try:
            items.append((f"params/{name}", _host(value)))
        except (TypeError, ValueError):
            continue
    return items


def saveproblem(prob, filename=None):
    """``saveproblem(prob, filename)`` / ``saveproblem(out)`` (src/output.jl:129-142):
    eqn (L, dims, T), clock (dt), grid and params."""
    if isinstance(prob, Output):
        prob, filename = prob.prob, prob.path
    eqn = prob.eqn
    items = [("eqn/L", _host(eqn.L)), ("eqn/dims", np.asarray(eqn.dims, dtype=np.int64)),
             ("eqn/T", np.asarray(str(np.dtype(eqn.T))))]
    items.append(("clock/dt", np.float64(prob.clock.dt)))
    items += _grid_items(prob.grid)
    items += _param_items(prob.params)
    _flush_pending(filename)
    _write(filename, items)


def savediagnostic(diagnostic, diagname, filename):
    """``savediagnostic(diag, name, filename)`` (src/output.jl:150-158): steps, t
    and data up to the diagnostic's last saved entry."""
    i = diagnostic.i
    data = [_host(v) for v in diagnostic.data[:i]]
    _flush_pending(filename)
    _write(filename, [(f"diagnostics/{diagname}/steps", np.asarray(diagnostic.steps[:i], dtype=np.int64)),
                      (f"diagnostics/{diagname}/t", np.asarray(diagnostic.t[:i], dtype=np.float64)),
                      (f"diagnostics/{diagname}/data", np.stack(data) if data else np.zeros(0))])
