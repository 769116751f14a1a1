"""Solution savers of ImmersedBodies.jl (src/io/io.jl:72-179, src/io/hdf5.jl) for
``solve`` / ``solve_``.

* ``ArrayValue``, ``MultiDomainValue``, ``TimestepRange``, ``SolutionValues``,
  ``SolutionSaver``, ``SaveNothing``                      io/io.jl:72-179
* ``fluid_group``, ``body_group``                          io/hdf5.jl:21-50
* ``SaveNpz``: the layout of ``SaveHDF5`` (io/hdf5.jl:1-356) in a ``.npz`` file

HDF5 itself needs h5py / libhdf5, which this environment lacks, so ``SaveHDF5`` raises
when a solve starts. ``SaveNpz`` writes the same tree: a group path becomes the
``/``-joined key of each dataset, an attribute ``name`` of the object at ``path`` the key
``path@name``. As with ``SaveHDF5`` in mode "cw", an existing file is extended: a solve
over new timesteps grows the time axis, one over old timesteps overwrites them, and
entries the savers do not own are kept. Arrays keep the reference's index order (the time
index last), and problem descriptions are numpy structured arrays.
"""
from __future__ import annotations

import os
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from . import problem as P


class SolutionValue:
    """A function of the State to be saved (io/io.jl:72-78)."""


class ArrayValue(SolutionValue):
    """``f(y, state)`` sets ``y``, an array of ``dims`` and ``dtype`` (io/io.jl:80-91)."""

    def __init__(self, f: Callable, dtype, dims: Tuple[int, ...]):
        self.f = f
        self.dtype = np.dtype(dtype)
        self.dims = tuple(int(d) for d in dims)


class MultiDomainValue(SolutionValue):
    """``f(y, state)`` sets a 3-D array ``y[i, j, level]`` at ``(x[i], y[j]) =
    coords[level]`` (io/io.jl:93-113)."""

    def __init__(self, f: Callable, coords):
        self.f = f
        self.coords = [(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
                       for x, y in coords]
        self.dims = (len(self.coords[0][0]), len(self.coords[0][1]), len(self.coords))
        self.dtype = np.dtype(np.float64)


class Timesteps:
    """The timesteps to save at (io/io.jl:115-120)."""


class TimestepRange(Timesteps):
    """Every ``step``-th timestep from ``start``; at most ``length`` of them
    (io/io.jl:115-143)."""

    def __init__(self, *, step: int, start: int = 1, length: Optional[int] = None):
        self.start, self.step, self.length = int(start), int(step), length

    def index_of(self, index: int):
        """(inside, 1-based position in the range) of a timestep index."""
        steps, rem = divmod(index - self.start, self.step)
        i = steps + 1
        return rem == 0 and i >= 1 and (self.length is None or i <= self.length), i

    def count(self, index: int) -> int:
        """Timesteps of the range at or before ``index``."""
        c = max(0, (self.step + index - self.start) // self.step)
        return c if self.length is None else min(self.length, c)


class SolutionValues:
    """Values to save at each time of ``times``: ``name => value`` pairs (io/io.jl:145-160)."""

    def __init__(self, times: TimestepRange, *vals):
        if len(vals) == 1 and isinstance(vals[0], (list, tuple)) and vals[0] and \
                isinstance(vals[0][0], (list, tuple)):
            vals = tuple(vals[0])
        self.times = times
        self.vals: List[Tuple[str, SolutionValue]] = [(str(n), v) for n, v in vals]


class SolutionSaver:
    """A backend for saving SolutionValues (io/io.jl:162-166). ``init(max_timestep)``
    returns an object with ``wants(index)``, ``update(state)`` and ``finalize()``."""

    def init(self, max_timestep: int):
        raise NotImplementedError


class _NothingSaver:
    def wants(self, index):
        return False

    def update(self, state):
        pass

    def finalize(self):
        pass


class SaveNothing(SolutionSaver):
    """Do not save anything (io/io.jl:168-176)."""

    def init(self, max_timestep):
        return _NothingSaver()


def fluid_group(prob, times: TimestepRange, funcs) -> SolutionValues:
    """Field accessors (x_velocity_, y_velocity_, vorticity_) on every level, named without
    the trailing underscore (io/hdf5.jl:21-26)."""
    out = []
    for f in funcs:
        out.append((f.__name__.rstrip("_"), MultiDomainValue(f, P.coords(prob.fluid.grid, f))))
    return SolutionValues(times, *out)


def body_group(prob, times: TimestepRange, funcs, bodies=None) -> SolutionValues:
    """Boundary accessors per body, named ``body/func`` (io/hdf5.jl:28-50)."""
    names = list(prob.bodies.byname) if bodies is None else [str(b) for b in bodies]
    out = []
    for body in names:
        r = prob.bodies.panel_range(body)
        for f in funcs:
            dims = P.arraysize(f, len(r))
            out.append((f"{body}/{f.__name__.rstrip('_')}",
                        ArrayValue((lambda y, st, f=f, r=r: f(y, st, slice(r.start, r.stop))),
                                   np.float64, dims)))
    return SolutionValues(times, *out)


# ---- the SaveHDF5 layout in a .npz file ----

def _fluid_record(fluid):
    g = fluid.grid
    (x0, x1), (y0, y1) = P.extents(g.base)
    lim = np.dtype([("min", "f8"), ("max", "f8")])
    grid = np.dtype([("h", "f8"), ("xlims", lim), ("ylims", lim), ("nlevel", "i8")])
    dt = np.dtype([("grid", grid), ("Re", "f8")])
    return np.array(((g.base.h, (x0, x1), (y0, y1), g.nlevel), fluid.Re), dtype=dt)


def _coord_matrix(coords):
    """2 x nlevel (start, step, length) records of each level's x and y coordinates
    (io/hdf5.jl:236-239)."""
    dt = np.dtype([("start", "f8"), ("step", "f8"), ("length", "i8")])
    out = np.zeros((2, len(coords)), dtype=dt)
    for lev, xy in enumerate(coords):
        for a in range(2):
            r = xy[a]
            step = float(r[1] - r[0]) if len(r) > 1 else 0.0
            out[a, lev] = (float(r[0]), step, len(r))
    return out


class _NpzFile:
    def __init__(self, path):
        self.path = path
        self.data: Dict[str, np.ndarray] = {}
        if os.path.exists(path):
            with np.load(path, allow_pickle=False) as f:
                self.data = {k: f[k] for k in f.files}

    def put_or_check(self, key, value):
        """_write_or_check_dataset / attribute (io/hdf5.jl:256-335)."""
        value = np.asarray(value)
        if key in self.data:
            old = self.data[key]
            if old.dtype != value.dtype or old.shape != value.shape or not np.array_equal(old, value):
                raise ValueError(f"{key!r} has an incorrect value")
        else:
            self.data[key] = value

    def time_dataset(self, key, dtype, dims, n):
        """_create_or_get_time_dataset (io/hdf5.jl:275-315): the time axis last, grown to n."""
        if key in self.data:
            a = self.data[key]
            if a.dtype != dtype:
                raise ValueError(f"dataset {key!r} has an incorrect datatype")
            if a.shape[:-1] != tuple(dims):
                raise ValueError(f"dataset {key!r} has an incorrect dataspace")
            if a.shape[-1] < n:
                b = np.zeros(tuple(dims) + (n,), dtype=dtype, order="F")
                b[..., :a.shape[-1]] = a
                self.data[key] = b
        else:
            self.data[key] = np.zeros(tuple(dims) + (n,), dtype=dtype, order="F")
        return key

    def write(self):
        tmp = self.path + ".tmp.npz"
        np.savez(tmp, **self.data)
        os.replace(tmp, self.path)


class _NpzSaver:
    def __init__(self, save, max_timestep):
        self.file = _NpzFile(save.file)
        self.updaters = []
        self.ranges = []
        for path, val in save.paths:
            path = str(path).strip("/")
            if isinstance(val, P.Problem):
                self._problem(path, val)
            elif isinstance(val, SolutionValues):
                self._values(path, val, max_timestep)
            else:
                raise TypeError(f"cannot save {type(val).__name__} at {path!r}")

    def _problem(self, path, prob):
        f = self.file
        f.put_or_check(f"{path}/fluid", _fluid_record(prob.fluid))
        for name, body, _, _ in prob.bodies.sections:
            b = f"{path}/bodies/{name}"
            f.put_or_check(f"{b}@type", np.array(type(body).__name__))
            f.put_or_check(f"{b}/positions", np.asfortranarray(body.panels.xb))
            f.put_or_check(f"{b}/lengths", body.panels.ds)
        f.put_or_check(f"{path}/scheme", np.array((prob.scheme.dt,), dtype=[("dt", "f8")]))
        f.put_or_check(f"{path}/scheme@type", np.array(type(prob.scheme).__name__))

    def _values(self, path, vals, max_timestep):
        f = self.file
        n = vals.times.count(max_timestep)
        tkey = f.time_dataset(f"{path}/_timestep_time", np.dtype(np.float64), (), n)
        ikey = f.time_dataset(f"{path}/_timestep_index", np.dtype(np.int64), (), n)
        savers = []
        for name, val in vals.vals:
            key = f.time_dataset(f"{path}/{name}", val.dtype, val.dims, n)
            if isinstance(val, MultiDomainValue):
                f.put_or_check(f"{key}@coords", _coord_matrix(val.coords))
            savers.append((key, val))
        self.ranges.append(vals.times)

        def update(state):
            inside, i = vals.times.index_of(state.index)
            if not inside:
                return
            f.data[tkey][i - 1] = state.t
            f.data[ikey][i - 1] = state.index
            for key, val in savers:
                y = np.zeros(val.dims, dtype=val.dtype, order="F")
                val.f(y, state)
                f.data[key][..., i - 1] = y

        self.updaters.append(update)

    def wants(self, index):
        return any(r.index_of(index)[0] for r in self.ranges)

    def update(self, state):
        for u in self.updaters:
            u(state)

    def finalize(self):
        self.file.write()


class SaveNpz(SolutionSaver):
    """SaveHDF5's layout (io/hdf5.jl:1-20) in the .npz file ``file``: ``paths`` are
    ``path => Problem`` or ``path => SolutionValues`` pairs."""

    def __init__(self, file: str, *paths):
        self.file = str(file)
        self.paths = list(paths)

    def init(self, max_timestep):
        return _NpzSaver(self, max_timestep)


class SaveHDF5(SolutionSaver):
    """SaveHDF5(file, paths...) (io/hdf5.jl:1-20). HDF5 is not available here (no h5py or
    libhdf5): starting a solve with it raises; SaveNpz writes the same layout."""

    def __init__(self, file, *paths, mode="cw"):
        self.file, self.paths, self.mode = file, list(paths), mode

    def init(self, max_timestep):
        try:
            import h5py  # noqa: F401
        except ImportError as e:
            raise RuntimeError("SaveHDF5 needs h5py and libhdf5, which are not installed; "
                               "SaveNpz writes the same layout to a .npz file") from e
        raise RuntimeError("SaveHDF5 is not implemented in this package; use SaveNpz")
