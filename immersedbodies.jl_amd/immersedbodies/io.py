"""State checkpoint in the byte format of ImmersedBodies.jl's save_state / load_state!
(io.jl:1-71): the header string, Int64 index, Float64 t0, then the column-major Float64
arrays q, q0, Gamma, psi, nonlin[1], nonlin[2], panels.xb, panels.ds, panels.ub, F~b and
freestream_vel, native byte order. load_state! recomputes t (set_time_index!) and fb
(update_vars!). Files written here load in the reference and vice versa.
"""
from __future__ import annotations

import io as _io
from typing import BinaryIO, Union

import numpy as np

from .problem import Problem, State, gridstep, timestep

STATE_FILE_HEADER = "@ImmersedBodies.jl State v1@".encode("utf-8")


def _write(f: BinaryIO, a):
    f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def save_state(dst: Union[str, BinaryIO], state: State) -> None:
    """save_state(filename_or_io, state) (io.jl:10-25)."""
    if isinstance(dst, str):
        with open(dst, "wb") as f:
            return save_state(f, state)
    f = dst
    f.write(STATE_FILE_HEADER)
    f.write(np.array([state.index], dtype="<i8").tobytes())
    f.write(np.array([state.t0], dtype="<f8").tobytes())
    for a in (state.q, state.q0, state.Gamma, state.psi, state.nonlin[0], state.nonlin[1]):
        _write(f, a)  # (nlevel, n) row blocks == column-major n x nlevel
    _write(f, state.xb.T)  # column-major npanel x 2
    _write(f, state.ds)
    _write(f, state.ub.T)
    _write(f, state.Ftilde)
    _write(f, state.freestream_vel)


def _read(f: BinaryIO, n: int) -> np.ndarray:
    b = f.read(8 * n)
    if len(b) != 8 * n:
        raise EOFError("state file is truncated")
    return np.frombuffer(b, dtype="<f8").astype(np.float64)


def load_state_(src: Union[str, BinaryIO], state: State) -> State:
    """load_state!(filename_or_io, state) (io.jl:42-71)."""
    if isinstance(src, str):
        with open(src, "rb") as f:
            return load_state_(f, state)
    f = src
    if f.read(len(STATE_FILE_HEADER)) != STATE_FILE_HEADER:
        raise ValueError("State file is missing header")
    index = int(np.frombuffer(f.read(8), dtype="<i8")[0])
    state.t0 = float(np.frombuffer(f.read(8), dtype="<f8")[0])
    state.set_time_index(index)
    for name in ("q", "q0", "Gamma", "psi"):
        a = getattr(state, name)
        setattr(state, name, _read(f, a.size).reshape(a.shape))
    state.nonlin = [_read(f, a.size).reshape(a.shape) for a in state.nonlin]
    n = state.xb.shape[0]
    state.xb = _read(f, 2 * n).reshape(2, n).T.copy()
    state.ds = _read(f, n)
    state.ub = _read(f, 2 * n).reshape(2, n).T.copy()
    state.Ftilde = _read(f, 2 * n)
    state.freestream_vel = _read(f, 2)
    # update_vars!: fb = F~b h / dt (problem.jl:676-685)
    h, dt = gridstep(state.prob), timestep(state.prob)
    state.fb = (state.Ftilde.reshape(2, n).T * h / dt) if n else np.zeros((0, 2))
    return state


def load_state(src: Union[str, BinaryIO], prob: Problem) -> State:
    """load_state(filename_or_io, prob) (io.jl:35)."""
    return load_state_(src, State(prob))


def state_bytes(state: State) -> bytes:
    buf = _io.BytesIO()
    save_state(buf, state)
    return buf.getvalue()
