"""Checkpoint and resume of device-resident states, and the reference's MAT files.

Two formats:

* ``save_state`` / ``load_state``: this project's snapshot of one MPS (every
  site in the reference layout ``[Dl, d, Dr]``, complex128, plus bonds, local
  dims and a JSON dictionary of caller metadata such as the step index, time
  or a trajectory's random-generator state) in one ``.npz`` file, written to a
  temporary name and renamed, so an interrupted write never leaves a
  half-written snapshot under the final name.  Loading uses numpy's default
  ``allow_pickle=False``.  A state resumed from a snapshot continues exactly
  as the uninterrupted run: an RK4 step reads nothing but its input state,
  the MPOs and the environments it builds itself.
* ``write_data_file`` / ``read_data_file``: the MAT v5 file the reference's
  density-matrix driver writes every ``measure_int*20`` steps and at the end
  (``examples/rydberg_density_matrix_gpu.jl:251-255, 273-316``) and its restart
  driver reads back (``rydberg_pollution_density_matrix_gpu_restart.jl:
  151-161``): the run's parameters, ``i_last``, the measurement arrays and
  ``rho`` as a cell array of site tensors; through ``scipy.io``.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import core as M

FORMAT = "mpship-mps-snapshot-1"


def _atomic(path, write):
    tmp = f"{path}.tmp{os.getpid()}"
    try:
        write(tmp)
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


def save_state(path: str, mps: "M.MPS", **meta) -> None:
    """Snapshot of a device MPS (downloaded once) and JSON metadata to ``path`` (.npz)."""
    save_sites(path, mps.to_host(), **meta)


def save_sites(path: str, sites, **meta) -> None:
    """Snapshot of host sites [Dl, d, Dr] and JSON metadata to ``path`` (.npz)."""
    sites = [np.asarray(a, dtype=np.complex128) for a in sites]
    arrays = {f"site_{j}": a for j, a in enumerate(sites)}
    arrays["format"] = np.array(FORMAT)
    arrays["dims"] = np.asarray([sites[0].shape[0]] + [a.shape[2] for a in sites], dtype=np.int64)
    arrays["d"] = np.asarray([a.shape[1] for a in sites], dtype=np.int64)
    arrays["meta"] = np.array(json.dumps(meta, sort_keys=True))

    def write(tmp):
        with open(tmp, "wb") as f:
            np.savez(f, **arrays)

    _atomic(path, write)


def load_sites(path: str):
    """(sites, meta) of a snapshot, on the host."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not an {FORMAT} file")
        dims = [int(x) for x in z["dims"]]
        sites = [np.asarray(z[f"site_{j}"], dtype=np.complex128) for j in range(len(dims) - 1)]
        meta = json.loads(str(z["meta"]))
    for j, a in enumerate(sites):
        if a.shape[0] != dims[j] or a.shape[2] != dims[j + 1]:
            raise ValueError(f"{path}: site {j} shape {a.shape} does not match bonds {dims[j]}, {dims[j + 1]}")
    return sites, meta


def load_state(path: str, ctx: "M.Context | None" = None, into: "M.MPS | None" = None):
    """(device MPS, meta) from a snapshot; ``into``: upload into this existing
    state (same shapes) instead of creating one."""
    sites, meta = load_sites(path)
    if into is not None:
        into.upload(sites)
        return into, meta
    return M.MPS.from_host(sites, ctx), meta


# ------------------------------------------------------------- MAT v5 files
def _cell(arrays):
    c = np.empty((len(arrays), 1), dtype=object)
    for j, a in enumerate(arrays):
        c[j, 0] = np.asarray(a)
    return c


def write_data_file(path: str, i: int, t_m, observables: dict, rho, times=None, params: dict | None = None):
    """write_data_file (rydberg_density_matrix_gpu.jl:273-316): parameters,
    ``i_last``, ``t_m``, the measurement arrays of ``observables`` (e.g. e_pop,
    e_pop_j, s_pop, s_pop_j, tr_rho, I_r, I2_r, I_l, I2_l) and ``rho`` (a device
    MPS or a list of host sites) as a cell array, to a MAT v5 file."""
    from scipy.io import savemat

    sites = rho.to_host() if isinstance(rho, M.MPS) else [np.asarray(a) for a in rho]
    out = {"TN": "Complex{Float64}", "TA": "libmpship (HIP, gfx950)", "i_last": int(i),
           "t_m": np.asarray(t_m, dtype=np.float64), "rho": _cell(sites)}
    for k, v in (params or {}).items():
        out[k] = v
    for k, v in observables.items():
        out[k] = np.asarray(v)
    if times is not None:
        out["times"] = np.asarray(times, dtype=np.float64)

    def write(tmp):
        with open(tmp, "wb") as f:
            savemat(f, out, do_compression=False, oned_as="column")

    _atomic(path, write)


def read_data_file(path: str) -> dict:
    """A write_data_file MAT file back: ``rho`` as a list of [Dl, d, Dr] host
    arrays, ``i_last`` as an int, the other fields as numpy arrays (squeezed)."""
    from scipy.io import loadmat

    raw = loadmat(path, squeeze_me=False)
    out = {}
    for k, v in raw.items():
        if k.startswith("__"):
            continue
        if k == "rho":
            out["rho"] = [np.asarray(v[j, 0], dtype=np.complex128) for j in range(v.shape[0])]
        elif k == "i_last":
            out["i_last"] = int(np.asarray(v).ravel()[0])
        else:
            out[k] = np.squeeze(v)
    # MATLAB drops trailing singleton dimensions; a [Dl, d, 1] site read by
    # other MAT readers may come back as [Dl, d]
    if "rho" in out:
        out["rho"] = [a if a.ndim == 3 else a.reshape(a.shape[0], -1, 1, order="F") for a in out["rho"]]
    return out
