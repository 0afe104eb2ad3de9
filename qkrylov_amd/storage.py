"""Result storage: solver results to and from disk under the dataset names of
the reference's HDF5 example (examples/hdf5_storage.py:5-13): "energy" and
"eigenvector" for a Lanczos result, "eigenvalues" and "eigenvector_{i}" for
a Davidson result. Continued-fraction and FTLM results, which that example
does not store, use "alphas", "betas", "norm_phi0" and "beta",
"partition_function", "internal_energy", "specific_heat".

Files ending in .h5 / .hdf5 are written with h5py when it is importable;
every other name (and .h5 without h5py) is an uncompressed .npz archive
holding the same datasets plus a "kind" entry. Loading never unpickles.
"""
import os

import numpy as np

from .api import DavidsonResult, DynamicsResult, FTLMResult, LanczosResult

_KIND = {LanczosResult: "lanczos", DavidsonResult: "davidson", DynamicsResult: "dynamics", FTLMResult: "ftlm"}


def _datasets(result):
    """(kind, {dataset name: array}) for one solver result."""
    kind = _KIND.get(type(result))
    if kind is None:
        raise TypeError(f"cannot store a {type(result).__name__}")
    d = {}
    if kind == "lanczos":
        d["energy"] = np.float64(result.energy)
        if result.eigenvector is not None:
            d["eigenvector"] = np.asarray(result.eigenvector)
        for name in ("iterations", "converged"):
            if getattr(result, name, None) is not None:
                d[name] = np.int64(getattr(result, name))
        for name in ("alphas", "betas"):
            if getattr(result, name, None) is not None:
                d[name] = np.asarray(getattr(result, name), dtype=np.float64)
    elif kind == "davidson":
        d["eigenvalues"] = np.asarray(result.eigenvalues)
        for i, v in enumerate(result.eigenvectors):
            d[f"eigenvector_{i}"] = np.asarray(v)
        d["iterations"] = np.int64(result.iterations)
        d["converged"] = np.int64(result.converged)
    elif kind == "dynamics":
        d["alphas"] = np.asarray(result.alphas, dtype=np.float64)
        d["betas"] = np.asarray(result.betas, dtype=np.float64)
        d["norm_phi0"] = np.float64(result.norm_phi0)
    else:
        for name in ("beta", "partition_function", "internal_energy", "specific_heat"):
            d[name] = np.asarray(getattr(result, name), dtype=np.float64)
    return kind, d


def _is_h5(filename):
    return str(filename).lower().endswith((".h5", ".hdf5"))


def _h5py():
    try:
        import h5py
    except ImportError:
        return None
    return h5py


def save_result(filename, result):
    """Writes one LanczosResult / DavidsonResult / DynamicsResult /
    FTLMResult. Returns the path written (".npz" is appended by numpy when
    an .npz archive gets a name without that suffix)."""
    kind, d = _datasets(result)
    h5 = _h5py() if _is_h5(filename) else None
    if h5 is not None:
        with h5.File(filename, "w") as f:
            for k, v in d.items():
                f.create_dataset(k, data=v)
            f.attrs["kind"] = kind
        return str(filename)
    path = str(filename)
    if not path.endswith(".npz"):
        path += ".npz"
    np.savez(path, kind=np.array(kind), **d)
    return path


def _read(filename):
    h5 = _h5py() if _is_h5(filename) else None
    if h5 is not None and os.path.exists(filename):
        with h5.File(filename, "r") as f:
            d = {k: np.asarray(f[k][()]) for k in f.keys()}
            kind = f.attrs.get("kind")
        return (str(kind) if kind is not None else None), d
    path = str(filename)
    if not os.path.exists(path) and os.path.exists(path + ".npz"):
        path += ".npz"
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    kind = str(d.pop("kind")) if "kind" in d else None
    return kind, d


def load_result(filename):
    """Reads a file written by save_result (or by the reference's example,
    whose HDF5 files carry no kind: it is inferred from the dataset names)."""
    kind, d = _read(filename)
    if kind is None:
        kind = ("lanczos" if "energy" in d else "davidson" if "eigenvalues" in d
                else "ftlm" if "partition_function" in d else "dynamics")
    if kind == "lanczos":
        r = LanczosResult(float(d["energy"]), int(d["iterations"]) if "iterations" in d else None,
                          bool(d["converged"]) if "converged" in d else None, d.get("eigenvector"),
                          d.get("alphas"), d.get("betas"))
        return r
    if kind == "davidson":
        n = len(d["eigenvalues"])
        return DavidsonResult(d["eigenvalues"], [d[f"eigenvector_{i}"] for i in range(n)],
                              int(d.get("iterations", 0)), bool(d.get("converged", False)))
    if kind == "dynamics":
        return DynamicsResult(d["alphas"], d["betas"], float(d["norm_phi0"]))
    if kind == "ftlm":
        vals = [d[k] for k in ("beta", "partition_function", "internal_energy", "specific_heat")]
        vals = [float(v) if np.ndim(v) == 0 else v for v in vals]
        return FTLMResult(*vals)
    raise ValueError(f"unknown result kind {kind!r} in {filename}")
