"""qkrylov_amd -- MI355X-native engine for qkrylov's matrix-free H.x apply and
its Krylov loops (Lanczos, Davidson, continued fraction, FTLM), behind the
reference's C ABI (include/qkrylov/c_api.h) and Python names
(bindings/python/qkrylov/__init__.py)."""
from ._lib import load as _load_library  # noqa: F401
from .api import (  # noqa: F401
    B, Basis, Bdag, C, Cdag, CDn, CdagDn, CdagUp, CUp, DavidsonResult, DeviceArray, DynamicsResult, FermionBasis,
    FermionSite, FTLMResult, HubbardBasis, HubbardSite, LanczosResult, MatrixFreeHamiltonian, N, Ndn, Nup, Nupdn,
    Op, OpSum, Site, SpinHalfBasis, SpinHalfSite, Sm, Sp, Sx, Sy, Sz, TJBasis, TJSite, continued_fraction_batched,
    continued_fraction_coeffs, davidson_lowest, evaluate_spectral_function, ftlm, ftlm_batched, lanczos_ground_state,
    OpProduct, OpTermSum, StructureFactorResult, structure_factor, tridiag_eigen, tridiag_lowest,
    get_start_mode, reference_normal_vector, reference_start_vector, set_start_mode,
)
from . import models  # noqa: F401
from .storage import load_result, save_result  # noqa: F401

__version__ = "0.1.0"


def gpu_count():
    """Number of visible HIP devices."""
    return int(_load_library().qkrylov_amd_device_count())


def find_gpu():
    """'hip' when a HIP device is visible, else None (the reference returns
    its build's GPU backend name, bindings/python/qkrylov/__init__.py:30-34)."""
    return "hip" if gpu_count() > 0 else None


__all__ = [
    "Op", "OpSum",
    "Sz", "Sp", "Sm", "Sx", "Sy", "CdagUp", "CUp", "CdagDn", "CDn", "Nup", "Ndn", "Nupdn", "Bdag", "B", "N",
    "Site", "SpinHalfSite", "FermionSite", "HubbardSite", "TJSite",
    "Basis", "SpinHalfBasis", "FermionBasis", "HubbardBasis", "TJBasis",
    "MatrixFreeHamiltonian",
    "LanczosResult", "lanczos_ground_state", "DavidsonResult", "davidson_lowest", "DynamicsResult",
    "continued_fraction_coeffs", "evaluate_spectral_function", "FTLMResult", "ftlm",
    "ftlm_batched", "continued_fraction_batched", "DeviceArray", "structure_factor", "StructureFactorResult",
    "OpProduct", "OpTermSum", "tridiag_eigen", "tridiag_lowest",
    "set_start_mode", "get_start_mode", "reference_start_vector", "reference_normal_vector",
    "find_gpu", "gpu_count", "save_result", "load_result",
]
