"""qkrylov_amd -- MI355X-native engine for qkrylov's matrix-free H.x apply and
its Krylov loops (Lanczos, Davidson, continued fraction, FTLM), behind the
reference's C ABI (include/qkrylov/c_api.h) and Python names."""
from ._lib import load as _load_library  # noqa: F401
from .api import (  # noqa: F401
    B, Bdag, C, Cdag, CDn, CdagDn, CdagUp, CUp, DavidsonResult, DeviceArray, DynamicsResult, FermionBasis,
    FermionSite, FTLMResult, HubbardBasis, HubbardSite, LanczosResult, MatrixFreeHamiltonian, N, Ndn, Nup, Nupdn,
    OpSum, SpinHalfBasis, SpinHalfSite, Sm, Sp, Sx, Sy, Sz, TJBasis, TJSite, continued_fraction_batched,
    continued_fraction_coeffs, davidson_lowest, evaluate_spectral_function, ftlm, ftlm_batched, lanczos_ground_state,
)
from . import models  # noqa: F401


def gpu_count():
    return int(_load_library().qkrylov_amd_device_count())
