"""The multi-column apply (qkrylov_hamiltonian_apply_columns_device_f64):
nv real row-major columns, column c multiplied by H with its diagonal scaled
by s_c. Against the oracle (off-diagonal part plus s_c times the diagonal,
from the oracle's CSR, within 1e-13 relative); a column's bits do not depend
on the block width or on the other columns; refused operators."""
import numpy as np
import pytest

import oracle_lib as O
import qkrylov_amd as qa
from qkrylov_amd import models

pytestmark = pytest.mark.gpu


def field_terms(L, h):
    return [(h * (1 + 0.1 * i), [("Sz", i)]) for i in range(L)]


MODELS = {
    # XXZ ring at Delta = 1 (the serving loop's engine) and a 4x4 J1-J2
    # torus with a site-dependent field (the whole diagonal is scaled)
    "xxz_ring14": lambda: (14, 0, models.xxz_chain(14, 1.0)),
    "j1j2_4x4_field": lambda: (16, 2, models.j1j2_square(4, 4) + field_terms(16, 0.3)),
}


@pytest.mark.parametrize("name", sorted(MODELS))
@pytest.mark.parametrize("split", [-1, 6])
def test_columns_vs_oracle_and_width_independent(name, split):
    L, sz2, terms = MODELS[name]()
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(L, sz2=sz2), qa.SpinHalfSite(), qa.OpSum(terms), split_bits=split)
    om = O.Model(O.SPIN, L, O.sector(use_sz=1, sz2=sz2), O.SPIN, terms)
    D = H.dimension
    assert D == om.dim
    diag = om.csr()[3].real
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (D, 8))
    scales = np.array([1.0, 0.5, 1.5, 0.0, -0.75, 1.25, 0.9, 2.0])
    ref = np.empty((D, 8))
    for c in range(8):
        y_full = om.apply(X[:, c]).real
        ref[:, c] = y_full + (scales[c] - 1.0) * diag * X[:, c]
    out = {}
    for nv in (1, 2, 4, 8):
        for c0 in range(0, 8, nv):
            Xd = qa.DeviceArray.from_host(X[:, c0:c0 + nv])
            Yd = qa.DeviceArray(D * nv)
            print(f"apply_columns {name} split={split} D={D} nv={nv} c0={c0}", flush=True)
            H.apply_columns_device(Xd.ptr, Yd.ptr, scales[c0:c0 + nv])
            Y = Yd.to_host().reshape(D, nv)
            for j in range(nv):
                c = c0 + j
                err = np.linalg.norm(Y[:, j] - ref[:, c]) / np.linalg.norm(ref[:, c])
                assert err < 1e-13, (name, nv, c, err)
                if c in out:
                    assert np.array_equal(out[c], Y[:, j]), (name, nv, c)   # same bits at every width
                else:
                    out[c] = Y[:, j].copy()
            Xd.free()
            Yd.free()


def test_columns_refuse_what_they_do_not_cover():
    X, Y = qa.DeviceArray(3 * 16), qa.DeviceArray(3 * 16)
    H = qa.MatrixFreeHamiltonian(qa.SpinHalfBasis(6, sz2=0), qa.SpinHalfSite(), qa.OpSum(models.xxz_chain(6, 1.0)))
    with pytest.raises(RuntimeError):
        H.apply_columns_device(X.ptr, Y.ptr, np.ones(3))                  # width 3
    Hf = qa.MatrixFreeHamiltonian(qa.HubbardBasis(4, conserve_nup=True, nup=2, conserve_ndn=True, ndn=2),
                                  qa.HubbardSite(), qa.OpSum(models.hubbard_chain(4)))
    assert Hf.dimension == 36
    Xb, Yb = qa.DeviceArray(36 * 2), qa.DeviceArray(36 * 2)
    with pytest.raises(RuntimeError, match="fermion signs"):
        Hf.apply_columns_device(Xb.ptr, Yb.ptr, np.ones(2))
