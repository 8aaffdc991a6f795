"""The autodiff oracle of the adjoint (oracle/adjoint_torch.py) against the
C oracle: its forward traces must equal the reference restatement's."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_em2_torch_restatement_equals_c_oracle():
    from oracle.adjoint_torch import em2_cfg_inputs, em2_traces
    from oracle.oracle_ctypes import run_oracle
    from seidart_hip import models
    cfg = models.em2(nx=40, nz=30, steps=120, src=(20, 10), nrec=6)
    coef, src, A, dt, dx, dz, sij, rij = em2_cfg_inputs(cfg)
    tr = em2_traces(torch.tensor(coef), torch.tensor(src[0]), torch.tensor(src[1]), A, dt, dx, dz, sij, rij)
    rec, _, _, us = run_oracle(cfg)
    assert us == 0
    o = rec[:, :, [0, 2]]
    assert np.abs(o).max() > 1e-3
    # same operations in the same order: equal to the last bit
    assert np.array_equal(tr.numpy(), o)


def test_em2_coefficients_torch_equals_oracle():
    from oracle.adjoint_torch import em2_coefficients_np
    from seidart_hip import fwi, models
    A = models.em2(nx=30, nz=20, steps=10, src=(15, 10), nrec=4)["arrays"]
    names = ("e11", "e13", "e33", "s11", "s13", "s33")
    ref = em2_coefficients_np(*(A[n] for n in names))
    got = fwi.em2_coefficients([torch.tensor(A[n]) for n in names[:3]], [torch.tensor(A[n]) for n in names[3:]])
    assert np.array_equal(got.numpy(), ref)


@pytest.mark.parametrize("density", ["arithmetic", "harmonic", "geometric"])
@pytest.mark.parametrize("stencil", ["reference", "corrected"])
def test_el2_torch_restatement_equals_c_oracle(density, stencil):
    from oracle.adjoint_torch import el2_cfg_inputs, el2_traces
    from oracle.oracle_ctypes import run_oracle
    from seidart_hip import models
    cfg = models.elastic2(nx=40, nz=36, steps=100, src=(20, 12), nrec=6, density_method=density)
    cfg["stencil"] = stencil
    cfg["sources"]["xx"] = cfg["sources"]["z"] * 0.3
    coef, src, A, dt, dx, dz, sij, rij = el2_cfg_inputs(cfg)
    tr = el2_traces(torch.tensor(coef), torch.tensor(src), A, dt, dx, dz, sij, rij,
                    density_code={"harmonic": 1, "geometric": 2, "arithmetic": 3}[density],
                    corrected=stencil == "corrected")
    rec, _, _, us = run_oracle(cfg)
    assert us == 0
    o = rec[:, :, [0, 2]]
    assert np.abs(o).max() > 0
    assert np.array_equal(tr.numpy(), o)
