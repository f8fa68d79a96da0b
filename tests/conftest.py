import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def rand(rng, shape, cplx=False):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a


def hermitian_env(rng, D, w, cplx=False, scale=1.0):
    """Random environment E[a, w, a'] whose MPO-leg slices are Hermitian in (a, a')
    when paired with a Hermitian MPO (E[:, k, :]^dag = E[:, k, :])."""
    E = rand(rng, (D, w, D), cplx) * scale
    return 0.5 * (E + E.conj().transpose(2, 1, 0))


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from finitemptools_amd.engine import engine
    return engine(0)


# MPO channel pairs (i, j) with W_{.i}^dag = W_{.j} in the Schur layout of
# finitemptools_amd.models: the environments of a Hermitian MPO satisfy
# E[:, j, :] = E[:, i, :]^dag for these and E[:, k, :] Hermitian otherwise.
ADJOINT_CHANNELS = {"xxz": [(1, 2)], "tfim": [], "hubbard": [(1, 3), (2, 4)], "j1j2": [(1, 2)]}


def model_env(rng, D, w, model, cplx=False, scale=1.0):
    E = hermitian_env(rng, D, w, cplx, scale)
    for i, j in ADJOINT_CHANNELS[model]:
        E[:, i, :] = rand(rng, (D, D), cplx) * scale
        E[:, j, :] = E[:, i, :].conj().T
    return E
