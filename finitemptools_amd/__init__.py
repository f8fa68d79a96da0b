"""MI355X-native effective-Hamiltonian engine for finite-MPS algorithms.

The HIP library ``libmpsheff.so`` (C ABI in include/mpsheff.h) evaluates the
contractions behind H_eff.x and the environment updates; this package is the
host-side mirror of the reference's interface for that path
(algorithms/derivatives.jl, mpo/transfer.jl, envs/finiteenv.jl,
auxiliary/simple_lanczos.jl, algorithms/dmrg.jl, algorithms/tdvp.jl).
"""
from ._build import build, lib_path  # noqa: F401

__all__ = ["build", "lib_path"]
