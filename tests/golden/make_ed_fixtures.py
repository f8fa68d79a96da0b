"""Writes exact-diagonalisation ground-state energies used as fixtures by the
GPU DMRG tests (computed here by the oracle's independent Kronecker-product
Hamiltonians and scipy's sparse eigensolver).

    python tests/golden/make_ed_fixtures.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ed  # noqa: E402

CASES = {
    "xxz_L20_J1_D1": (lambda: ed.xxz_hamiltonian(20, 1.0, 1.0, 0.0)),
    "xxz_L12_J1_D1": (lambda: ed.xxz_hamiltonian(12, 1.0, 1.0, 0.0)),
    "tfim_L12_J1_h0.7": (lambda: ed.tfim_hamiltonian(12, 1.0, 0.7)),
    "j1j2_L12_J2_0.25": (lambda: ed.j1j2_hamiltonian(12, 1.0, 0.25)),
    "j1j2_L12_J2_0.5": (lambda: ed.j1j2_hamiltonian(12, 1.0, 0.5)),
    "hubbard_L6_t1_U4": (lambda: ed.hubbard_hamiltonian(6, 1.0, 4.0)),
}

# lowest energy inside one charge sector (n_up, n_dn)
SECTOR_CASES = {
    "hubbard_L6_t1_U4_N3_3": (6, 3, 3),
    "hubbard_L8_t1_U4_N4_4": (8, 4, 4),
}


def main():
    out = {}
    for name, build in CASES.items():
        t0 = time.time()
        H = build()
        e = float(ed.ground_energy(H)[0])
        out[name] = e
        print(f"{name}: {e:.15f} ({time.time() - t0:.1f} s)", flush=True)
    for name, (L, nup, ndn) in SECTOR_CASES.items():
        t0 = time.time()
        out[name] = float(ed.hubbard_sector_energy(L, nup, ndn))
        print(f"{name}: {out[name]:.15f} ({time.time() - t0:.1f} s)", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ed_energies.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
