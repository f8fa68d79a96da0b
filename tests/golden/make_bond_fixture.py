"""Write hubbard_L64_D2048_bonds.json: the charge sectors of three bonds of
the Hubbard chain (L = 64, t = 1, U = 4, half filling, U1 x U1) after a
two-site DMRG ramp D = 256, 512, 1024, 2048 (one sweep each, eps = 1e-10)
from the charge-density-wave product state.  For each bond (key: the left
site of the two-site tensor) it stores the two-site tensor's left and right
bond sectors and the environments' sectors and MPO channel charges; the
benchmark builds its middle-bond H_eff plan from them with seeded random
block contents.  Needs the GPU engine:  python tests/golden/make_bond_fixture.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main():
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import models
    from finitemptools_amd.dmrg import DMRG2
    from finitemptools_amd.truncation import MPSTruncation
    L = 64
    phys = bm.hubbard_charges()
    psi = bm.prodmps(phys, [phys.index((0, 0)), phys.index((1, 1))] * (L // 2))
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    env = bm.BlockExpectationCache(mpo, psi)
    for D in (256, 512, 1024, 2048):
        bm.sweep(env, DMRG2(trunc=MPSTruncation(D=D, eps=1e-10)))
        print(f"D={D} done", flush=True)
    out = {"bonds": {}}
    for site in (L // 2 - 1, L // 4, 8):
        x = bm.two_site(psi[site], psi[site + 1])
        Le, Re = env.hstorage[site], env.hstorage[site + 2]
        out["bonds"][site] = {
            "x_l": x.layout.lspace.sectors, "x_r": x.layout.rspace.sectors,
            "L_space": Le.layout.space.sectors, "L_qw": Le.layout.qw,
            "R_space": Re.layout.space.sectors, "R_qw": Re.layout.qw}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hubbard_L64_D2048_bonds.json")
    with open(path, "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
