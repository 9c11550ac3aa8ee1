"""Writes salt_local_mesh.npz and salt_local_inparam.json: the reference's
SEG salt local example (examples/03_salt_body_SEG_local/input1D: its
Exodus mesh local_mesh__SEG_salt__5Hz.e and the inparam keys the GPU path
takes), read with axisem3d_amd.exodus and preloop.load_inparam; and
prem_regional_mesh.npz: the reference's regional PREM mesh (examples/
02_3d_crust_S362ANI_regional/input_share/regional_mesh__prem_ani__50s.e,
spherical and semi-spherical elements, transversely isotropic with Q);
and prem_global_mesh.npz: its global PREM mesh of the anisotropy example
(examples/05_anisotropy_global/PREM_anisotropy_w_and_wo_full_Cij_50s/
sim1_ani_prem_mesh/input/AxiSEM_prem_ani_50.e, with the fluid outer
core and linear elements at the centre).
Each mesh keeps the connectivity, coordinates, side sets, global
variables, the crdsys record and the elemental variables element_type,
fluid and the model at the element nodes.

usage (where the reference tree is): python tests/golden/make_exodus_fixture.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

SRC = "/root/reference/examples/03_salt_body_SEG_local/input1D"
ELEM_VARS = ["element_type", "fluid"] + [f"{v}_{k}" for v in
                                         ("RHO", "VP", "VS")
                                         for k in range(4)]
PREM_REGIONAL = ("/root/reference/examples/02_3d_crust_S362ANI_regional/"
                 "input_share/regional_mesh__prem_ani__50s.e")
PREM_GLOBAL = ("/root/reference/examples/05_anisotropy_global/"
               "PREM_anisotropy_w_and_wo_full_Cij_50s/sim1_ani_prem_mesh/"
               "input/AxiSEM_prem_ani_50.e")
PREM_VARS = ["element_type", "fluid"] + [
    f"{v}_{k}" for v in ("RHO", "VPV", "VPH", "VSV", "VSH", "ETA", "QKAPPA",
                         "QMU") for k in range(4)]


def main():
    from axisem3d_amd import exodus, preloop
    inp = preloop.load_inparam(SRC)
    ex = exodus.ExodusMesh(os.path.join(SRC, inp["exodus_mesh"]))
    np.savez_compressed(os.path.join(HERE, "salt_local_mesh.npz"),
                        **ex.to_arrays(ELEM_VARS))
    with open(os.path.join(HERE, "salt_local_inparam.json"), "w") as f:
        json.dump(inp, f, indent=1, sort_keys=True)
        f.write("\n")
    prem = exodus.ExodusMesh(PREM_REGIONAL)
    np.savez_compressed(os.path.join(HERE, "prem_regional_mesh.npz"),
                        **prem.to_arrays(PREM_VARS))
    glob = exodus.ExodusMesh(PREM_GLOBAL)
    np.savez_compressed(os.path.join(HERE, "prem_global_mesh.npz"),
                        **glob.to_arrays(PREM_VARS))


if __name__ == "__main__":
    main()
