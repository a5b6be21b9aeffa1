"""Write a VsvbSystem as the raw stream file tests/fortran/vsvb_drive.F90 reads:
the 14 int32 sizes of vsvb_system, then its arrays in struct order."""
import numpy as np

from valence_amd.system import pack

ORDER = ["atom_type", "coords", "nuc_charge", "type_shell_start", "shell_l", "shell_prim_start", "prim_exp",
         "prim_coef", "orb_natom", "orb_atoms", "orb_start", "xpset", "coeff", "pair_sc", "coeff_sc"]
SIZES = ["natom", "ntype", "nshell_type", "nprim", "norbs", "mxctr", "totlen", "npair", "nunpd", "ndocc", "ndf",
         "nspinc", "ntol_d", "ntol_i"]


def write(s, path):
    pk = pack(s)
    with open(path, "wb") as fh:
        fh.write(np.array([pk.scalars[k] for k in SIZES], np.int32).tobytes())
        for name in ORDER:
            if name == "pair_sc" and pk.scalars["npair"] == 0:
                continue
            a = pk.i32.get(name)
            if a is None:
                a = pk.f64[name]
            if name == "pair_sc":  # the reference's storage order, pair_sc(npair, 2, nspinc)
                a = np.asarray(a, np.int32).transpose(0, 2, 1)
            fh.write(np.ascontiguousarray(a).ravel().tobytes())
