"""ctypes mirrors of include/mdff_types.h (plain data, no library load)."""
import ctypes as C

NT = 16
MAX_LEN = 80

_d, _i = C.c_double, C.c_int
_T2d = (_d * NT) * NT
_T3d = ((_d * NT) * NT) * NT
_T3i = ((_i * NT) * NT) * NT

# enum values of mdff_types.h / mdff_gpu.h
NVE_LF, NVE_VV, NVT_NHCN = 0, 1, 2
TRUNC_NO, TRUNC_LINEAR, TRUNC_QUADRATIC = 0, 1, 2
PIM_SCF, PIM_SCFKO, PIM_PCG = 0, 1, 2
FP64, MIXED = 0, 1
# matrix-core arithmetic of the mixed form's k-space sums (mdff_set_kspace_modes)
KS_F32, KS_F16X3 = 2, 3
STATUS = {0: "ok", 1: "bad argument", 2: "invalid configuration", 3: "I/O error",
          4: "HIP runtime error", 5: "neighbour list overflow", 6: "non-finite result",
          7: "SCF unconverged", 8: "unsupported", 9: "out of memory",
          10: "larger than the fixed scratch of a no-allocation entry point"}
ERR_SCF = 7
ES = dict(LQ=1 << 0, LD=1 << 1, LT=1 << 2, LDAMP=1 << 3, FORCES=1 << 4, STRESS=1 << 5,
          EF=1 << 6, EFG=1 << 7, DIR=1 << 8, REC=1 << 9, INPIM=1 << 10, UPDATE_SF=1 << 11)


class Params(C.Structure):
    _fields_ = [
        ("ntype", _i),
        ("Fposff", _i), ("lverletL", _i), ("skindiff", _d),
        ("lstatic", _i), ("lrdf", _i), ("lpstress", _i), ("lreduced", _i),
        ("cutshortrange", _d), ("cutlongrange", _d),
        ("integrator", _i), ("dt", _d), ("temp", _d), ("tauTberendsen", _d),
        ("timesca_thermo", _d), ("timesca_baro", _d),
        ("npas", _i), ("nprint", _i), ("fprint", _i), ("cprint", _i), ("tprint", _i),
        ("tstart", _i), ("nequil", _i), ("nequilT", _i),
        ("nhc_n", _i), ("nhc_mults", _i), ("nhc_yosh_order", _i),
        ("lnmlj", _i), ("lbhmft", _i), ("lbhmftd", _i), ("lcoulombic", _i), ("lsymmetric", _i),
        ("nmassit", _i), ("nqit", _i), ("ndipit", _i), ("nquadit", _i), ("npolit", _i),
        ("massit", _d * NT), ("qit", _d * NT), ("dipit", (_d * 3) * NT),
        ("quadit", ((_d * 3) * 3) * NT), ("polit", ((_d * 3) * 3) * NT),
        ("lpolar", _i * NT),
        ("lpoldamping", _T3i), ("pol_damp_b", _T3d), ("pol_damp_c", _T3d), ("pol_damp_k", _T3i),
        ("alphaES", _d), ("kES", _i * 3), ("epsw", _d), ("lautoES", _i),
        ("Abhmftd", _T2d), ("Bbhmftd", _T2d), ("Cbhmftd", _T2d), ("Dbhmftd", _T2d), ("BDbhmftd", _T2d),
        ("sigmalj", _T2d), ("epslj", _T2d), ("plj", _T2d), ("qlj", _T2d), ("trunc", _i),
        ("algo_pim", _i), ("algo_extrapolate_dipole", _i), ("extrapolate_order", _i),
        ("min_scf_pim_iter", _i), ("max_scf_pim_iter", _i),
        ("conv_tol_ind", _d), ("omegakO", _d),
    ]


class Thermo(C.Structure):
    _fields_ = [
        ("istep", _i),
        ("e_kin", _d), ("e_nvt", _d),
        ("u_lj", _d), ("u_bhmftd", _d), ("u_coul", _d), ("u_pol", _d),
        ("pvir_lj", _d), ("pvir_bhmftd", _d), ("pvir_coul", _d),
        ("tau_lj", (_d * 3) * 3), ("tau_bhmftd", (_d * 3) * 3), ("tau_coul", (_d * 3) * 3),
        ("temp_r", _d), ("u_tot", _d), ("e_tot", _d), ("h_tot", _d), ("pressure", _d),
        ("n_scf", _i), ("rmsd", _d), ("rmsd_fp64", _d), ("verlet_updates", _i),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                v = [[v[i][j] for j in range(3)] for i in range(3)]
            out[name] = v
        return out


class Config(C.Structure):
    _fields_ = [
        ("nion", _i), ("ntype", _i), ("name", C.c_char * (MAX_LEN + 1)),
        ("A", (_d * 3) * 3),
        ("atypit", (C.c_char * (MAX_LEN + 1)) * NT),
        ("nionit", _i * NT),
        ("typia", C.POINTER(_i)),
        ("rx", C.POINTER(_d)), ("ry", C.POINTER(_d)), ("rz", C.POINTER(_d)),
        ("vx", C.POINTER(_d)), ("vy", C.POINTER(_d)), ("vz", C.POINTER(_d)),
        ("fx", C.POINTER(_d)), ("fy", C.POINTER(_d)), ("fz", C.POINTER(_d)),
    ]
