"""Force-field parameters of the PAS CG model in internal units.

The values are the defaults of the reference's parameter file (cg/default.yml)
and the unit system of cg/include/cg/utils/units.h: length 1 = 5 A, time 1 =
tau, energy 1 = eps, k_B = 1, charge 1 = e. `default_params()` returns a plain
dict; the C-ABI wrappers pack it into their own structs.

Only data lives here: no physics is computed in this module.
"""
from __future__ import annotations

import copy
import math
import struct

# ---------------------------------------------------------------- units ----
# cg/include/cg/utils/units.h:8-59, evaluated in the same order.
F77UNIT = 1.0
# 2^(-1/6) as cg/include/cg/utils/math.h:6 writes it
C216_INV = 0.89089871814033927
ANGSTROM = F77UNIT / 5.0
NANOMETER = 10.0 * ANGSTROM
METER = NANOMETER * 1.0e9
NANOSECOND = 1.0
TAU = 1.0 * NANOSECOND
SECOND = NANOSECOND * 1.0e9
ATOM = 1.0
MOL = 6.02214076e23 / ATOM
EPS = 1.0
KCAL = EPS * MOL / 1.57824959
JOULE = KCAL / 4184.0
EPS_KB = 1.0
KB = EPS / EPS_KB
KG = JOULE * SECOND * SECOND / (METER * METER)
AMU = KG * 0.99999999965e-3 / MOL
ECHARGE = 1.0
COULOMB = ECHARGE / 1.602176634e-19
AMPERE = COULOMB / SECOND
CSPEED = 299792458.0 * METER / SECOND
HENRY = KG * METER * METER / (SECOND * SECOND * AMPERE * AMPERE)
MU_0 = 1.25663706212e-6 * HENRY / METER
EPS_0 = 1.0 / (MU_0 * CSPEED * CSPEED)
RAD = 1.0
DEG = (2.0 * math.pi / 360.0) * RAD


def f32(x: float) -> float:
    """Value of a float-parsed quantity such as `5e-3f` (utils/quantity.cpp:96-100)."""
    return struct.unpack("f", struct.pack("f", x))[0]


# -------------------------------------------------------- amino acids ----
# Residue codes in aa_code order (cg/include/cg/amino/amino_acid.h:17-39).
AA_NAMES = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLU", "GLN", "GLY", "HIS", "ILE",
            "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL",
            "PHO"]
AA_LETTERS = "ARNDCEQGHILKMFPSTWYVO"
AA_INDEX = {n: i for i, n in enumerate(AA_NAMES)}
LETTER_INDEX = {c: i for i, c in enumerate(AA_LETTERS)}
GLY, PRO, CYS = AA_INDEX["GLY"], AA_INDEX["PRO"], AA_INDEX["CYS"]

# cg/default.yml:839-1049 "amino acids": mass (amu), polarization,
# charge (e), contact limits (back, side all, side polar, side hydrophobic).
_AA_TABLE = {
    "ALA": (71.037113805, "hydrophobic", 0, (2, 4, 1, 1)),
    "ARG": (156.101111050, "polar", 1, (2, 4, 2, 0)),
    "ASN": (114.042927470, "polar", 0, (2, 3, 2, 0)),
    "ASP": (115.026943065, "polar", -1, (2, 2, 2, 0)),
    "CYS": (103.009184505, "hydrophobic", 0, (2, 5, 2, 2)),
    "GLU": (129.042593135, "polar", -1, (2, 2, 2, 0)),
    "GLN": (128.058577540, "polar", 0, (2, 3, 2, 0)),
    "GLY": (57.021463735, None, 0, (2, 0, 0, 0)),
    "HIS": (137.058911875, "polar", 0, (2, 4, 2, 0)),
    "ILE": (113.084064015, "hydrophobic", 0, (2, 7, 2, 4)),
    "LEU": (113.084064015, "hydrophobic", 0, (2, 7, 2, 4)),
    "LYS": (128.094963050, "polar", 1, (2, 3, 2, 0)),
    "MET": (131.040484645, "hydrophobic", 0, (2, 7, 1, 1)),
    "PHE": (147.068413945, "hydrophobic", 0, (2, 7, 2, 4)),
    "PRO": (97.052763875, None, 0, (1, 0, 0, 0)),
    "SER": (87.032028435, "polar", 0, (2, 2, 2, 0)),
    "THR": (101.047678505, "polar", 0, (2, 3, 2, 0)),
    "TRP": (186.079312980, "hydrophobic", 0, (2, 8, 3, 4)),
    "TYR": (163.063328575, "hydrophobic", 0, (2, 7, 2, 2)),
    "VAL": (99.068413945, "hydrophobic", 0, (2, 6, 1, 4)),
    "PHO": (129.116, "polar", -2, (2, 2, 2, 0)),
}
_PTYPE = {None: 0, "polar": 1, "hydrophobic": 2}

# ss_pairs_csv (cg/default.yml:427-449): LJ r_min (A) for side-chain pairs.
SS_PAIRS_CSV = """type,ALA,ARG,ASN,ASP,CYS,GLN,GLU,GLY,HIS,ILE,LEU,LYS,MET,PHE,PHO,PRO,SER,THR,TRP,TYR,VAL
ALA,6.41596,7.99254,7.02481,6.73004,6.96894,7.39072,7.03513,0.0,7.49807,7.4515,7.64571,7.25636,7.90635,8.167,6.73004,6.84509,6.52979,6.94089,8.6489,8.01558,7.06164
ARG,7.99254,10.0222,8.87054,9.12439,8.24142,9.25989,9.51825,0.0,9.2293,8.75517,8.98363,9.49215,9.27252,9.2578,9.12439,8.6967,8.27447,8.49644,9.82199,9.50922,8.31488
ASN,7.02481,8.87054,7.74478,7.49738,7.49217,8.18851,7.99731,0.0,8.36404,7.95509,8.13966,8.11461,8.48918,8.64859,7.49738,7.55152,7.18018,7.45743,9.10847,8.96106,7.53705
ASP,6.73004,9.12439,7.49738,7.3958,7.17592,8.15441,7.81457,0.0,8.49948,7.65258,7.85839,8.58582,8.14884,8.5134,7.3958,7.23515,6.99305,7.19339,9.10307,9.34604,7.22267
CYS,6.96894,8.24142,7.49217,7.17592,7.55733,7.72232,7.44878,0.0,8.1679,7.94719,8.07406,7.83117,8.29132,8.49679,7.17592,7.52633,6.97079,7.4017,8.94558,8.25576,7.55818
GLN,7.39072,9.25989,8.18851,8.15441,7.72232,8.6339,8.40626,0.0,8.64366,8.24492,8.43726,8.68554,8.83875,8.95351,8.15441,7.90612,7.63811,7.76803,9.58092,9.27035,7.8121
GLU,7.03513,9.51825,7.99731,7.81457,7.44878,8.40626,8.37524,0.0,8.8443,7.97275,8.19823,8.90461,8.61045,8.74517,7.81457,7.53281,7.40681,7.50872,9.47919,9.47761,7.49567
GLY,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0,0.0
HIS,7.49807,9.2293,8.36404,8.49948,8.1679,8.64366,8.8443,0.0,8.82674,8.36691,8.56678,8.57814,8.9267,8.97502,8.49948,7.90923,7.87748,7.97856,9.65746,9.38343,7.91666
ILE,7.4515,8.75517,7.95509,7.65258,7.94719,8.24492,7.97275,0.0,8.36691,8.5263,8.77193,8.15891,8.9472,9.11118,7.65258,7.87539,7.52021,7.92881,9.57259,8.77521,8.06246
LEU,7.64571,8.98363,8.13966,7.85839,8.07406,8.43726,8.19823,0.0,8.56678,8.77193,8.93237,8.39261,9.14397,9.33507,7.85839,8.01034,7.68498,8.12374,9.78508,9.01661,8.29202
LYS,7.25636,9.49215,8.11461,8.58582,7.83117,8.68554,8.90461,0.0,8.57814,8.15891,8.39261,8.93192,8.79787,8.79341,8.58582,7.80572,7.73488,7.79424,9.20968,9.03734,7.68742
MET,7.90635,9.27252,8.48918,8.14884,8.29132,8.83875,8.61045,0.0,8.9267,8.9472,9.14397,8.79787,9.29013,9.55493,8.14884,8.28535,7.94113,8.33046,10.0219,9.28102,8.48089
PHE,8.167,9.2578,8.64859,8.5134,8.49679,8.95351,8.74517,0.0,8.97502,9.11118,9.33507,8.79341,9.55493,9.73068,8.5134,8.35495,8.24345,8.58396,10.168,9.56181,8.68731
PHO,6.73004,9.12439,7.49738,7.3958,7.17592,8.15441,7.81457,0.0,8.49948,7.65258,7.85839,8.58582,8.14884,8.5134,7.3958,7.23515,6.99305,7.19339,9.10307,9.34604,7.22267
PRO,6.84509,8.6967,7.55152,7.23515,7.52633,7.90612,7.53281,0.0,7.90923,7.87539,8.01034,7.80572,8.28535,8.35495,7.23515,7.41516,7.04405,7.31767,8.73009,8.60035,7.47655
SER,6.52979,8.27447,7.18018,6.99305,6.97079,7.63811,7.40681,0.0,7.87748,7.52021,7.68498,7.73488,7.94113,8.24345,6.99305,7.04405,6.64507,6.96918,8.753,8.35846,7.16628
THR,6.94089,8.49644,7.45743,7.19339,7.4017,7.76803,7.50872,0.0,7.97856,7.92881,8.12374,7.79424,8.33046,8.58396,7.19339,7.31767,6.96918,7.30031,9.13956,8.5783,7.54035
TRP,8.6489,9.82199,9.10847,9.10307,8.94558,9.58092,9.47919,0.0,9.65746,9.57259,9.78508,9.20968,10.0219,10.168,9.10307,8.73009,8.753,9.13956,10.8456,10.0829,9.22171
TYR,8.01558,9.50922,8.96106,9.34604,8.25576,9.27035,9.47761,0.0,9.38343,8.77521,9.01661,9.03734,9.28102,9.56181,9.34604,8.60035,8.35846,8.5783,10.0829,9.34099,8.38777
VAL,7.06164,8.31488,7.53705,7.22267,7.55818,7.8121,7.49567,0.0,7.91666,8.06246,8.29202,7.68742,8.48089,8.68731,7.22267,7.47655,7.16628,7.54035,9.22171,8.38777,7.64602
"""

# heurestic bond angles coefficients (cg/default.yml:374-385): a0..a6 per
# (type1, type2) in eps/rad**k, types G, P, X.
HEUR_ANG_CSV = """G,G,20872.75597,-63260.52963,79318.301,-52680.17088,19554.21897,-3847.670279,313.6716916
G,P,8222.83155,-25178.89003,31841.70634,-21290.04519,7941.501624,-1567.87285,128.0761621
G,X,20872.75597,-63260.52963,79318.301,-52680.17088,19554.21897,-3847.670279,313.6716916
P,G,34646.70029,-109957.2324,144423.9036,-100525.9874,39127.31928,-8079.214542,691.8417699
P,P,10744.12043,-34148.94233,44818.66284,-31110.67875,12060.69185,-2479.723349,211.6367439
P,X,34646.70029,-109957.2324,144423.9036,-100525.9874,39127.31928,-8079.214542,691.8417699
X,G,15883.02041,-48923.1471,62360.6974,-42110.86572,15891.78309,-3178.490602,263.2916319
X,P,16912.27207,-53570.09757,70150.19389,-48602.41198,18791.04978,-3844.690523,325.3085829
X,X,15883.02041,-48923.1471,62360.6974,-42110.86572,15891.78309,-3178.490602,263.2916319
"""

# heurestic dihedral coefficients (cg/default.yml:392-403): const, sin, cos,
# sin2, cos2, sin_cos (eps) per (type2, type3).
HEUR_DIH_CSV = """G,G,0.133672207,-0.007576316,0.003786907,-0.124627339,0.425373566,-0.060606303
G,P,0.935285048,0.928786712,-0.18516837,0.015857805,0.2861951,0.072728001
G,X,0.210489196,-0.00606094,0.202709724,-0.160512736,0.461339767,0.13333598
P,G,0.233402207,-0.101516187,0.109235732,0.14906496,0.151803103,-0.742423775
P,P,1.810497634,1.171212546,0.091084321,-0.254152251,0.557619284,-1.569694253
P,X,0.661379307,0.115151794,0.428904959,0.200723546,0.100490651,-0.803028162
X,G,0.198889776,0.018181703,-0.070746181,0.122076238,0.178719533,-0.624241103
X,P,1.254229713,0.739393723,0.686217752,0.219188398,0.083115678,-0.790909154
X,X,0.275933756,0.00606063,0.257226522,0.15460117,0.146208844,-0.448484074
"""


def _heur_index(t1: str, t2: str) -> int:
    o = {"G": 0, "P": 1, "X": 2}
    return 3 * o[t1] + o[t2]


def parse_pair_table(csv_text: str) -> list[list[float]]:
    """Parses a per-pair table (`type,ALA,...` header) into [aa1][aa2]."""
    lines = [ln for ln in csv_text.strip().splitlines() if ln.strip()]
    cols = lines[0].split(",")[1:]
    out = [[0.0] * 21 for _ in range(21)]
    for ln in lines[1:]:
        cells = ln.split(",")
        a1 = AA_INDEX[cells[0]]
        for name, val in zip(cols, cells[1:]):
            out[a1][AA_INDEX[name]] = float(val)
    return out


SS_RMIN_A = parse_pair_table(SS_PAIRS_CSV)


def default_params() -> dict:
    """Defaults of cg/default.yml, internal units."""
    A = ANGSTROM
    p: dict = {}
    p["rep_cutoff"] = 5.0 * A                # general.repulsive cutoff
    p["counting_factor"] = 1.3
    p["nl_cutoff"] = 20.0 * A                # neighbor list
    p["nl_pad"] = 10.0 * A
    p["tether_on"] = 1
    p["tether_H1"] = 100.0 * EPS / (A ** 2)
    p["tether_H2"] = 0.0
    p["lrep_on"] = 1
    p["lrep_depth"] = 1.0
    p["chir_on"] = 0
    p["e_chi"] = 1.0
    p["nat_ang_on"] = 1
    p["CBA"] = 30.0
    p["heur_ang_on"] = 1
    poly = [[0.0] * 9 for _ in range(7)]
    for ln in HEUR_ANG_CSV.strip().splitlines():
        c = ln.split(",")
        k = _heur_index(c[0], c[1])
        for d in range(7):
            poly[d][k] = float(c[2 + d])
    p["heur_ang_poly"] = poly
    p["nat_dih_on"] = 1
    p["nat_dih_variant"] = 0
    p["CDA"] = 0.66
    p["CDB"] = 0.66
    p["CDH"] = 3.33
    p["heur_dih_on"] = 1
    coeffs = [[0.0] * 9 for _ in range(6)]
    for ln in HEUR_DIH_CSV.strip().splitlines():
        c = ln.split(",")
        k = _heur_index(c[0], c[1])
        for d in range(6):
            coeffs[d][k] = float(c[2 + d])
    p["heur_dih_coeffs"] = coeffs
    p["pauli_on"] = 1
    p["pauli_depth"] = 1.0
    p["nat_cont_on"] = 1
    p["nat_cont_depth"] = 1.0
    # native contacts "disulfide bond force": harmonic H1=100 eps/A^2, nat_r 6 A
    p["nc_ss_kind"] = 1
    p["nc_ss_H1"] = 100.0 / (A ** 2)
    p["nc_ss_H2"] = 0.0
    p["nc_ss_nat_r"] = 6.0 * A
    p["nc_ss_depth"] = 4.0
    p["nc_ss_r_min"] = 6.0 * A
    p["dh_on"] = 1
    p["dh_variant"] = 0
    p["dh_screen_dist"] = 15.0 * A
    p["dh_cutoff"] = 20.0 * A
    p["dh_perm"] = 80.0 * EPS_0
    p["qa_on"] = 1
    p["qa_include4"] = 1
    p["qa_phase_dur"] = 10.0
    p["qa_formation_tol"] = 0.0
    p["qa_min_cos_hr"] = 0.92
    p["qa_min_cos_hh"] = 0.75
    p["qa_max_cos_nr"] = 0.5
    set_qa_forces(p, "lj")
    p["pid_on"] = 0
    p["pid_include4"] = 1
    p["pid_lambda_variant"] = 1  # algebraic
    p["pid_bbp_alpha"], p["pid_bbp_psi0"] = 6.4, 1.05
    p["pid_bbm_alpha"], p["pid_bbm_psi0"] = 6.0, -1.44
    p["pid_ss_alpha"], p["pid_ss_psi0"] = 1.2, -0.23
    set_pid_forces(p, "lj")
    p["afm_on"] = 0
    p["afm_H1"] = 30.0 / (A ** 2)
    p["afm_H2"] = 0.0
    # afm.average stats over: the window of the tips' moving averages
    p["afm_stats_window"] = 100.0 * TAU
    # simulation box walls (sbox.walls): none by default; depth of the "solid"
    # type and the wall threshold
    p["n_walls"] = 0
    p["wall_origin"] = [[0.0, 0.0, 0.0] for _ in range(6)]
    p["wall_normal"] = [[0.0, 0.0, 0.0] for _ in range(6)]
    p["wall_depth"] = 4.0 * EPS
    p["wall_min_dist"] = 5.0 * A
    # per plane: kind (0 solid, 1 harmonic, 2 lj), box side (-1 none) and
    # connection limit; harmonic and LJ walls off, as state::traj_init leaves
    # them (state.cpp:111)
    p["wall_kind"] = [0] * 6
    p["wall_side"] = [-1] * 6
    p["wall_limit"] = [0] * 6
    p["hw_on"] = 0
    p["ljw_on"] = 0
    p["wall_hh1"] = 30.0 * EPS / (A ** 2)
    p["wall_lj_depth"] = 4.0 * EPS
    p["lang_on"] = 1
    p["gamma"] = f32(2.0)
    p["temperature"] = f32(0.35)
    p["dt"] = f32(5e-3)
    set_lj_wall_timing(p)
    p["mass"] = [1.0] * 21  # input.normalize mass: true
    p["charge"] = [float(_AA_TABLE[n][2]) for n in AA_NAMES]
    p["ptype"] = [_PTYPE[_AA_TABLE[n][1]] for n in AA_NAMES]
    p["base_sync"] = [list(_AA_TABLE[n][3]) for n in AA_NAMES]
    return p


def real_masses() -> list[float]:
    return [_AA_TABLE[n][0] * AMU for n in AA_NAMES]


def set_qa_forces(p: dict, variant: str = "lj") -> None:
    """Sink-LJ table for QA contact types (thread_setup.cpp:456-515):
    0 NONE, 1 bb, 2 bs, 3 sb, 4 + a1 + 21 a2 side-side."""
    A = ANGSTROM
    depth = [0.0] * 445
    lo = [0.0] * 445
    hi = [0.0] * 445

    def put(k, d, rmin):
        depth[k] = d
        if variant == "lj":
            lo[k] = hi[k] = rmin
        else:  # "sink lj" with inherited values: r_low = repulsive cutoff
            lo[k] = p["rep_cutoff"]
            hi[k] = rmin

    put(1, 1.0, 5.0 * A)
    put(2, 1.0, 6.8 * A)
    put(3, 1.0, 6.8 * A)
    for a1 in range(21):
        for a2 in range(21):
            put(4 + a1 + 21 * a2, 1.0, SS_RMIN_A[a1][a2] * A)
    p["qa_lj_depth"], p["qa_lj_r_low"], p["qa_lj_r_high"] = depth, lo, hi


def set_pid_forces(p: dict, variant: str = "lj", bbp_depth=None, bbm_depth=None,
                   ss_depth_table=None) -> None:
    """bb+/bb-/ss sink-LJ of the PID term (thread_setup.cpp:610-665)."""
    A = ANGSTROM
    bbp = [0.2, 5.6 * A, 5.6 * A]
    bbm = [0.2, 6.2 * A, 6.2 * A]
    ss_d = [1.0] * 441
    ss_lo = [0.0] * 441
    ss_hi = [0.0] * 441
    for a1 in range(21):
        for a2 in range(21):
            r = SS_RMIN_A[a1][a2] * A
            ss_lo[a1 * 21 + a2] = r
            ss_hi[a1 * 21 + a2] = r
    if variant == "sink lj":
        bbp = [bbp_depth if bbp_depth is not None else 0.2, 0.0, 5.6 * A]
        bbm = [bbm_depth if bbm_depth is not None else 0.2, 0.0, 6.2 * A]
        for k in range(441):
            ss_lo[k] = 0.0
        if ss_depth_table is not None:
            for a1 in range(21):
                for a2 in range(21):
                    ss_d[a1 * 21 + a2] = ss_depth_table[a1][a2]
    p["pid_bbp_lj"], p["pid_bbm_lj"] = bbp, bbm
    p["pid_ss_depth"], p["pid_ss_r_low"], p["pid_ss_r_high"] = ss_d, ss_lo, ss_hi


def clone(p: dict) -> dict:
    return copy.deepcopy(p)


def set_solid_walls(p: dict, lo, hi, axes="xyz") -> None:
    """Solid walls on the faces of the box [lo, hi] along the given axes: on
    each axis a wall at lo facing +axis and one at hi facing -axis, in that
    order (state::setup_walls with the walls moved to the box faces,
    cg/src/cg/simul/state.cpp:588-640)."""
    origin, normal = [], []
    for a in axes:
        k = "xyz".index(a)
        e = [0.0, 0.0, 0.0]
        e[k] = 1.0
        o_lo = [0.0, 0.0, 0.0]
        o_lo[k] = float(lo[k])
        o_hi = [0.0, 0.0, 0.0]
        o_hi[k] = float(hi[k])
        origin += [o_lo, o_hi]
        normal += [e, [-x for x in e]]
    p["n_walls"] = len(origin)
    p["wall_origin"] = origin + [[0.0, 0.0, 0.0]] * (6 - len(origin))
    p["wall_normal"] = normal + [[0.0, 0.0, 0.0]] * (6 - len(normal))


def set_lj_wall_timing(p: dict, breaking_dist_factor: float = 1.5, cycle_dur: float = 0.0) -> None:
    """The LJ wall constants thread::setup_lj_walls binds
    (cg/src/cg/simul/thread_setup.cpp:787-790): breaking distance
    factor * 2^(-1/6) * threshold, saturation step dt / cycle duration (a zero
    cycle duration gives an infinite step: connections saturate, or break, at
    once). Call again after changing dt or the threshold."""
    p["ljw_breaking_dist"] = breaking_dist_factor * math.pow(0.5, 1.0 / 6.0) * p["wall_min_dist"]
    p["ljw_saturation_diff"] = p["dt"] / cycle_dur if cycle_dur != 0.0 else math.inf


def auto_wall_limit(num_res: int, threshold: float, target_density: float) -> int:
    """The "auto" connection limit of harmonic and LJ walls
    (cg/src/cg/simul/state.cpp:613-620): C int conversion of a double
    quotient, then an integer halving, both toward zero."""
    import numpy as np
    x = float(np.cbrt(float(num_res))) - threshold * float(np.cbrt(float(target_density)))
    lim = int((num_res - float(round(math.pow(x, 3.0)))) / 3)
    return int(lim / 2)


AXIS_TYPES = ("void", "periodic", "solid", "harmonic", "lj")


def set_box_walls(p: dict, axes: dict, num_res: int, target_density: float = 1e-3 / ANGSTROM ** 3,
                  limits: dict | None = None) -> list:
    """Walls of state::setup_walls (cg/src/cg/simul/state.cpp:570-700) from the
    per-axis types of sbox.walls ({"x": "solid", "y": "periodic", ...};
    missing axes are "void"): for each walled axis, in x, y, z order, a plane
    facing +axis on the -axis side and one facing -axis on the +axis side,
    through the origin (pcg_box_setup moves them to the box faces). limits
    gives the connection limit of "harmonic" and "lj" ({"harmonic": 10}); the
    default is the auto limit of the target density. Returns the periodic
    flag of each axis (the pbc field of the box)."""
    limits = limits or {}
    origin, normal, kind, side, limit = [], [], [], [], []
    pbc = [0, 0, 0]
    for k, a in enumerate("xyz"):
        t = axes.get(a, "void")
        if t not in AXIS_TYPES:
            raise ValueError(f"unknown wall type {t!r}")
        if t == "periodic":
            pbc[k] = 1
        if t in ("void", "periodic"):
            continue
        e = [0.0, 0.0, 0.0]
        e[k] = 1.0
        lim = limits.get(t)
        if lim is None:
            lim = auto_wall_limit(num_res, p["wall_min_dist"], target_density) if t != "solid" else 0
        for s_, nrm in ((2 * k, e), (2 * k + 1, [-x for x in e])):
            origin.append([0.0, 0.0, 0.0])
            normal.append(list(nrm))
            kind.append({"solid": 0, "harmonic": 1, "lj": 2}[t])
            side.append(s_)
            limit.append(int(lim))
    pad = 6 - len(origin)
    p["n_walls"] = len(origin)
    p["wall_origin"] = origin + [[0.0, 0.0, 0.0]] * pad
    p["wall_normal"] = normal + [[0.0, 0.0, 0.0]] * pad
    p["wall_kind"] = kind + [0] * pad
    p["wall_side"] = side + [-1] * pad
    p["wall_limit"] = limit + [0] * pad
    return pbc


def box_params(box_min, box_max, pbc=(0, 0, 0), total_time: float = 3e6 * TAU, **kw) -> dict:
    """The simulation-box protocol parameters (pcg_box) with the defaults of
    the "simulation box" section of cg/default.yml in internal units;
    keyword arguments override fields by name."""
    A = ANGSTROM
    b = {
        "box_min": [float(x) for x in box_min], "box_max": [float(x) for x in box_max],
        "pbc": [int(x) for x in pbc], "total_time": total_time,
        "rest_period": 10e3 * TAU, "target_vel": 5e-3 * A / TAU,
        "accel_dist": 12.0 * A, "avg_forces_over": 100.0 * TAU,
        "squeezing": 0, "target_density": 1e-3 / A ** 3, "vel_above_2V": 2e-2 * A / TAU,
        "force_min": 0, "force_for_max_vel": 0.4 * EPS / A,
        "oscillations": 0, "osc_shear": 0, "num_of_cycles": 6, "angular_freq": 1e-4 / TAU,
        "amp_relative": 0, "amp_abs": 10.0 * A, "amp_rel": 0.1,
    }
    for k, v in kw.items():
        if k not in b:
            raise KeyError(k)
        b[k] = v
    return b


def initial_box(pos, kind: str = "sufficient", bond: float = 3.8 * ANGSTROM, pad_x_bond: float = 2.0,
                max_density: float | None = 1e-4 / ANGSTROM ** 3, cubic: bool = False,
                squeezing: bool = False, cryst1=None):
    """state::compile_model's box (cg/src/cg/simul/state.cpp:173-225):
    "from CRYST1 record" [0, cryst1]; "sufficient": the cube of the maximum
    density (if given) widened to every residue + pad (pad = pad_x_bond *
    bond, the mean tether length), as a cube when cubic or squeezing.
    Returns (box_min, box_max). As in the reference, the non-cubic case starts
    its residue loop at index 1."""
    import numpy as np
    pos = np.asarray(pos, np.float64)
    n = len(pos)
    if kind == "from CRYST1 record":
        return [0.0, 0.0, 0.0], [float(x) for x in cryst1]
    if kind != "sufficient":
        raise ValueError(kind)
    pad = pad_x_bond * bond
    side = 0.0
    lo, hi = [1.0, 1.0, 1.0], [-1.0, -1.0, -1.0]
    if max_density is not None:
        vol = n / max_density
        side = float(np.cbrt(vol)) / 2.0
        lo, hi = [-side] * 3, [side] * 3
    if cubic or squeezing:
        for i in range(n):
            for c in range(3):
                side = max(side, abs(float(pos[i, c])) + pad)
        return [-side] * 3, [side] * 3
    for i in range(1, n):
        for c in range(3):
            lo[c] = min(lo[c], float(pos[i, c]) - pad)
            hi[c] = max(hi[c], float(pos[i, c]) + pad)
    return lo, hi
