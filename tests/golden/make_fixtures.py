"""Writes the golden fixtures under tests/golden/ from the reference's own
parity test inputs and traces (parity/tests/*/ in the reference tree).

Run in a container where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference

Each fixture holds the compiled model (positions in internal units, residue
types, chains, native angles/dihedrals, native contacts, PBC cell), the seed
and SAW settings, and the reference's per-step total potential energy trace
(V_fort and V_cxx columns of comp.csv) for the first steps. Only data is
written; nothing of the reference's code is copied.

Input semantics restated here (cited):
  * PDB ATOM records, hydrogens skipped, chains in chain-id order,
    residues in order of appearance: cg/src/cg/input/pdb_file.cpp:14-84,336-363.
  * native bond angle / dihedral of consecutive CA atoms:
    cg/include/cg/utils/angles.h:5-19, pdb_file.cpp:365-397.
  * contacts "from all atoms": pdb_file.cpp:486-538 (heavy atoms of residues
    with |seq_num difference| >= 3 whose distance is within
    (r1 + r2) (26/7)^(1/6); radii by longest name prefix,
    cg/src/cg/amino/aa_data.cpp:16-31,102-110), one contact per residue
    pair with the CA-CA length, pdb_file.cpp:421-447.
  * sequence files: cg/src/cg/input/seq_file.cpp:8-138.
"""
from __future__ import annotations

import csv
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ANGSTROM = 0.2

AA_NAMES = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLU", "GLN", "GLY", "HIS", "ILE",
            "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL",
            "PHO"]
AA_LETTERS = "ARNDCEQGHILKMFPSTWYVO"

# default atom data and per-residue overrides (cg/default.yml:800-1049), A
DEF_ATOMS = {"C": 1.61, "CA": 1.88, "CB": 1.88, "CD": 1.88, "CD1": 1.61,
             "CD2": 1.76, "CE": 1.88, "CE1": 1.76, "CE2": 1.76, "CE3": 1.76,
             "CG": 1.88, "CG1": 1.88, "CG2": 1.88, "CH2": 1.76, "CZ": 1.61,
             "CZ2": 1.76, "CZ3": 1.73, "N": 1.64, "ND1": 1.64, "ND2": 1.64,
             "NE": 1.64, "NE1": 1.61, "NE2": 1.64, "NH1": 1.64, "NH2": 1.64,
             "NZ": 1.64, "O": 1.42, "OD1": 1.42, "OD2": 1.42, "OE1": 1.42,
             "OE2": 1.42, "OG": 1.46, "OG1": 1.46, "OH": 1.46, "SG": 1.77,
             "SD": 1.77}
OVERRIDES = {"ASP": {"CG": 1.61, "OD1": 1.46}, "GLU": {"CD": 1.88, "OE1": 1.46},
             "ILE": {"CD": 1.88}, "LEU": {"CD1": 1.88, "CD2": 1.88},
             "PHE": {"CZ": 1.76},
             "TRP": {"CD1": 1.76, "CD2": 1.61, "CE2": 1.64},
             "TYR": {"CD1": 1.76, "CG": 1.61}}


def _close(x, y):
    return abs(x - y) / abs(x) < 1e-5


def _correct(name, radius):
    """aa_data.cpp:16-31 radius correction."""
    if name[0] == "N":
        return 1.64 * ANGSTROM
    if name[0] == "S":
        return 1.77 * ANGSTROM
    if name[0] == "O" and not _close(radius, 1.42 * ANGSTROM) and not _close(radius, 1.46 * ANGSTROM):
        return 1.46 * ANGSTROM
    if (name[0] == "C" and not _close(radius, 1.88 * ANGSTROM)
            and not _close(radius, 1.76 * ANGSTROM) and not _close(radius, 1.61 * ANGSTROM)):
        return 1.88 * ANGSTROM
    return radius


def atom_radius(res, name):
    for n in range(len(name), 0, -1):
        pre = name[:n]
        ov = OVERRIDES.get(res, {})
        if pre in ov:
            return _correct(pre, ov[pre] * ANGSTROM)
        if pre in DEF_ATOMS:
            return _correct(pre, DEF_ATOMS[pre] * ANGSTROM)
    raise KeyError(name)


def parse_pdb(path):
    chains = {}  # chain id -> list of residues (in order of appearance)
    cryst1 = None
    with open(path) as f:
        for line in f:
            rec = line[:6]
            if rec.startswith("END") and not rec.startswith("ENDMDL"):
                break
            if rec == "CRYST1":
                cryst1 = [float(line[6:15]) * ANGSTROM, float(line[15:24]) * ANGSTROM,
                          float(line[24:33]) * ANGSTROM]
            if rec != "ATOM  ":
                continue
            name = line[12:16].strip()
            if name[0] == "H":
                continue
            resn = line[17:20].strip()
            cid = line[21]
            seqn = int(line[22:26])
            pos = [float(line[30:38]) * ANGSTROM, float(line[38:46]) * ANGSTROM,
                   float(line[46:54]) * ANGSTROM]
            res_list = chains.setdefault(cid, [])
            if not res_list or res_list[-1]["seq"] != seqn:
                res_list.append({"seq": seqn, "name": resn, "atoms": []})
            res_list[-1]["atoms"].append((name, pos))
    return chains, cryst1


def _sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _norm(a):
    return math.sqrt(_dot(a, a))


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
            a[0] * b[1] - a[1] * b[0]]


def _acos(x):
    return math.acos(-1.0 if x < -1.0 else (x if x < 1.0 else 1.0))


def bond_angle(r1, r2, r3):
    u1, u2 = _sub(r2, r1), _sub(r2, r3)
    return _acos(_dot(u1, u2) / (_norm(u1) * _norm(u2)))


def dihedral_angle(r1, r2, r3, r4):
    u1, u2, u3 = _sub(r2, r1), _sub(r3, r2), _sub(r4, r3)
    v1, v2 = _cross(u1, u2), _cross(u2, u3)
    phi = _acos(_dot(v1, v2) / (_norm(v1) * _norm(v2)))
    if _dot(v1, u3) < 0:
        phi = -phi
    return phi


def model_from_pdb(path, all_atom_contacts=False, load_structure=True):
    chains, cryst1 = parse_pdb(path)
    pos, types, chain_idx, seq_idx, seqnum = [], [], [], [], []
    ang, dih, teth = [], [], []
    res_atoms = []
    for ci, cid in enumerate(sorted(chains)):
        first = len(pos)
        for k, res in enumerate(chains[cid]):
            ca = [p for (n, p) in res["atoms"] if n == "CA"][0]
            pos.append(ca)
            types.append(AA_NAMES.index(res["name"]))
            chain_idx.append(ci)
            seq_idx.append(k)
            seqnum.append(res["seq"])
            res_atoms.append(res)
        last = len(pos)
        for i in range(first, last - 1):
            teth.append([i, i + 1])
            if i + 2 < last:
                th = bond_angle(pos[i], pos[i + 1], pos[i + 2])
                ang.append([i, i + 1, i + 2, 1 if load_structure else 0,
                            th if load_structure else 0.0])
                if i + 3 < last:
                    ph = dihedral_angle(pos[i], pos[i + 1], pos[i + 2], pos[i + 3])
                    dih.append([i, i + 1, i + 2, i + 3, 1 if load_structure else 0,
                                ph if load_structure else 0.0])
    cont = []
    if all_atom_contacts and load_structure:
        alpha = (26.0 / 7.0) ** (1.0 / 6.0)
        atoms = []
        for ri, res in enumerate(res_atoms):
            for (n, p) in res["atoms"]:
                atoms.append((ri, atom_radius(res["name"], n), p))
        seen = set()
        for (r1, rad1, p1) in atoms:
            for (r2, rad2, p2) in atoms:
                if r1 == r2 or abs(seqnum[r1] - seqnum[r2]) < 3:
                    continue
                if _norm(_sub(p1, p2)) <= (rad1 + rad2) * alpha:
                    key = (min(r1, r2), max(r1, r2))
                    if key in seen:
                        continue
                    seen.add(key)
                    cont.append([key[0], key[1],
                                 _norm(_sub(pos[key[0]], pos[key[1]])), 0])
        cont.sort()
    return {"n": len(pos), "pos": pos, "orig_pos": pos, "atype": types,
            "chain": chain_idx, "seq": seq_idx, "teth": teth, "ang": ang,
            "dih": dih, "cont": cont, "cryst1": cryst1}


def model_from_seq(seqs):
    pos, types, chain_idx, seq_idx, teth, ang, dih = [], [], [], [], [], [], []
    for ci, s in enumerate(seqs):
        first = len(pos)
        for k, c in enumerate(s):
            pos.append([0.0, 0.0, 0.0])
            types.append(AA_LETTERS.index(c))
            chain_idx.append(ci)
            seq_idx.append(k)
        last = len(pos)
        for i in range(first, last - 1):
            teth.append([i, i + 1])
            if i + 2 < last:
                ang.append([i, i + 1, i + 2, 0, 0.0])
                if i + 3 < last:
                    dih.append([i, i + 1, i + 2, i + 3, 0, 0.0])
    return {"n": len(pos), "pos": pos, "orig_pos": pos, "atype": types,
            "chain": chain_idx, "seq": seq_idx, "teth": teth, "ang": ang,
            "dih": dih, "cont": [], "cryst1": None}


def read_pair_csv(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    cols = rows[0][1:]
    out = [[0.0] * 21 for _ in range(21)]
    for row in rows[1:]:
        if not row:
            continue
        a1 = AA_NAMES.index(row[0])
        for name, v in zip(cols, row[1:]):
            out[a1][AA_NAMES.index(name)] = float(v)
    return out


def read_trace(path, cols, max_steps):
    out = {c: [] for c in cols}
    steps = []
    with open(path) as f:
        rd = csv.DictReader(f)
        for row in rd:
            s = int(row["step_idx"])
            if s >= max_steps:
                break
            steps.append(s)
            for c in cols:
                out[c].append(float(row[c]))
    out["step"] = steps
    return out


def main(ref):
    pt = os.path.join(ref, "parity", "tests")
    fixtures = {}

    m = model_from_pdb(os.path.join(pt, "start2/cxx_7/start_7_2.pdb"))
    fixtures["start2"] = dict(model=m, seed=1, pbc=True, saw=None,
                              trace=read_trace(os.path.join(pt, "start2/comp.csv"),
                                               ["V_fort_7", "V_cxx_7"], 20000))

    m = model_from_pdb(os.path.join(pt, "TEST/cxx/start.pdb"), load_structure=False)
    fixtures["TEST"] = dict(model=m, seed=1, pbc=True, saw=None,
                            md05=read_pair_csv(os.path.join(pt, "TEST/cxx/md05.csv")),
                            trace=read_trace(os.path.join(pt, "TEST/comp.csv"),
                                             ["V_fort", "V_cxx"], 200))

    m = model_from_pdb(os.path.join(pt, "input_test/cxx/start_6_2.pdb"),
                       load_structure=False)
    fixtures["input_test"] = dict(
        model=m, seed=1, pbc=True, saw=None,
        md05=read_pair_csv(os.path.join(pt, "input_test/cxx/md05.csv")),
        trace=read_trace(os.path.join(pt, "input_test/comp.csv"),
                         ["V_fort", "V_cxx"], 200))

    m = model_from_pdb(os.path.join(pt, "sample/cxx/sample.pdb"), load_structure=False)
    fixtures["sample"] = dict(model=m, seed=1, pbc=True, saw=None,
                              md05=read_pair_csv(os.path.join(pt, "sample/cxx/md05.csv")),
                              trace=read_trace(os.path.join(pt, "sample/comp.csv"),
                                               ["V_fort", "V_cxx"], 20))

    with open(os.path.join(pt, "9aac/cxx/9aac.yml")) as f:
        seq = "".join(ln.strip() for ln in f.read().split("|", 1)[1].splitlines())
    m = model_from_seq([seq])
    fixtures["9aac"] = dict(model=m, seed=77977, pbc=False,
                            saw=dict(bond=3.8, bond_f32=True, intersection=4.56),
                            md05=read_pair_csv(os.path.join(pt, "9aac/cxx/md05.csv")),
                            trace=read_trace(os.path.join(pt, "9aac/comp.csv"),
                                             ["V_fort", "V_cxx"], 200))

    m = model_from_pdb(os.path.join(pt, "1ubq/cxx/1ubq.pdb"), all_atom_contacts=True)
    fixtures["1ubq"] = dict(model=m, seed=448, pbc=False,
                            saw=dict(bond="avg_tether", intersection=4.56),
                            trace=read_trace(os.path.join(pt, "1ubq/comp.csv"),
                                             ["V_fort", "V_cxx"], 200))

    for name, fx in fixtures.items():
        path = os.path.join(HERE, f"{name}.json")
        with open(path, "w") as f:
            json.dump(fx, f, separators=(",", ":"))
        print(name, fx["model"]["n"], "residues,", len(fx["model"]["cont"]),
              "contacts,", len(fx["trace"]["step"]), "steps ->", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
