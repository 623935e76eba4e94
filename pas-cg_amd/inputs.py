"""The reference's input formats, read into the library's model dicts.

* PDB files: ATOM records of the first model (hydrogens skipped), CRYST1,
  SSBOND; chains in chain-id order, residues in order of appearance
  (cg/src/cg/input/pdb_file.cpp:14-84). The model takes the C-alpha atoms,
  native tether lengths, bond angles and dihedrals of consecutive residues
  (pdb_file.cpp:336-397, angles of cg/include/cg/utils/angles.h:5-19).
  Native contacts are the SSBOND pairs (type SSBOND), then the contacts
  derived from atom overlaps (pdb_file.cpp:470-518): heavy atoms (or only
  C-alpha atoms with residue radii) of residues whose sequence numbers differ
  by 3 or more, within (r1 + r2) (26/7)^(1/6); one contact per residue pair,
  with the C-alpha distance and a type from the first overlapping atom pair
  (backbone N, CA, C, O, OXT or side chain; pdb_file.cpp:399-452).
* Sequence files (cg/src/cg/input/seq_file.cpp:8-138) with contact maps
  (cg/src/cg/input/map_file.cpp:26-53): YAML with `model: chains:`, each
  chain a `seq` string and optional `maps`, each map inline or
  `(at path): file` or `__source:` with a `shift`; the map's contacts,
  angles and dihedrals are CSV tables (i1,i2,length / i1,i2,i3,theta /
  i1,i2,i3,i4,phi) in internal units and radians. Residues start at the
  origin; morph_into_saw draws the starting conformation.
* The self-avoiding-walk start (model::morph_into_saw, model.cpp:108-234)
  through the library's pcg_morph_saw.
* Amino-acid pair tables (md05.csv) for the PID and QA side-chain depths.

Positions and lengths are in the library's internal units (1 A = 0.2).
Parsing is host-side input preparation: nothing here runs per step.
"""
from __future__ import annotations

import csv
import io
import math
import os

import numpy as np

from . import model as M
from . import params as P

A = P.ANGSTROM

# nat_cont::type (cg/include/cg/nat_cont/type.h:4)
ANY, BACK_BACK, BACK_SIDE, SIDE_BACK, SIDE_SIDE, SSBOND = range(6)
BACKBONE = ("N", "CA", "C", "O", "OXT")

# heavy-atom radii in A and per-residue overrides (the reference's default
# amino-acid data, cg/default.yml "default atom data" / "amino acids")
ATOM_RADII = {"C": 1.61, "CA": 1.88, "CB": 1.88, "CD": 1.88, "CD1": 1.61,
              "CD2": 1.76, "CE": 1.88, "CE1": 1.76, "CE2": 1.76, "CE3": 1.76,
              "CG": 1.88, "CG1": 1.88, "CG2": 1.88, "CH2": 1.76, "CZ": 1.61,
              "CZ2": 1.76, "CZ3": 1.73, "N": 1.64, "ND1": 1.64, "ND2": 1.64,
              "NE": 1.64, "NE1": 1.61, "NE2": 1.64, "NH1": 1.64, "NH2": 1.64,
              "NZ": 1.64, "O": 1.42, "OD1": 1.42, "OD2": 1.42, "OE1": 1.42,
              "OE2": 1.42, "OG": 1.46, "OG1": 1.46, "OH": 1.46, "SG": 1.77,
              "SD": 1.77}
ATOM_RADII_BY_RES = {"ASP": {"CG": 1.61, "OD1": 1.46}, "GLU": {"CD": 1.88, "OE1": 1.46},
                     "ILE": {"CD": 1.88}, "LEU": {"CD1": 1.88, "CD2": 1.88},
                     "PHE": {"CZ": 1.76},
                     "TRP": {"CD1": 1.76, "CD2": 1.61, "CE2": 1.64},
                     "TYR": {"CD1": 1.76, "CG": 1.61}}
RESIDUE_RADII = {"ALA": 3.35, "ARG": 3.95, "ASN": 3.65, "ASP": 3.5, "CYS": 3.7, "GLU": 3.65,
                 "GLN": 3.9, "GLY": 3.15, "HIS": 4.0, "ILE": 4.5, "LEU": 4.6, "LYS": 3.65,
                 "MET": 4.5, "PHE": 4.6, "PRO": 3.7, "SER": 3.3, "THR": 3.6, "TRP": 4.7,
                 "TYR": 4.5, "VAL": 4.0, "PHO": 3.5}


def _close(x: float, y: float) -> bool:
    return abs(x - y) / abs(x) < 1e-5


def _corrected(name: str, r: float) -> float:
    """Radius correction by element (cg/src/cg/amino/aa_data.cpp:16-31)."""
    if name[0] == "N":
        return 1.64 * A
    if name[0] == "S":
        return 1.77 * A
    if name[0] == "O" and not _close(r, 1.42 * A) and not _close(r, 1.46 * A):
        return 1.46 * A
    if (name[0] == "C" and not _close(r, 1.88 * A) and not _close(r, 1.76 * A)
            and not _close(r, 1.61 * A)):
        return 1.88 * A
    return r


def atom_radius(res: str, name: str) -> float:
    """Radius of atom `name` of residue `res`: the longest known name prefix,
    the residue's override first (aa_data.cpp:102-110)."""
    over = ATOM_RADII_BY_RES.get(res, {})
    for k in range(len(name), 0, -1):
        pre = name[:k]
        if pre in over:
            return _corrected(pre, over[pre] * A)
        if pre in ATOM_RADII:
            return _corrected(pre, ATOM_RADII[pre] * A)
    raise KeyError(f"no radius for atom {name} of {res}")


# ------------------------------------------------------------------- PDB ----
class PdbStructure:
    """Chains (chain id -> residues in order of appearance, each a dict with
    seq, name and atoms [(serial, name, pos)]), the CRYST1 cell and the
    SSBOND residue pairs [((chain, seq), (chain, seq))]."""

    def __init__(self):
        self.chains: dict[str, list[dict]] = {}
        self.cryst1: list[float] | None = None
        self.ssbonds: list[tuple] = []


def read_pdb(source) -> PdbStructure:
    """Fixed-column PDB records (pdb_file.cpp:14-84): ATOM of the first
    model (records after its ENDMDL are skipped), hydrogens (names starting
    with H) skipped, CRYST1, SSBOND; reading stops at END. Residues of a
    chain are told apart by their sequence numbers. `source` is a path or
    the file's text."""
    text = source if "\n" in source else open(source).read()
    st = PdbStructure()
    first_model_done = False
    for line in text.splitlines():
        rec = line[:6]
        if rec.startswith("END") and not rec.startswith("ENDMDL"):
            break
        if rec == "ENDMDL":
            first_model_done = True
        elif rec == "CRYST1":
            st.cryst1 = [float(line[6:15]) * A, float(line[15:24]) * A, float(line[24:33]) * A]
        elif rec == "SSBOND":
            st.ssbonds.append(((line[15], int(line[17:21])), (line[29], int(line[31:35]))))
        elif rec == "ATOM  " and not first_model_done:
            name = line[12:16].strip()
            if name[0] == "H":
                continue
            resn = line[17:20].strip()
            cid = line[21]
            seqn = int(line[22:26])
            pos = (float(line[30:38]) * A, float(line[38:46]) * A, float(line[46:54]) * A)
            res = st.chains.setdefault(cid, [])
            if not res or res[-1]["seq"] != seqn:
                res.append({"seq": seqn, "name": resn, "atoms": []})
            res[-1]["atoms"].append((int(line[6:11]), name, pos))
    return st


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _norm(a):
    return math.sqrt(_dot(a, a))


def _acos_ratio(num, den):
    """acos(num / den) clamped to [-1, 1]; NaN for a degenerate geometry
    (den 0), as the reference's FP division gives there."""
    if den == 0.0:
        return math.nan
    return math.acos(min(max(num / den, -1.0), 1.0))


def bond_angle(r1, r2, r3) -> float:
    """cg/include/cg/utils/angles.h:5-9"""
    u1, u2 = _sub(r2, r1), _sub(r2, r3)
    return _acos_ratio(_dot(u1, u2), _norm(u1) * _norm(u2))


def dihedral_angle(r1, r2, r3, r4) -> float:
    """cg/include/cg/utils/angles.h:11-19"""
    u1, u2, u3 = _sub(r2, r1), _sub(r3, r2), _sub(r4, r3)
    v1, v2 = _cross(u1, u2), _cross(u2, u3)
    phi = _acos_ratio(_dot(v1, v2), _norm(v1) * _norm(v2))
    return -phi if _dot(v1, u3) < 0 else phi


def _model_dict(pos, atype, chain, seq, ang, ang_has, ang_theta, dih, dih_has, dih_phi,
                cont, cont_len, cont_type, cell) -> dict:
    n = len(pos)
    pos = np.asarray(pos, np.float64).reshape(n, 3)
    chain = np.asarray(chain, np.int32)
    m = {"n": n, "pos": pos, "orig_pos": pos.copy(),
         "atype": np.asarray(atype, np.uint8), "chain": chain,
         "seq": np.asarray(seq, np.int32),
         "teth": np.asarray([[i, i + 1] for i in range(n - 1) if chain[i] == chain[i + 1]],
                            np.int32).reshape(-1, 2),
         "ang": np.asarray(ang, np.int32).reshape(-1, 3),
         "ang_has": np.asarray(ang_has, np.uint8),
         "ang_theta": np.asarray(ang_theta, np.float64),
         "dih": np.asarray(dih, np.int32).reshape(-1, 4),
         "dih_has": np.asarray(dih_has, np.uint8),
         "dih_phi": np.asarray(dih_phi, np.float64),
         "cont": np.asarray(cont, np.int32).reshape(-1, 2),
         "cont_len": np.asarray(cont_len, np.float64),
         "cont_type": np.asarray(cont_type, np.int32),
         "cell": np.asarray(cell if cell is not None else [0.0, 0.0, 0.0], np.float64)}
    return M.empty_tips(m)


def model_from_pdb(source, contacts: str = "all_atoms", load_structure: bool = True,
                   pbc: bool = False) -> dict:
    """Model of a PDB file (pdb_file::to_model, pdb_file.cpp:336-455).

    contacts: "all_atoms" (heavy-atom overlaps), "ca" (C-alpha atoms with
    residue radii) or "none" (SSBOND records only). load_structure False
    leaves the native angles and dihedrals unset (the heuristic terms take
    them), as an input without a structure would. pbc takes the CRYST1 cell.
    """
    st = read_pdb(source)
    pos, atype, chain, seq, seqnum, residues = [], [], [], [], [], []
    ang, ang_th, dih, dih_ph = [], [], [], []
    key = {}
    for ci, cid in enumerate(sorted(st.chains)):
        first = len(pos)
        for k, res in enumerate(st.chains[cid]):
            ca = [p for (_, nm, p) in res["atoms"] if nm == "CA"]
            if not ca:
                raise ValueError(f"residue {res['name']} {res['seq']} of chain {cid} has no CA")
            key[(cid, res["seq"])] = len(pos)
            pos.append(ca[0])
            atype.append(P.AA_INDEX[res["name"]])
            chain.append(ci)
            seq.append(k)
            seqnum.append(res["seq"])
            residues.append(res)
        last = len(pos)
        for i in range(first, last - 2):
            ang.append([i, i + 1, i + 2])
            ang_th.append(bond_angle(pos[i], pos[i + 1], pos[i + 2]))
            if i + 3 < last:
                dih.append([i, i + 1, i + 2, i + 3])
                dih_ph.append(dihedral_angle(pos[i], pos[i + 1], pos[i + 2], pos[i + 3]))
    has_a = [1 if load_structure else 0] * len(ang)
    has_d = [1 if load_structure else 0] * len(dih)
    if not load_structure:
        ang_th = [0.0] * len(ang)
        dih_ph = [0.0] * len(dih)

    cont, clen, ctype, taken = [], [], [], set()
    for (a, b) in st.ssbonds:
        i, j = sorted((key[a], key[b]))
        if (i, j) in taken:
            continue
        taken.add((i, j))
        cont.append([i, j])
        clen.append(_norm(_sub(pos[i], pos[j])))
        ctype.append(SSBOND)
    if contacts != "none" and load_structure:
        alpha = (26.0 / 7.0) ** (1.0 / 6.0)
        all_atoms = contacts == "all_atoms"
        atoms = []   # (residue index, name, radius, pos), chains in id order, atoms in serial order
        for ri, res in enumerate(residues):
            for (serial, nm, p) in res["atoms"]:
                if not all_atoms and nm != "CA":
                    continue
                rad = atom_radius(res["name"], nm) if all_atoms else RESIDUE_RADII[res["name"]] * A
                atoms.append((ri, nm, rad, p))
        for (r1, n1, rad1, p1) in atoms:
            for (r2, n2, rad2, p2) in atoms:
                if r1 == r2 or abs(seqnum[r1] - seqnum[r2]) < 3:
                    continue
                if _norm(_sub(p1, p2)) <= (rad1 + rad2) * alpha:
                    i, j = min(r1, r2), max(r1, r2)
                    if (i, j) in taken:
                        continue
                    taken.add((i, j))
                    b1, b2 = n1 in BACKBONE, n2 in BACKBONE
                    cont.append([i, j])
                    clen.append(_norm(_sub(pos[i], pos[j])))
                    ctype.append(BACK_BACK if b1 and b2 else BACK_SIDE if b1 else
                                 SIDE_BACK if b2 else SIDE_SIDE)
    order = sorted(range(len(cont)), key=lambda k: (cont[k][0], cont[k][1]))
    cont = [cont[k] for k in order]
    clen = [clen[k] for k in order]
    ctype = [ctype[k] for k in order]
    cell = st.cryst1 if (pbc and st.cryst1) else None
    return _model_dict(pos, atype, chain, seq, ang, has_a, ang_th, dih, has_d, dih_ph,
                       cont, clen, ctype, cell)


# --------------------------------------------------------- sequence files ----
def _resolve(node, base: str):
    """An `(at path): file` node becomes the file's content (YAML parsed,
    other files as text), paths relative to the including file."""
    import yaml
    if isinstance(node, dict) and set(node) == {"(at path)"}:
        path = os.path.join(base, node["(at path)"])
        with open(path) as f:
            text = f.read()
        if path.endswith((".yml", ".yaml")):
            return yaml.safe_load(text), os.path.dirname(path)
        return text, os.path.dirname(path)
    return node, base


def _table(node, base: str) -> list[dict]:
    node, _ = _resolve(node, base)
    if node is None:
        return []
    rows = list(csv.DictReader(io.StringIO(node.strip())))
    return [{k.strip(): v.strip() for k, v in r.items()} for r in rows]


def read_map(node, base: str = ".", shift: int = 0) -> dict:
    """A contact map (map_file.cpp:26-53): contacts, angles, dihedrals, each
    index shifted by `shift` (map_file.cpp:6-24)."""
    node, base = _resolve(node, base)
    out = {"contacts": [], "angles": [], "dihedrals": []}
    for r in _table(node.get("contacts"), base):
        out["contacts"].append((int(r["i1"]) + shift, int(r["i2"]) + shift, float(r["length"])))
    for r in _table(node.get("angles"), base):
        out["angles"].append((int(r["i1"]) + shift, int(r["i2"]) + shift, int(r["i3"]) + shift,
                              float(r["theta"])))
    for r in _table(node.get("dihedrals"), base):
        out["dihedrals"].append((int(r["i1"]) + shift, int(r["i2"]) + shift,
                                 int(r["i3"]) + shift, int(r["i4"]) + shift, float(r["phi"])))
    return out


def model_from_seq_file(source) -> dict:
    """Model of a sequence file (seq_file::load, seq_file.cpp:8-138).

    Per chain: residues from the one-letter sequence (whitespace removed),
    the maps' contacts (type ANY, chain-relative indices) and native angles
    and dihedrals first, then a tether for every consecutive pair and a
    heuristic angle/dihedral for every triple/quadruple no map gave. All
    residues sit at the origin: morph_into_saw gives the start.
    """
    import yaml
    if "\n" in source:
        doc, base = yaml.safe_load(source), "."
    else:
        with open(source) as f:
            doc = yaml.safe_load(f)
        base = os.path.dirname(os.path.abspath(source))
    atype, chain, seq = [], [], []
    ang, ang_has, ang_th, dih, dih_has, dih_ph = [], [], [], [], [], []
    cont, clen = [], []
    for ci, ch in enumerate(doc["model"]["chains"]):
        ch, cbase = _resolve(ch, base)
        codes = "".join(str(ch["seq"]).split())
        first = len(atype)
        for k, c in enumerate(codes):
            atype.append(P.LETTER_INDEX[c])
            chain.append(ci)
            seq.append(k)
        taken_a, taken_d = set(), set()
        for entry in ch.get("maps") or []:
            if isinstance(entry, dict) and "__source" in entry:
                mp = read_map(entry["__source"], cbase, int(entry.get("shift", 0)))
            else:
                mp = read_map(entry, cbase)
            for (i1, i2, length) in mp["contacts"]:
                cont.append([first + i1, first + i2])
                clen.append(length)
            for (i1, i2, i3, th) in mp["angles"]:
                ang.append([first + i1, first + i2, first + i3])
                ang_has.append(1)
                ang_th.append(th)
                taken_a.add((i1, i2, i3))
            for (i1, i2, i3, i4, ph) in mp["dihedrals"]:
                dih.append([first + i1, first + i2, first + i3, first + i4])
                dih_has.append(1)
                dih_ph.append(ph)
                taken_d.add((i1, i2, i3, i4))
        n = len(codes)
        for i in range(n - 2):
            if (i, i + 1, i + 2) not in taken_a:
                ang.append([first + i, first + i + 1, first + i + 2])
                ang_has.append(0)
                ang_th.append(0.0)
            if i + 3 < n and (i, i + 1, i + 2, i + 3) not in taken_d:
                dih.append([first + i, first + i + 1, first + i + 2, first + i + 3])
                dih_has.append(0)
                dih_ph.append(0.0)
    pos = [[0.0, 0.0, 0.0]] * len(atype)
    return _model_dict(pos, atype, chain, seq, ang, ang_has, ang_th, dih, dih_has, dih_ph,
                       cont, clen, [ANY] * len(cont), None)


# ------------------------------------------------------------------ SAW ----
def morph_into_saw(m: dict, seed: int, bond: float | None = None,
                   intersection: float = 4.56 * A, box_size: float = 0.0,
                   density: float = 0.0, pbc: bool = False,
                   update_orig: bool | None = None) -> tuple[dict, int]:
    """The model with its chains redrawn as self-avoiding walks
    (model::morph_into_saw through pcg_morph_saw) and the ran2 draws used,
    for pcg_options.ran2_skip. bond None takes the mean native tether
    length of the model (model.cpp:112-125); update_orig (default: True when
    a bond length is given, as for a sequence file, whose tethers have no
    native length) makes the original positions follow the walk
    (state.cpp:134-135)."""
    from . import pcg as G
    if bond is None:
        t = m["teth"]
        d = np.linalg.norm(m["orig_pos"][t[:, 1]] - m["orig_pos"][t[:, 0]], axis=1)
        bond = float(d.sum() / len(d))
        upd = False if update_orig is None else update_orig
    else:
        upd = True if update_orig is None else update_orig
    pos, draws = G.morph_saw(m["chain"], seed, bond, intersection, box_size, density, pbc)
    out = dict(m)
    out["pos"] = pos
    if upd:
        out["orig_pos"] = pos.copy()
    if pbc:
        side = box_size if box_size > 0 else (m["n"] / density) ** (1.0 / 3.0)
        out["cell"] = np.array([side, side, side], np.float64)
    return out, draws


# ------------------------------------------------------------ pair tables ----
def read_pair_table(source) -> list[list[float]]:
    """A 21x21 amino-acid pair table (md05.csv: a header of residue names,
    then one row per residue)."""
    text = source if "\n" in source else open(source).read()
    return P.parse_pair_table(text)
