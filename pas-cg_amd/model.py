"""Protein models handed to the C-ABI: residues, chains and per-term lists.

A model is a dict of numpy arrays, the compiled form of the reference's
`input::model` after `state::compile_model` (cg/src/cg/simul/state.cpp:144-249):

  n           number of residues
  pos         [n,3] float64  starting positions (internal units, 1 = 5 A)
  orig_pos    [n,3] float64  positions the tether lengths come from
  atype       [n]   uint8    residue type (aa_code order)
  chain, seq  [n]   int32    chain index and index within the chain
  teth        [k,2] int32    consecutive residues of a chain
  ang         [k,3] int32,  ang_has [k] uint8, ang_theta [k] float64
  dih         [k,4] int32,  dih_has [k] uint8, dih_phi [k] float64
  cont        [k,2] int32,  cont_len [k] float64, cont_type [k] int32
  cell        [3]   float64  PBC cell (0 turns an axis off)
  tip_res [t] int32, tip_orig [t,3], tip_vel [t,3]   velocity AFM tips

`synthetic_protein` generates the seeded synthetic chains of the benchmark
workloads. It is a host-side input generator, not part of the per-step loop.
"""
from __future__ import annotations

import math

import numpy as np

from . import params as P

ANGSTROM = P.ANGSTROM


def empty_tips(m: dict) -> dict:
    m["tip_res"] = np.zeros(0, np.int32)
    m["tip_orig"] = np.zeros((0, 3), np.float64)
    m["tip_vel"] = np.zeros((0, 3), np.float64)
    m["ftip_res"] = np.zeros(0, np.int32)
    m["ftip_force"] = np.zeros((0, 3), np.float64)
    return m


def from_fixture(fm: dict, pbc: bool = True) -> dict:
    """Model from a golden fixture's JSON model block."""
    n = fm["n"]
    m = {"n": n,
         "pos": np.asarray(fm["pos"], np.float64).reshape(n, 3),
         "orig_pos": np.asarray(fm["orig_pos"], np.float64).reshape(n, 3),
         "atype": np.asarray(fm["atype"], np.uint8),
         "chain": np.asarray(fm["chain"], np.int32),
         "seq": np.asarray(fm["seq"], np.int32),
         "teth": np.asarray(fm["teth"], np.int32).reshape(-1, 2)}
    ang = np.asarray(fm["ang"], np.float64).reshape(-1, 5)
    m["ang"] = ang[:, :3].astype(np.int32)
    m["ang_has"] = ang[:, 3].astype(np.uint8)
    m["ang_theta"] = ang[:, 4].copy()
    dih = np.asarray(fm["dih"], np.float64).reshape(-1, 6)
    m["dih"] = dih[:, :4].astype(np.int32)
    m["dih_has"] = dih[:, 4].astype(np.uint8)
    m["dih_phi"] = dih[:, 5].copy()
    cont = np.asarray(fm["cont"], np.float64).reshape(-1, 4)
    m["cont"] = cont[:, :2].astype(np.int32)
    m["cont_len"] = cont[:, 2].copy()
    m["cont_type"] = cont[:, 3].astype(np.int32)
    c = fm.get("cryst1")
    m["cell"] = np.asarray(c if (pbc and c) else [0.0, 0.0, 0.0], np.float64)
    return empty_tips(m)


def chain_lists(chain: np.ndarray):
    """Tethers, angles and dihedrals over consecutive residues of each chain
    (pdb_file.cpp:365-397, seq_file.cpp:103-136)."""
    n = len(chain)
    teth, ang, dih = [], [], []
    start = 0
    while start < n:
        end = start
        while end < n and chain[end] == chain[start]:
            end += 1
        for i in range(start, end - 1):
            teth.append((i, i + 1))
            if i + 2 < end:
                ang.append((i, i + 1, i + 2))
                if i + 3 < end:
                    dih.append((i, i + 1, i + 2, i + 3))
        start = end
    return (np.asarray(teth, np.int32).reshape(-1, 2),
            np.asarray(ang, np.int32).reshape(-1, 3),
            np.asarray(dih, np.int32).reshape(-1, 4))


def _bond_angle(r1, r2, r3):
    u1, u2 = r2 - r1, r2 - r3
    c = float(np.dot(u1, u2) / (np.linalg.norm(u1) * np.linalg.norm(u2)))
    return math.acos(min(max(c, -1.0), 1.0))


def _dihedral(r1, r2, r3, r4):
    u1, u2, u3 = r2 - r1, r3 - r2, r4 - r3
    v1, v2 = np.cross(u1, u2), np.cross(u2, u3)
    c = float(np.dot(v1, v2) / (np.linalg.norm(v1) * np.linalg.norm(v2)))
    phi = math.acos(min(max(c, -1.0), 1.0))
    return -phi if np.dot(v1, u3) < 0 else phi


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def job_seed(j: int, base: int = 448) -> int:
    """Seed of synthetic job j: splitmix64(448 xor j); 448 is the reference's
    default seed (cg/default.yml:22)."""
    return splitmix64(base ^ j)


def _compact_saw(n: int, rng: np.random.Generator, bond: float, excl: float,
                 trials: int = 24) -> np.ndarray:
    """Self-avoiding chain with fixed bond length grown towards its own centre
    of mass (a Rosenbluth-style biased growth): among `trials` random bond
    directions giving a C-alpha bond angle between 85 and 150 degrees (so that
    i, i+2 stay beyond 5 A), the one closest to the running centroid that keeps
    every |i-j| >= 3 pair farther than `excl` is taken. When the chain traps
    itself it backs up a few residues (more after repeated traps) and regrows."""
    pos = np.zeros((n, 3))
    csum = np.zeros(3)
    i, fails = 1, 0
    while i < n:
        cen = csum / i
        dirs = rng.normal(size=(trials, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        if i >= 2:
            prev = pos[i - 1] - pos[i - 2]
            prev /= np.linalg.norm(prev)
            c = dirs @ prev  # = -cos(bond angle)
            dirs = dirs[(c > -0.0872) & (c < 0.866)]
        cand = pos[i - 1] + bond * dirs
        if i >= 3 and len(cand):
            d = np.linalg.norm(cand[:, None, :] - pos[None, :i - 2, :], axis=2)
            cand = cand[(d > excl).all(axis=1)]
        if len(cand) == 0:
            fails += 1
            if fails > 50 * n:
                raise RuntimeError("could not grow a self-avoiding chain")
            back = min(i - 1, 2 + int(rng.integers(0, 4 + fails // 8)))
            for k in range(i - back, i):
                csum -= pos[k]
            i -= back
            continue
        k = int(np.argmin(np.linalg.norm(cand - cen, axis=1)))
        pos[i] = cand[k]
        csum += pos[i]
        i += 1
    return pos


def synthetic_protein(n: int, seed: int, native_cutoff_A: float = 7.5,
                      bond_A: float = 3.8, charged: bool = True,
                      heuristic_frac: float = 0.0) -> dict:
    """Seeded synthetic C-alpha protein: residue types uniform over the 20
    standard amino acids, a compact self-avoiding native conformation with
    3.8 A bonds, native contacts |i-j| >= 3 with native distance < 7.5 A
    (type ANY), native bond angles and dihedrals from that conformation.
    The simulation starts from the native conformation.

    heuristic_frac > 0 drops the native value of that (seeded) fraction of
    the bond angles and dihedrals, which then take the heuristic potentials
    of their residue pair (a model read from a sequence file without a map
    has none, seq_file.cpp:103-136; state.cpp:337-352)."""
    rng = np.random.Generator(np.random.PCG64(seed & 0xFFFFFFFFFFFFFFFF))
    atype = rng.integers(0, 20, size=n).astype(np.uint8)
    bond = P.f32(bond_A) * ANGSTROM
    pos = _compact_saw(n, rng, bond, 4.56 * ANGSTROM)
    pos -= pos.mean(axis=0)
    chain = np.zeros(n, np.int32)
    seq = np.arange(n, dtype=np.int32)
    teth, ang, dih = chain_lists(chain)
    ang_theta = np.array([_bond_angle(pos[a], pos[b], pos[c]) for a, b, c in ang])
    dih_phi = np.array([_dihedral(pos[a], pos[b], pos[c], pos[d]) for a, b, c, d in dih])
    cut = native_cutoff_A * ANGSTROM
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    ii, jj = np.nonzero(np.triu(d < cut, k=3))
    cont = np.stack([ii, jj], axis=1).astype(np.int32)
    m = {"n": n, "pos": pos, "orig_pos": pos.copy(), "atype": atype, "chain": chain,
         "seq": seq, "teth": teth, "ang": ang,
         "ang_has": np.ones(len(ang), np.uint8), "ang_theta": ang_theta,
         "dih": dih, "dih_has": np.ones(len(dih), np.uint8), "dih_phi": dih_phi,
         "cont": cont, "cont_len": d[ii, jj].astype(np.float64),
         "cont_type": np.zeros(len(cont), np.int32),
         "cell": np.zeros(3)}
    if heuristic_frac > 0:
        hr = np.random.Generator(np.random.PCG64((seed ^ 0x4E55) & 0xFFFFFFFFFFFFFFFF))
        m["ang_has"][hr.uniform(size=len(ang)) < heuristic_frac] = 0
        m["dih_has"][hr.uniform(size=len(dih)) < heuristic_frac] = 0
    return empty_tips(m)


def add_disulfides(m: dict, count: int, seed: int = 0) -> dict:
    """Marks `count` native contacts (seeded choice among those whose residues
    are not in another marked contact) as disulfide bonds: both residues
    become CYS and the contact type SSBOND (5), as the PDB reader types the
    contacts of SSBOND records (pdb_file.cpp:404-420)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    used: set = set()
    picks = []
    for k in rng.permutation(len(m["cont"])):
        a, b = (int(x) for x in m["cont"][k])
        if a in used or b in used:
            continue
        picks.append(int(k))
        used.update((a, b))
        if len(picks) == count:
            break
    m["cont_type"] = np.asarray(m["cont_type"], np.int32).copy()
    m["atype"] = np.asarray(m["atype"], np.uint8).copy()
    for k in picks:
        m["cont_type"][k] = 5
        m["atype"][m["cont"][k]] = P.AA_INDEX["CYS"]
    return m


def replicate(m: dict, copies: int, box_side: float, seed: int,
              min_sep: float = 4.56 * ANGSTROM) -> dict:
    """`copies` randomly placed and rotated copies of one chain in a cubic
    periodic box of side `box_side` (each copy is its own chain), no two
    residues of different copies closer than `min_sep`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = m["n"]
    out = {k: [] for k in ("pos", "atype", "chain", "seq", "teth", "ang", "ang_has",
                           "ang_theta", "dih", "dih_has", "dih_phi", "cont",
                           "cont_len", "cont_type")}
    for c in range(copies):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, cc, d = q
        R = np.array([[a*a+b*b-cc*cc-d*d, 2*(b*cc-a*d), 2*(b*d+a*cc)],
                      [2*(b*cc+a*d), a*a-b*b+cc*cc-d*d, 2*(cc*d-a*b)],
                      [2*(b*d-a*cc), 2*(cc*d+a*b), a*a-b*b-cc*cc+d*d]])
        placed = np.concatenate(out["pos"]) if out["pos"] else np.zeros((0, 3))
        for _try in range(2000):
            shift = rng.uniform(-box_side / 2, box_side / 2, size=3)
            cand = m["pos"] @ R.T + shift
            if len(placed) == 0:
                break
            d = cand[:, None, :] - placed[None, :, :]
            d -= box_side * np.round(d / box_side)
            if np.sqrt((d * d).sum(-1)).min() > min_sep:
                break
        else:
            raise RuntimeError("box too crowded to place the copies")
        out["pos"].append(cand)
        off = c * n
        out["atype"].append(m["atype"])
        out["chain"].append(np.full(n, c, np.int32))
        out["seq"].append(m["seq"])
        for k in ("teth", "ang", "dih", "cont"):
            out[k].append(m[k] + off)
        for k in ("ang_has", "ang_theta", "dih_has", "dih_phi", "cont_len", "cont_type"):
            out[k].append(m[k])
    r = {k: np.concatenate(v) for k, v in out.items()}
    r["n"] = n * copies
    r["orig_pos"] = r["pos"].copy()
    r["cell"] = np.array([box_side] * 3, np.float64)
    return empty_tips(r)


def add_pulling(m: dict, vel_A_per_tau: float = 0.005) -> dict:
    """Constant-velocity pulling of the last residue of chain 0 away from the
    first one, whose position is held by a resting tip
    (state::setup_afm, cg/src/cg/simul/state.cpp:548-568)."""
    first = 0
    last = int(np.nonzero(m["chain"] == m["chain"][0])[0][-1])
    r0, r1 = m["pos"][first], m["pos"][last]
    u = (r1 - r0) / np.linalg.norm(r1 - r0)
    m["tip_res"] = np.array([first, last], np.int32)
    m["tip_orig"] = np.stack([r0, r1]).astype(np.float64)
    m["tip_vel"] = np.stack([np.zeros(3), vel_A_per_tau * ANGSTROM * u])
    return m


def add_force_pulling(m: dict, force_eps_per_A: float = 1.0) -> dict:
    """Constant-force pulling (afm type "const force"): the first residue of
    chain 0 held by a resting velocity tip, the last one pulled by a force of
    the given size along the direction from the first to the last
    (state::setup_afm, cg/src/cg/simul/state.cpp:548-568)."""
    first = 0
    last = int(np.nonzero(m["chain"] == m["chain"][0])[0][-1])
    r0, r1 = m["pos"][first], m["pos"][last]
    u = (r1 - r0) / np.linalg.norm(r1 - r0)
    m["tip_res"] = np.array([first], np.int32)
    m["tip_orig"] = r0[None].astype(np.float64)
    m["tip_vel"] = np.zeros((1, 3), np.float64)
    m["ftip_res"] = np.array([last], np.int32)
    m["ftip_force"] = (force_eps_per_A / ANGSTROM * u)[None].astype(np.float64)
    return m
