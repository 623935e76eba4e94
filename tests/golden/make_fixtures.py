"""Writes the golden fixtures under tests/golden/ from the reference's own
parity test inputs and traces (parity/tests/*/ in the reference tree).

Run in a container where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference

Each fixture holds the compiled model (positions in internal units, residue
types, chains, native angles/dihedrals, native contacts, PBC cell), the seed
and SAW settings, and the reference's per-step total potential energy trace
(V_fort and V_cxx columns of comp.csv) for the first steps. Only data is
written; nothing of the reference's code is copied.

The models come from the package's own input readers
(pas-cg_amd/inputs.py: PDB and sequence files as the reference reads them,
cg/src/cg/input/pdb_file.cpp and seq_file.cpp), so the fixtures also pin
those readers against the reference's traces.
"""
from __future__ import annotations

import csv
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def _package():
    """The package's input readers (pas-cg_amd/inputs.py), which restate the
    reference's PDB and sequence-file semantics."""
    import pcg_import
    pcg_import.load()
    from pas_cg_amd import inputs, params
    return inputs, params


def _fixture_model(m, cryst1):
    """The JSON model block of a fixture from a package model dict."""
    ang = [[int(i), int(j), int(k), int(h), float(t)]
           for (i, j, k), h, t in zip(m["ang"], m["ang_has"], m["ang_theta"])]
    dih = [[int(i), int(j), int(k), int(l), int(h), float(p)]
           for (i, j, k, l), h, p in zip(m["dih"], m["dih_has"], m["dih_phi"])]
    cont = [[int(i), int(j), float(d), int(t)]
            for (i, j), d, t in zip(m["cont"], m["cont_len"], m["cont_type"])]
    return {"n": int(m["n"]), "pos": m["pos"].tolist(), "orig_pos": m["orig_pos"].tolist(),
            "atype": m["atype"].astype(int).tolist(), "chain": m["chain"].astype(int).tolist(),
            "seq": m["seq"].astype(int).tolist(), "teth": m["teth"].astype(int).tolist(),
            "ang": ang, "dih": dih, "cont": cont, "cryst1": cryst1}


def model_from_pdb(path, all_atom_contacts=False, load_structure=True):
    inputs, _ = _package()
    m = inputs.model_from_pdb(path, contacts="all_atoms" if all_atom_contacts else "none",
                              load_structure=load_structure)
    return _fixture_model(m, inputs.read_pdb(path).cryst1)


def model_from_seq(seq_file):
    inputs, _ = _package()
    return _fixture_model(inputs.model_from_seq_file(seq_file), None)


def read_pair_csv(path):
    inputs, _ = _package()
    return inputs.read_pair_table(path)


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
                                               ["V_fort_7"], 20000))

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

    m = model_from_seq(os.path.join(pt, "9aac/cxx/9aac.yml"))
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
