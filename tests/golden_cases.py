"""Parameter sets and loaders for the reference's parity fixtures
(parity/tests/<name>/cxx/inputfile.yml merged over cg/default.yml)."""
from __future__ import annotations

import json
import os

import numpy as np

from pas_cg_amd import model as M
from pas_cg_amd import params as P

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str) -> dict:
    with open(os.path.join(HERE, f"{name}.json")) as f:
        return json.load(f)


def case(name: str):
    """Returns (params, model, seed, saw, fixture)."""
    fx = load(name)
    p = P.default_params()
    if name == "start2":
        p.update(qa_on=0, nat_cont_on=0, dh_on=0, temperature=0.38)
    elif name in ("TEST", "sample"):
        p.update(qa_on=0, nat_cont_on=0, temperature=0.38, pid_on=1)
        P.set_pid_forces(p, "sink lj", ss_depth_table=fx["md05"])
    elif name == "input_test":
        p.update(qa_on=0, nat_cont_on=0, temperature=0.38, pid_on=1, pid_include4=0)
        P.set_pid_forces(p, "sink lj", ss_depth_table=fx["md05"])
    elif name == "9aac":
        p.update(qa_on=0, temperature=0.38, pid_on=1, pid_include4=0)
        P.set_pid_forces(p, "sink lj", bbp_depth=0.1, bbm_depth=0.1,
                         ss_depth_table=fx["md05"])
    elif name == "1ubq":
        p.update(qa_on=0, dh_on=0, rep_cutoff=4.0 * P.ANGSTROM, counting_factor=1.5)
    m = M.from_fixture(fx["model"], pbc=fx["pbc"])
    saw = None
    if fx["saw"]:
        s = fx["saw"]
        if s["bond"] == "avg_tether":
            d = np.linalg.norm(m["pos"][m["teth"][:, 1]] - m["pos"][m["teth"][:, 0]], axis=1)
            bond, upd = d.sum() / len(d), False
        else:
            bond, upd = P.f32(s["bond"]) * P.ANGSTROM, True
        saw = dict(bond=bond, intersection=s["intersection"] * P.ANGSTROM, update_orig=upd)
    return p, m, fx["seed"], saw, fx
