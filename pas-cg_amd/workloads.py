"""Seeded synthetic workloads of the per-step loop's benchmark.

Every workload is built from `model.synthetic_protein` (job j of a workload
uses the seed splitmix64(448 xor j); 448 is the reference's default seed,
cg/default.yml:22):

  c0      one chain of 56 residues; Verlet cutoff 18 A, pad 2 A, Pauli
          repulsion at 4 A, native LJ, tether, native angles and dihedrals,
          chirality; T = 0.35, gamma = 2, dt = 0.005 (float-parsed).
  c1      64 copies of one 100-residue chain in a periodic cubic box at 1 mM
          (side (64 / (1 mM N_A))^(1/3) = 473.3 A), adding Debye-Hueckel
          (screening 10 A, eps_r = 80), quasi-adiabatic contacts and PID.
  c2      1024 independent replicas of the c0 chain.
  c3      256 distinct proteins with lengths uniform in [30, 600], the full
          force field, 10 % of them pulled at 0.005 A/tau by a k = 0.1 eps/A^2
          tip on the terminal residue.
  c4      a Poisson stream of jobs: lengths lognormal (median 150, capped at
          2000), step counts log-uniform, 20 % pulling runs, the rest folding
          at T uniform in [0.3, 0.45].

The native conformation is the compact self-avoiding walk of
`synthetic_protein`, compacted by 2000 steps of Go-model annealing on the
GPU (`anneal_natives`): Langevin dynamics in the Go model of the walk itself
(tethers, native contacts of the walk, its bond angles and dihedrals, local
repulsion, Pauli exclusion) with the temperature lowered geometrically from
0.6 to 0.15 eps/kB over ten stages; the native map, angles and dihedrals are
then taken from the annealed conformation. These generators prepare input;
they are not part of the per-step loop.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import model as M
from . import params as P

A = P.ANGSTROM
AVOGADRO = 6.02214076e23


def base_params() -> dict:
    """Force field of the single-chain workload (c0)."""
    p = P.default_params()
    p.update(nl_cutoff=18.0 * A, nl_pad=2.0 * A, rep_cutoff=4.0 * A, chir_on=1,
             qa_on=0, pid_on=0, dh_on=0, afm_on=0, temperature=P.f32(0.35))
    return p


def full_params(pad_A: float = 2.0) -> dict:
    """Full force field of the ragged and served workloads (c3, c4): native LJ,
    Pauli, constant-permittivity Debye-Hueckel (screening 10 A, eps_r = 80),
    quasi-adiabatic contacts, PID, angles, dihedrals, chirality and a
    0.1 eps/A^2 velocity-AFM spring."""
    p = base_params()
    p.update(nl_pad=pad_A * A, dh_on=1, dh_variant=0, dh_screen_dist=10.0 * A,
             dh_perm=80.0 * P.EPS_0, qa_on=1, pid_on=1, afm_on=1,
             afm_H1=0.1 / (A * A), afm_H2=0.0)
    return p


def c0(chain: dict | None = None) -> tuple[dict, dict]:
    """The single chain (`chain`: the same chain after annealing)."""
    return base_params(), chain if chain is not None else M.synthetic_protein(56, M.job_seed(0))


def c1_chain(n: int = 100) -> dict:
    return M.synthetic_protein(n, M.job_seed(1))


def c1(copies: int = 64, n: int = 100, chain: dict | None = None) -> tuple[dict, dict]:
    """The periodic box of `copies` copies of one chain (`chain`: c1_chain
    after annealing)."""
    p = full_params()
    p.update(afm_on=0)
    side = (copies / (1e-3 * AVOGADRO * 1e3)) ** (1.0 / 3.0) * 1e10  # metres -> A
    one = chain if chain is not None else c1_chain(n)
    return p, M.replicate(one, copies, side * A, seed=1)


def c2(copies: int = 1024, chain: dict | None = None) -> tuple[dict, list[dict]]:
    p, m = c0(chain)
    return p, [m] * copies


def _c3_job(args):
    j, n, pulled = args
    m = M.synthetic_protein(int(n), M.job_seed(j))
    if pulled:
        m = M.add_pulling(m)
    return m


def ragged_lengths(count: int = 256, lo: int = 30, hi: int = 600, seed: int = 448) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.integers(lo, hi + 1, size=count)


def c3(count: int = 256, workers: int | None = None, pulled: bool = True
       ) -> tuple[dict, list[dict]]:
    """The ragged batch on the walks' own natives; every 10th protein is pulled
    (pulled=False: none)."""
    ns = ragged_lengths(count)
    jobs = [(j, ns[j], pulled and j % 10 == 0) for j in range(count)]
    workers = workers or min(8, os.cpu_count() or 1)
    if workers > 1 and count > 8:
        with ProcessPoolExecutor(workers) as ex:
            models = list(ex.map(_c3_job, jobs, chunksize=4))
    else:
        models = [_c3_job(a) for a in jobs]
    return full_params(), models


def c4_jobs(count: int, rate_per_s: float, steps_lo: int, steps_hi: int,
            seed: int = 448, max_n: int = 2000) -> list[dict]:
    """Job descriptions of a Poisson stream: arrival time (s), length,
    steps, temperature, pulled."""
    rng = np.random.Generator(np.random.PCG64(seed ^ 0xC4))
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=count))
    n = np.clip(np.round(150.0 * np.exp(0.6 * rng.normal(size=count))), 30, max_n)
    steps = np.exp(rng.uniform(math.log(steps_lo), math.log(steps_hi), size=count))
    pulled = rng.uniform(size=count) < 0.2
    T = rng.uniform(0.3, 0.45, size=count)
    return [{"id": j, "arrival": float(t[j]), "n": int(n[j]), "steps": int(steps[j]),
             "T": float(T[j]), "pulled": bool(pulled[j])} for j in range(count)]


def c4_model(job: dict) -> dict:
    """Job j's protein (not yet annealed, no tips: c4_finish adds them)."""
    return M.synthetic_protein(job["n"], M.job_seed(10_000 + job["id"]))


def c4_finish(jobs: list, models: list) -> list:
    """The stream's proteins on Go-annealed natives, pulled jobs with their
    tips. Needs the GPU."""
    models = anneal_natives(models)
    return [M.add_pulling(m) if j["pulled"] else m for j, m in zip(jobs, models)]


def go_params() -> dict:
    """Go model of the annealing: tether, native contacts, native angles and
    dihedrals, local repulsion and Pauli exclusion; no QA, PID, DH or AFM."""
    p = base_params()
    p.update(chir_on=0, heur_ang_on=0, heur_dih_on=0, pauli_on=1)
    return p


def renative(m: dict, pos: np.ndarray, native_cutoff_A: float = 7.5) -> dict:
    """The model with `pos` as its native (and starting) conformation: native
    contacts |i-j| >= 3 closer than the cutoff, bond angles and dihedrals
    from `pos` (one chain, as synthetic_protein makes them)."""
    out = dict(m)
    pos = np.asarray(pos, np.float64) - np.asarray(pos, np.float64).mean(axis=0)
    out["pos"] = pos
    out["orig_pos"] = pos.copy()
    out["ang_theta"] = np.array([M._bond_angle(pos[a], pos[b], pos[c]) for a, b, c in m["ang"]])
    out["dih_phi"] = np.array([M._dihedral(pos[a], pos[b], pos[c], pos[d])
                               for a, b, c, d in m["dih"]])
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    ii, jj = np.nonzero(np.triu(d < native_cutoff_A * A, k=3))
    out["cont"] = np.stack([ii, jj], axis=1).astype(np.int32)
    out["cont_len"] = d[ii, jj].astype(np.float64)
    out["cont_type"] = np.zeros(len(ii), np.int32)
    return M.empty_tips(out)


def anneal_natives(models: list, steps: int = 2000, t_hot: float = 0.6, t_cold: float = 0.15,
                   stages: int = 10, seed: int = 448) -> list:
    """Go-model annealing of the synthetic natives on the GPU: all models in
    one batch (mixed precision), `stages` equal stages of `steps` in all with
    the temperature lowered geometrically from t_hot to t_cold (by the noise
    scale sqrt(T / t_hot)); returns the models rebuilt on the annealed
    conformations (tips dropped: add them after)."""
    from . import pcg as G
    p = go_params()
    b = G.Batch(p, models, temps=[t_hot] * len(models), seeds=[seed + k for k in range(len(models))],
                precision=G.MIXED, philox_seed=seed, steps_per_launch=max(1, steps // stages))
    per = max(1, steps // stages)
    for k in range(stages):
        T = t_hot * (t_cold / t_hot) ** (k / max(1, stages - 1))
        b.set_noise_scale(math.sqrt(T / t_hot))
        b.run(per)
    out = [renative(m, b.get(s, 0)) for s, m in enumerate(models)]
    b.close()
    return out


def c3_annealed(count: int = 256, workers: int | None = None) -> tuple[dict, list[dict]]:
    """The ragged batch on Go-annealed natives; every 10th protein is pulled
    (its tips placed on the annealed conformation). Needs the GPU."""
    p, models = c3(count, workers, pulled=False)
    models = anneal_natives(models)
    return p, [M.add_pulling(m) if j % 10 == 0 else m for j, m in enumerate(models)]


# algorithmic bytes per list item of each term (index ints plus parameters)
ITEM_BYTES = {"chir": 24, "tether": 12, "lrep": 8, "nat_cont": 13, "pauli": 8,
              "const_dh": 12, "rel_dh": 12, "qa_cand": 12, "qa_cont": 30, "pid": 10,
              "nat_ang": 16, "heur_ang": 13, "heur_dih": 17, "cnd": 20, "snd": 20,
              "vel_afm": 24}
RESIDUE_BYTES = 361  # Gear y0..y5 r+w 288, positions 24, forces 24, orig 24, type 1


def term_enabled(p: dict) -> dict:
    dh = bool(p["dh_on"])
    return {"chir": bool(p["chir_on"]), "tether": bool(p["tether_on"]),
            "lrep": bool(p["lrep_on"]), "nat_cont": bool(p["nat_cont_on"]),
            "pauli": bool(p["pauli_on"]) and not p["qa_on"],
            "const_dh": dh and p["dh_variant"] == 0, "rel_dh": dh and p["dh_variant"] == 1,
            "qa_cand": bool(p["qa_on"]), "qa_cont": bool(p["qa_on"]),
            "pid": bool(p["pid_on"]), "nat_ang": bool(p["nat_ang_on"]),
            "heur_ang": bool(p["heur_ang_on"]),
            "heur_dih": bool(p["heur_dih_on"]),
            "cnd": bool(p["nat_dih_on"]) and p["nat_dih_variant"] == 0,
            "snd": bool(p["nat_dih_on"]) and p["nat_dih_variant"] == 1,
            "vel_afm": bool(p["afm_on"])}


def step_bytes(p: dict, sizes: dict, rebuilds_per_step: float = 0.0) -> float:
    """Algorithmic HBM bytes of one step of one system:
    N * 361 + sum_t L_t * b_t + 12 * |NL| per rebuild, amortised."""
    on = term_enabled(p)
    b = sizes["n"] * RESIDUE_BYTES
    for t, nb in ITEM_BYTES.items():
        if on[t]:
            b += sizes[t] * nb
    return b + 12.0 * sizes["nl"] * rebuilds_per_step
