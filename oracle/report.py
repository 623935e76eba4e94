"""Oracle of the report scalars: a numpy restatement of the reference's
make_report (cg/src/cg/output/make_report.cpp:132-247), for tests only.

Test infrastructure: only tests/ (and the bench's CPU leg) use anything under
oracle/; the library computes the same scalars on the device (pcg_report).
Each function follows the reference function it names, with numpy's LAPACK
SVD in place of Eigen::JacobiSVD (both give the singular values and vectors
of the same matrices to rounding).
"""
from __future__ import annotations

import numpy as np


def rmsd(orig_r: np.ndarray, cur_r: np.ndarray) -> float:
    """make_report::rmsd (make_report.cpp:132-158): centre both, H = orig^T cur,
    R = V diag(1, 1, d) U^T with d = sign det(V^T U), RMSD of cur R - orig."""
    o = orig_r - orig_r.mean(axis=0)
    c = cur_r - cur_r.mean(axis=0)
    H = o.T @ c
    U, _, Vt = np.linalg.svd(H)
    V = Vt.T
    d = np.sign(np.linalg.det(V.T @ U))
    R = V @ np.diag([1.0, 1.0, d]) @ U.T
    return float(np.sqrt(((c @ R - o) ** 2).sum(axis=1).mean()))


def kinetic_energy(v: np.ndarray, mass: np.ndarray) -> float:
    """make_report::kinetic_energy (make_report.cpp:160-166)."""
    return float((0.5 * mass * (v * v).sum(axis=1)).sum())


def gyration_radius(r: np.ndarray) -> float:
    """make_report::gyration_radius (make_report.cpp:168-179); the sum is
    divided by (float)N."""
    mean_r = r.sum(axis=0) / len(r)
    return float(np.sqrt(((r - mean_r) ** 2).sum() / float(np.float32(len(r)))))


def asphericity(r: np.ndarray) -> float:
    """make_report::asphericity (make_report.cpp:181-194): lambda = the square
    roots of the singular values of the centred positions,
    W = 1.5 lambda_x^2 - 0.5 |lambda|^2."""
    pos = r - r.mean(axis=0)
    lam = np.sqrt(np.linalg.svd(pos, compute_uv=False))
    return float(1.5 * lam[0] ** 2 - 0.5 * (lam ** 2).sum())


def scalars(t, V, r, orig_r, v, mass, nc_active=None, nc_type=None, qa_pairs=None, chain=None):
    """make_report::add_cur_scalars (make_report.cpp:196-247) for one system;
    qa_pairs are the (i1, i2) of the non-vacant QA contacts."""
    K = kinetic_energy(v, mass)
    out = {"TIME": t, "EPOT": V, "ETOT": V + K, "EKIN": K, "RG": gyration_radius(r),
           "L": float(np.linalg.norm(r[-1] - r[0])), "RMSD": rmsd(orig_r, r),
           "W": asphericity(r)}
    if nc_active is not None:
        out["ICN"] = int(np.count_nonzero(nc_active))
        out["ICNss"] = int(np.count_nonzero(np.asarray(nc_active) & (np.asarray(nc_type) == 5)))
    if qa_pairs is not None:
        same = [chain[a] == chain[b] for a, b in qa_pairs]
        out["INTRHC"] = int(sum(same))
        out["INTEHC"] = int(len(same) - sum(same))
    return out
