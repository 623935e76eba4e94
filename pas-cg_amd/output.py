"""The reference's text outputs, written from the library's reports.

* The `.out` file's per-trajectory `scalars` table
  (cg/src/cg/output/report.cpp:6-16, make_report.cpp:196-262): a comment
  line `#TRAJECTORY <idx>`, then the table in the SL4 layout
  (cg/src/cg/files/table/sl4.cpp:15-123): a header of upper-cased column
  names after '#', rows after ' ', every field right-aligned to its column's
  widest entry, fields three spaces apart, floats as printf "%g", integers
  as integers. The columns, in the reference's order and units:
  TIME[tau] EPOT[eps] ETOT[eps] RG[A] L[A] RMSD[A] W[A**2], then ICN ICNss
  with native contacts on, INTEHC INTRHC with QA on, and the six PID counts
  with PID on. Trajectories are separated by a blank line (sl4 div).
* The raw-data dump (cg/src/cg/output/print_raw_data.cpp:4-9): every step
  whose index is a multiple of the dump period, the line
  `traj_idx step_idx t V` of that step's force evaluation, 17 significant
  digits.

Host-side formatting only; the values come from pcg_report /
pcg_batch_report and pcg_get_energy.
"""
from __future__ import annotations

from . import params as P

A = P.ANGSTROM
PID_COLS = ["PID-INTRHC", "PID-B1-B1", "PID-B1-S1", "PID-INTEHC", "PID-B1-B2", "PID-S1-S2"]


def scalars_row(rep: dict, nat_cont: bool, qa: bool, pid: bool) -> dict:
    """One row of the scalars table from a report dict (REPORT_FIELDS)."""
    row = {"TIME[tau]": float(rep["TIME"]), "EPOT[eps]": float(rep["EPOT"]),
           "ETOT[eps]": float(rep["ETOT"]), "RG[A]": float(rep["RG"]) / A,
           "L[A]": float(rep["L"]) / A, "RMSD[A]": float(rep["RMSD"]) / A,
           "W[A**2]": float(rep["W"]) / (A * A)}
    if nat_cont:
        row["ICN"] = int(rep["ICN"])
        row["ICNss"] = int(rep["ICNss"])
    if qa:
        row["INTEHC"] = int(rep["INTEHC"])
        row["INTRHC"] = int(rep["INTRHC"])
    if pid:
        for k in PID_COLS:
            row[k] = int(rep[k])
    return row


def _field(v) -> str:
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, int):
        return "%d" % v
    if isinstance(v, float):
        return "%g" % v
    return str(v)


def sl4_table(rows: list[dict]) -> str:
    """A table in the SL4 layout (sl4.cpp:65-123); columns in the order the
    first row gives them."""
    if not rows:
        return ""
    cols = list(rows[0])
    cells = [[_field(r.get(c, "")) for c in cols] for r in rows]
    width = [max([len(c)] + [len(row[k]) for row in cells]) for k, c in enumerate(cols)]
    head = "#" + "   ".join(c.upper().rjust(w) for c, w in zip(cols, width))
    body = [" " + "   ".join(f.rjust(w) for f, w in zip(row, width)) for row in cells]
    return "\n".join([head] + body)


def out_file(trajectories: list[list[dict]]) -> str:
    """Text of a `.out` file: per trajectory its comment and scalars table."""
    divs = []
    for idx, rows in enumerate(trajectories):
        divs.append("#TRAJECTORY %d\n%s" % (idx, sl4_table(rows)))
    return "\n".join(divs)


def raw_data_line(traj_idx: int, step_idx: int, t: float, V: float) -> str:
    """print_raw_data's line (std::setprecision(17), default float format)."""
    return "%d %d %.17g %.17g\n" % (traj_idx, step_idx, t, V)


def run_with_raw_dump(ctx, steps: int, period: int, out, traj_idx: int = 0) -> None:
    """Advances a pcg.Context by `steps` steps, writing print_raw_data's line
    to `out` for every step whose index is a multiple of `period`."""
    if period <= 0:
        raise ValueError("dump period must be positive")
    done = 0
    while done < steps:
        t, step = ctx.time()
        if step % period == 0:
            ctx.step(1)
            done += 1
            V, _ = ctx.energy()
            out.write(raw_data_line(traj_idx, step, t, V))
            continue
        k = min(period - step % period, steps - done)
        ctx.step(k)
        done += k


# ------------------------------------------------------------ PDB snapshots --
# Fixed-column records as cg/src/cg/input/records.cpp writes them: 80-column
# lines, every field printf-formatted right-justified at its columns
# (cg/src/cg/input/fields.cpp), lines joined by newlines with none at the end.
def _put(line: list, col: int, text: str) -> None:
    """Writes text from 1-based column col, as format_ copies it."""
    beg = col - 1
    if beg + len(text) > len(line):
        line.extend(" " * (beg + len(text) - len(line)))
    line[beg:beg + len(text)] = list(text)


def _record(name: str) -> list:
    line = [" "] * 80
    _put(line, 1, "%6s" % name)
    return line


def _chain_scalars(t: float, pos, idx, cell) -> dict:
    """The per-chain row of make_report::emit_pdb (make_report.cpp:405-425)."""
    import numpy as np
    r = pos[idx]
    mean = r.sum(axis=0) / len(idx)
    rg = float(np.sqrt(((r - mean) ** 2).sum() / float(np.float32(len(idx)))))
    d = r[-1] - r[0]
    if cell is not None:
        c = np.asarray(cell, np.float64)
        inv = np.where(c != 0.0, 1.0 / np.where(c != 0.0, c, 1.0), 0.0)
        d = d - c * np.round(d * inv)
    cen = r - r.mean(axis=0)
    lam = np.sqrt(np.linalg.svd(cen, compute_uv=False)) if len(idx) > 1 else np.zeros(3)
    w = float(1.5 * lam[0] ** 2 - 0.5 * (lam ** 2).sum()) if len(lam) else 0.0
    return {"T[tau]": float(t), "Rg[A]": rg / A, "R_end_to_end[A]": float(np.linalg.norm(d)) / A,
            "N": len(idx), "W[A**2]": w / (A * A)}


def pdb_snapshots(snapshots, chain, atype, cell=None) -> str:
    """Text of the `.pdb` file of make_report::emit_pdb (make_report.cpp:385-465)
    for snapshots [(t, positions [n, 3])]: per snapshot MODEL, CRYST1 (the
    cell in A, zero without one), then per chain (ids A, B, ... in chain
    order) two REMARK 1 lines with that chain's T, Rg, end-to-end distance,
    N and W in the SL4 layout, its C-alpha ATOM records (serials running
    over the model, one more for each chain's TER, residue numbers from 1)
    and TER; ENDMDL."""
    import numpy as np
    chain = np.asarray(chain)
    ids = sorted(set(chain.tolist()))
    out = []
    for s, (t, pos) in enumerate(snapshots):
        pos = np.asarray(pos, np.float64)
        rows = [_chain_scalars(t, pos, np.nonzero(chain == c)[0], cell) for c in ids]
        cells = [[_field(r[k]) for k in rows[0]] for r in rows]
        width = [max([len(k)] + [len(row[j]) for row in cells]) for j, k in enumerate(rows[0])]
        head = " " + "   ".join(k.upper().rjust(w) for k, w in zip(rows[0], width))
        line = _record("MODEL ")
        _put(line, 11, "%4d" % (s + 1))
        out.append(line)
        line = _record("CRYST1")
        c3 = np.asarray(cell if cell is not None else [0.0, 0.0, 0.0], np.float64) / A
        for col, v in zip((7, 16, 25), c3):
            _put(line, col, "%9s" % ("%9.3f" % v))
        for col in (34, 41, 48):
            _put(line, col, "%7s" % ("%7.2f" % 90.0))
        _put(line, 56, "%11s" % "P 1")
        _put(line, 67, "%4d" % 1)
        out.append(line)
        serial = 1
        for ci, c in enumerate(ids):
            body = " " + "   ".join(f.rjust(w) for f, w in zip(cells[ci], width))
            for text in (head, body):
                for off in range(0, len(text), 69):
                    line = _record("REMARK")
                    _put(line, 8, "%3d" % 1)
                    _put(line, 12, "%69s" % text[off:off + 69])
                    out.append(line)
            members = np.nonzero(chain == c)[0]
            cid = chr(ord("A") + ci)
            for k, i in enumerate(members):
                line = _record("ATOM  ")
                _put(line, 7, "%5s" % str(serial))
                _put(line, 13, "%4s" % "CA")
                _put(line, 18, "%3s" % P.AA_NAMES[int(atype[i])])
                line[21] = cid
                _put(line, 23, "%4d" % (k + 1))
                for col, v in zip((31, 39, 47), pos[i] / A):
                    _put(line, col, "%8s" % ("%8.3f" % v))
                out.append(line)
                serial += 1
            line = _record("TER   ")
            _put(line, 7, "%5s" % str(serial))
            _put(line, 18, "%3s" % P.AA_NAMES[int(atype[members[-1]])])
            line[21] = cid
            _put(line, 23, "%4d" % len(members))
            out.append(line)
            serial += 1
        out.append(_record("ENDMDL"))
    return "\n".join("".join(line) for line in out)
