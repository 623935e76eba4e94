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
