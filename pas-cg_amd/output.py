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
    """A cell as the SL4 writer prints it (sl4.cpp:15-44): the value goes
    through its stream text and back, so a float prints as printf "%g" and
    a non-finite one, which the stream does not read back, as 0; booleans
    as T/F, integers as integers, text as is."""
    import math
    import numpy as np
    if isinstance(v, (bool, np.bool_)):
        return "T" if v else "F"
    if isinstance(v, (int, np.integer)):
        return "%d" % int(v)
    if isinstance(v, (float, np.floating)):
        v = float(v)
        return "%g" % v if math.isfinite(v) else "0"
    return str(v)


def sl4_table(rows: list[dict]) -> str:
    """A table in the SL4 layout (sl4.cpp:65-123): columns in the order they
    first appear, a cell a row does not set left blank, every column right-
    aligned to its widest entry; a table without rows prints its empty
    header, "#"."""
    cols = []
    for r in rows:
        for c in r:
            if c not in cols:
                cols.append(c)
    cells = [[_field(r[c]) if c in r else "" for c in cols] for r in rows]
    width = [max([len(c)] + [len(row[k]) for row in cells]) for k, c in enumerate(cols)]
    head = "#" + "   ".join(c.upper().rjust(w) for c, w in zip(cols, width))
    body = [" " + "   ".join(f.rjust(w) for f, w in zip(row, width)) for row in cells]
    return "\n".join([head] + body)


def _comment(*parts) -> str:
    """An SL4 comment (sl4.h:75-83): its parts as stream text (floats as
    "%g"), upper-cased, after '#'."""
    text = "".join(("%g" % p) if isinstance(p, float) else str(p) for p in parts)
    return "#" + text.upper()


def out_file(trajectories: list[list[dict]], afm: list | None = None,
             results: str | None = None) -> str:
    """Text of a `.out` file (make_report::emit_out, make_report.cpp:355-359):
    per trajectory its comment and scalars table, then, for a trajectory
    that reported AFM tips, the force AFM and velocity AFM divs
    (afm[idx] = (force rows, velocity rows) of afm_rows, accumulated over
    the reports; None without AFM); after the last trajectory the results
    div of at_simul_end when `results` (results_div) is given."""
    divs = []
    for idx, rows in enumerate(trajectories):
        parts = ["#TRAJECTORY %d" % idx, sl4_table(rows)]
        tips = afm[idx] if afm is not None and idx < len(afm) else None
        if tips is not None:
            parts += ["#FORCE AFM TIPS", sl4_table(tips[0]),
                      "#VELOCITY AFM TIPS", sl4_table(tips[1])]
        divs.append("\n".join(parts))
    if results is not None:
        divs.append(results)
    return "\n".join(divs)


def afm_rows(stats: dict, tips: dict) -> tuple[list[dict], list[dict]]:
    """One report's rows of the force AFM and velocity AFM tables
    (make_report::add_afm_stuff, make_report.cpp:264-311) from the tips'
    statistics (pcg.Context.afm_stats: velocity tips, then force tips) and
    the tips in place (tip_res, ftip_res): per force tip its residue and the
    norm of the average velocity in A/tau, per velocity tip its residue, the
    norm of the average force and the perpendicular force in eps/A; -1 for
    an average that has no value yet."""
    import numpy as np
    vres = list(np.asarray(tips.get("tip_res", []), np.int64))
    fres = list(np.asarray(tips.get("ftip_res", []), np.int64))
    nv = len(vres)
    vel_unit = A / P.TAU
    force_unit = P.EPS / A
    force_rows, vel_rows = [], []
    for k, res in enumerate(fres):
        j = nv + k
        avg = float(np.linalg.norm(stats["avg"][j])) / vel_unit if stats["has"][j] else -1.0
        force_rows.append({"res_idx": int(res), "avg_vel [A/tau]": avg})
    for k, res in enumerate(vres):
        has = bool(stats["has"][k])
        vel_rows.append({"res_idx": int(res),
                         "avg_force [eps/A]": (float(np.linalg.norm(stats["avg"][k])) / force_unit
                                               if has else -1.0),
                         "avg_perp_force [eps/A]": (float(stats["perp"][k]) / force_unit
                                                    if has else -1.0)})
    return force_rows, vel_rows


NC_TYPES = ["any", "bb", "bs", "sb", "ss", "ssbond"]


def nc_type_name(t: int) -> str:
    """format_nc_type (make_report.cpp:58-75) of a nat_cont::type value
    (ANY, BACK_BACK, BACK_SIDE, SIDE_BACK, SIDE_SIDE, SSBOND)."""
    return NC_TYPES[t] if 0 <= t < len(NC_TYPES) else ""


def qa_type_name(t: int) -> str:
    """format_qa_type (make_report.cpp:473-482) of a qa::contact_type value
    (1 back-back, 2 back-side, 3 side-back, side-side otherwise)."""
    return {1: "bb", 2: "bs", 3: "sb"}.get(int(t), "ss")


def results_div(contacts: dict | None = None, nc_times: list | None = None,
                unfold_times: list | None = None, num_traj: int = 1) -> str:
    """The results div the `.out` file ends with (make_report::at_simul_end,
    make_report.cpp:77-128). With the unfolding study's time measurement on
    (contacts given: i1, i2, type, nat_dist of every native contact), the
    average breaking/forming time of each contact over the trajectories
    that changed it (change_t >= 0 in nc_times, one array a trajectory;
    -1 if none did), rows sorted by (time, contact index), and the median
    unfolding time: entry num_traj // 2 of the sorted unfold_times (the
    latest change time of each trajectory whose every contact changed),
    N/A when there are not that many."""
    out = ["#RESULTS"]
    if contacts is None:
        return "\n".join(out)
    k = len(contacts["i1"])
    rows = []
    for idx in range(k):
        s, num = 0.0, 0
        for tr in nc_times or []:
            if tr[idx] >= 0.0:
                s += float(tr[idx])
                num += 1
        rows.append((s / num / P.TAU if num > 0 else -1.0, idx))
    rows.sort()
    tab = []
    for avg, idx in rows:
        tab.append({"i1": int(contacts["i1"][idx]), "i2": int(contacts["i2"][idx]),
                    "type": nc_type_name(int(contacts["type"][idx])),
                    "nat_dist[A]": float(contacts["nat_dist"][idx]) / A,
                    "avg_time[tau]": float(avg)})
    out += ["#AVERAGE BREAKING/FORMING TIMES", sl4_table(tab)]
    times = sorted(float(x) for x in (unfold_times or []))
    mid = num_traj // 2
    out.append("#MEDIAN (UN)FOLDING TIME: N/A" if mid >= len(times)
               else _comment("median (un)folding time: ", times[mid]))
    return "\n".join(out)


def unfolding_sample(active_change_t, num_changed: int):
    """What traj_end_step records of one trajectory for the unfolding study
    (thread_main.cpp:340-351): every native contact's change time, and the
    unfolding time (the latest change time) when every contact changed,
    else None. active_change_t = pcg.Context.nat_cont_state(k)[1]."""
    import numpy as np
    ct = np.asarray(active_change_t, np.float64).copy()
    t = None
    if num_changed == len(ct):
        t = 0.0
        for x in ct:
            t = max(t, float(x))
    return ct, t


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


# ---------------------------------------------------------------- map files --
# make_report::add_map_data (make_report.cpp:499-656): per report a div with
# the time, the native contacts, the QA contacts, the active PID bundles and
# every angle and dihedral of the model at the current positions.
SEQ_CHAIN_IDS = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz!\"#$%&\\'()*+,-./"


def _c_round(x):
    """C's round(): halves away from zero."""
    import numpy as np
    t = np.trunc(x)
    return np.where(np.abs(x - t) >= 0.5, t + np.sign(x), t)


def _wrapped_dist(pos, cell, i1, i2):
    """|pbc.wrap(r[i1], r[i2])| (cg/include/cg/sbox/pbc.h:39-50): each axis
    with a cell takes the nearest image."""
    import numpy as np
    c = np.asarray(cell, np.float64)
    inv = np.where(c != 0.0, 1.0 / np.where(c != 0.0, c, 1.0), 0.0)
    d = pos[np.asarray(i2)] - pos[np.asarray(i1)]
    d = d - c * _c_round(d * inv)
    return np.sqrt((d * d).sum(axis=-1))


def _unit(v):
    import numpy as np
    return (1.0 / np.sqrt((v * v).sum(axis=-1)))[..., None] * v


def map_angles(pos, ang):
    """angle_value (make_report.cpp:484-487): the angle between the bond
    vectors r2 - r1 and r3 - r2 of each triple, in radians."""
    import numpy as np
    ang = np.asarray(ang, np.int64).reshape(-1, 3)
    u1 = _unit(pos[ang[:, 1]] - pos[ang[:, 0]])
    u2 = _unit(pos[ang[:, 2]] - pos[ang[:, 1]])
    with np.errstate(invalid="ignore"):
        return np.arccos((u1 * u2).sum(axis=-1))


def map_dihedrals(pos, dih):
    """dihedral_value (make_report.cpp:489-497): the angle between the
    planes' normals r12 x r23 and r23 x r34, negative when r12 x r23 points
    against r34."""
    import numpy as np
    dih = np.asarray(dih, np.int64).reshape(-1, 4)
    r12 = pos[dih[:, 1]] - pos[dih[:, 0]]
    r23 = pos[dih[:, 2]] - pos[dih[:, 1]]
    r34 = pos[dih[:, 3]] - pos[dih[:, 2]]
    x1, x2 = np.cross(r12, r23), np.cross(r23, r34)
    with np.errstate(invalid="ignore"):
        phi = np.arccos((_unit(x1) * _unit(x2)).sum(axis=-1))
    return np.where((x1 * r34).sum(axis=-1) < 0.0, -phi, phi)


def pid_active(pos, cell, i1, i2, atype, prm: dict):
    """pid::eval_forces::is_active (pid/eval_forces.cpp:204-212) of bundles
    (i1, i2): distance <= counting factor * the side-chain pair's r_min *
    2^(-1/6)."""
    import numpy as np
    i1, i2 = np.asarray(i1, np.int64), np.asarray(i2, np.int64)
    at = np.asarray(atype, np.int64)
    r_high = np.asarray(prm["pid_ss_r_high"], np.float64).reshape(-1)[at[i1] * 21 + at[i2]]
    sigma = r_high * P.C216_INV
    return _wrapped_dist(pos, cell, i1, i2) <= prm["counting_factor"] * sigma


def _res_cells(m: dict, chain_ids: str, i: int, k: str) -> dict:
    ch = int(m["chain"][i])
    return {"chain" + k: chain_ids[ch], "seq" + k: int(m["seq"][i]) + 1,
            "res" + k: P.AA_NAMES[int(m["atype"][i])]}


def _pair_rows(m, chain_ids, pos, cell, i1, i2):
    rows = []
    dist = _wrapped_dist(pos, cell, i1, i2) if len(i1) else []
    for a, b, d in zip(i1, i2, dist):
        row = _res_cells(m, chain_ids, int(a), "1")
        row.update(_res_cells(m, chain_ids, int(b), "2"))
        # the reference labels a same-chain pair "inter" (make_report.cpp:531)
        same = int(m["chain"][int(a)]) == int(m["chain"][int(b)])
        row["inter/intra"] = "inter" if same else "intra"
        rows.append((row, float(d) / A))
    return rows


def map_div(t: float, pos, m: dict, cell=None, chain_ids: str | None = None,
            nat: dict | None = None, qa: dict | None = None, pid: dict | None = None) -> str:
    """One report's div of the `.map` file. m is the model (chain, seq,
    atype, ang/ang_has/ang_theta, dih/dih_has/dih_phi); chain ids default to
    the model's (chain_ids) or the sequence-file order A, B, ... Sections
    present only when their term is on: nat (all native contacts: i1, i2,
    type, active, change_t), qa (live contacts in slot order: i1, i2,
    type), pid (all bundles: i1, i2, active)."""
    import numpy as np
    pos = np.asarray(pos, np.float64)
    cell = np.asarray(cell if cell is not None else m.get("cell", [0.0, 0.0, 0.0]), np.float64)
    ids = chain_ids or m.get("chain_ids") or SEQ_CHAIN_IDS
    out = [_comment("T = ", float(t))]
    if nat is not None:
        rows = []
        for (row, d), ty, act, ct in zip(_pair_rows(m, ids, pos, cell, nat["i1"], nat["i2"]),
                                         nat["type"], nat["active"], nat["change_t"]):
            row["type"] = nc_type_name(int(ty))
            row["dist[A]"] = d
            row["active"] = bool(act)
            row["time[tau]"] = float(ct) / P.TAU
            rows.append(row)
        out += ["#NATIVE CONTACTS", _comment("n = ", int(np.count_nonzero(nat["active"]))),
                sl4_table(rows)]
    if qa is not None:
        rows = []
        for (row, d), ty in zip(_pair_rows(m, ids, pos, cell, qa["i1"], qa["i2"]), qa["type"]):
            row["type"] = qa_type_name(int(ty))
            row["dist[A]"] = d
            rows.append(row)
        out += ["#QA", _comment("n = ", len(rows)), sl4_table(rows)]
    if pid is not None:
        act = np.asarray(pid["active"], bool)
        i1 = np.asarray(pid["i1"])[act]
        i2 = np.asarray(pid["i2"])[act]
        rows = []
        for row, d in _pair_rows(m, ids, pos, cell, i1, i2):
            row["dist[A]"] = d
            rows.append(row)
        out += ["#PID", _comment("n = ", len(rows)), sl4_table(rows)]
    rows = []
    ang = np.asarray(m["ang"], np.int64).reshape(-1, 3)
    theta = map_angles(pos, ang)
    for k in range(len(ang)):
        i = int(ang[k, 0])
        row = {"i1": i}
        row.update(_res_cells(m, ids, i, "1"))
        row["theta"] = float(theta[k])
        if m["ang_has"][k]:
            row["native"] = float(m["ang_theta"][k])
        rows.append(row)
    out += ["#ANGLES", sl4_table(rows)]
    rows = []
    dih = np.asarray(m["dih"], np.int64).reshape(-1, 4)
    phi = map_dihedrals(pos, dih)
    for k in range(len(dih)):
        i = int(dih[k, 0])
        row = {"i1": i}
        row.update(_res_cells(m, ids, i, "1"))
        row["phi"] = float(phi[k])
        if m["dih_has"][k]:
            row["native"] = float(m["dih_phi"][k])
        rows.append(row)
    out += ["#DIHEDRALS", sl4_table(rows)]
    return "\n".join(out)


def map_file(trajectories: list[list[str]]) -> str:
    """Text of a `.map` file (make_report::emit_map, make_report.cpp:658-662):
    per trajectory its comment, then its reports' divs (map_div)."""
    return "\n".join("\n".join(["#TRAJECTORY %d" % idx] + divs)
                     for idx, divs in enumerate(trajectories))


def map_snapshot(ctx, m: dict, prm: dict, cell=None, chain_ids: str | None = None) -> str:
    """map_div of a pcg.Context's current state: positions and time, the
    native contacts' state (pcg_nat_cont_state), the live QA contacts with
    their types (pcg_get_list kind 7) and the PID bundles (kind 4), each
    when its term is on."""
    import numpy as np
    t, _ = ctx.time()
    pos = ctx.r()
    nat = qa = pid = None
    cont = np.asarray(m["cont"], np.int64).reshape(-1, 2)
    if prm.get("nat_cont_on") and len(cont):
        act, ct, _ = ctx.nat_cont_state(len(cont))
        nat = {"i1": cont.min(axis=1), "i2": cont.max(axis=1), "type": m["cont_type"],
               "active": act, "change_t": ct}
    if prm.get("qa_on"):
        i1, i2, ty, vac = ctx.list(7)
        live = vac == 0
        qa = {"i1": i1[live], "i2": i2[live], "type": ty[live].astype(np.int64)}
    c = np.asarray(cell if cell is not None else m.get("cell", [0.0, 0.0, 0.0]), np.float64)
    if prm.get("pid_on"):
        i1, i2, _, _ = ctx.list(4)
        pid = {"i1": i1, "i2": i2, "active": pid_active(pos, c, i1, i2, m["atype"], prm)}
    return map_div(t, pos, m, c, chain_ids, nat, qa, pid)


# ----------------------------------------------------------- report driver --
class Reporter:
    """The reference's report files for a run of a pcg.Context, as
    make_report (cg/src/cg/output/make_report.cpp:29-56) writes them: every
    `stats_every` of simulated time a row of the trajectory's scalars table,
    its AFM rows (AFM on) and a `.map` div, then `<prefix>.out` and
    `<prefix>.map` rewritten; every `struct_every` a snapshot, then
    `<prefix>.pdb` rewritten with all of them. A trajectory's end adds one
    of each (emit_all, thread_main.cpp:336-356) and the unfolding study's
    sample; `simul_end` appends the results div (make_report.cpp:77-128).

    The reference reports inside a step, after its force evaluation and
    before the QA finish and the Langevin step (thread_main.cpp:503-520),
    once t > 0. `run` advances the context to the first step whose time
    passes the next report time and reports from the state between steps:
    the same positions, time, velocities and QA contacts, and the energies
    and native-contact flags of the evaluation one step earlier (the step
    that produced those positions)."""

    def __init__(self, prefix: str, model: dict, params: dict, stats_every: float | None = None,
                 struct_every: float | None = None, chain_ids: str | None = None,
                 measure_times: bool = False, num_traj: int = 1):
        import numpy as np
        self.prefix, self.m, self.p = prefix, model, params
        self.stats_every, self.struct_every = stats_every, struct_every
        self.chain_ids = chain_ids
        self.measure_times, self.num_traj = measure_times, num_traj
        cont = np.asarray(model["cont"], np.int64).reshape(-1, 2)
        self.contacts = {"i1": cont.min(axis=1) if len(cont) else cont[:, 0],
                         "i2": cont.max(axis=1) if len(cont) else cont[:, 0],
                         "type": np.asarray(model["cont_type"]), "nat_dist": np.asarray(model["cont_len"])}
        self.trajs: list[dict] = []
        self.nc_times: list = []
        self.unfold_times: list = []
        self.results: str | None = None

    # -- bookkeeping of make_report (report.cpp:6-16)
    def traj_init(self):
        import math
        self.trajs.append({"rows": [], "afm": None, "maps": [], "snaps": []})
        self.stats_last = self.struct_last = -math.inf

    def _cur(self) -> dict:
        if not self.trajs:
            self.traj_init()
        return self.trajs[-1]

    def _flags(self):
        p = self.p
        return (bool(p.get("nat_cont_on")) and len(self.contacts["i1"]) > 0, bool(p.get("qa_on")),
                bool(p.get("pid_on")))

    def add_stats(self, ctx, tips: dict | None = None):
        """add_cur_scalars, add_afm_stuff and add_map_data, then the files."""
        tr = self._cur()
        nat, qa, pid = self._flags()
        tr["rows"].append(scalars_row(ctx.report(), nat, qa, pid))
        if self.p.get("afm_on"):
            f, v = afm_rows(ctx.afm_stats(), tips or {})
            if tr["afm"] is None:
                tr["afm"] = ([], [])
            tr["afm"][0].extend(f)
            tr["afm"][1].extend(v)
        cell = self.m.get("cell")
        tr["maps"].append(map_snapshot(ctx, self.m, self.p, cell, self.chain_ids))
        self.emit_out()
        self.emit_map()

    def add_snapshot(self, ctx):
        """add_cur_snapshot, then the `.pdb` file."""
        import numpy as np
        t, _ = ctx.time()
        self._cur()["snaps"].append((t, np.asarray(ctx.r(), np.float64).copy()))
        self.emit_pdb()

    def emit_out(self):
        with open(self.prefix + ".out", "w") as f:
            f.write(out_file([tr["rows"] for tr in self.trajs], afm=[tr["afm"] for tr in self.trajs],
                             results=self.results))

    def emit_map(self):
        with open(self.prefix + ".map", "w") as f:
            f.write(map_file([tr["maps"] for tr in self.trajs]))

    def emit_pdb(self):
        snaps = [s for tr in self.trajs for s in tr["snaps"]]
        cell = self.m.get("cell")
        cell = cell if cell is not None and any(float(c) != 0.0 for c in cell) else None
        with open(self.prefix + ".pdb", "w") as f:
            f.write(pdb_snapshots(snaps, self.m["chain"], self.m["atype"], cell))

    # -- make_report::operator() (make_report.cpp:29-46)
    def __call__(self, ctx, tips: dict | None = None):
        t, _ = ctx.time()
        if not t > 0:
            return
        if self.stats_every is not None and t >= self.stats_last + self.stats_every:
            self.add_stats(ctx, tips)
            self.stats_last = t
        if self.struct_every is not None and t >= self.struct_last + self.struct_every:
            self.add_snapshot(ctx)
            self.struct_last = t

    def _next_steps(self, t: float, dt: float) -> int:
        """Steps until the next report falls due (at least one); 0 with no
        reports configured."""
        from .protocol import steps_until
        due = []
        if self.stats_every is not None:
            due.append(self.stats_last + self.stats_every)
        if self.struct_every is not None:
            due.append(self.struct_last + self.struct_every)
        if not due:
            return 0
        return max(steps_until(t, min(due), dt), 1)

    def run(self, ctx, steps: int, tips: dict | None = None):
        """Advances the context by `steps` steps, reporting whenever a
        report falls due on the way."""
        self._cur()
        dt = self.p["dt"]
        self(ctx, tips)
        done = 0
        while done < steps:
            t, _ = ctx.time()
            k = self._next_steps(t, dt)
            k = steps - done if k == 0 else min(k, steps - done)
            ctx.step(k)
            done += k
            self(ctx, tips)

    def traj_end(self, ctx, tips: dict | None = None):
        """thread::traj_end_step: the unfolding study's sample, then emit_all."""
        if self.measure_times and len(self.contacts["i1"]):
            act, ct, changed = ctx.nat_cont_state(len(self.contacts["i1"]))
            times, t = unfolding_sample(ct, changed)
            self.nc_times.append(times)
            if t is not None:
                self.unfold_times.append(t)
        self.add_stats(ctx, tips)
        self.add_snapshot(ctx)

    def simul_end(self):
        """at_simul_end: the results div, then the `.out` file."""
        self.results = results_div(self.contacts if self.measure_times else None, self.nc_times,
                                   self.unfold_times, self.num_traj)
        self.emit_out()
