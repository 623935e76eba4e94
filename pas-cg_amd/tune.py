"""Autotuner of the batched form: Verlet skin (NL pad) and workgroup size.

The workgroup size only changes the order of the force additions, so it is
tuned in place by `pcg_batch_autotune` (trials on the live state, restored
after each).

The pad is a physical parameter whenever QA or PID is on. The reference
resets total_disp to 0 after every verify (cg/src/cg/nl/verify.cpp:31-32),
so the prunes `orig_dist - total_disp > max_req_dist` of
qa/loop_over_candidates.cpp:20 and pid/eval_forces.cpp:22 test a pair's
distance at the last rebuild, and the pad decides how long a pair that
started far away can approach before QA or PID sees it. A seeded ensemble
run on the GPU (96 proteins of 60-300 residues with QA on, the same Philox
noise at 2 A and at 12 A) measured 108.8 QA contacts per protein at 2 A and
104.3 at 12 A, a paired difference of -4.6 +- 0.6. The pad is therefore kept
at its configured value when QA or PID is on, and swept only when neither
is (every other list keeps the pairs its force can reach, so the pad then
changes the cost per step and nothing else).
"""
from __future__ import annotations

from . import pcg as G
from . import params as P


def pad_is_physical(params: dict) -> bool:
    """Whether the NL pad changes the dynamics (QA or PID on)."""
    return bool(params.get("qa_on")) or bool(params.get("pid_on"))


def tune(params: dict, models: list, pads_A=(4.0, 6.0, 8.0, 10.0, 12.0), trial_steps: int = 500,
         precision: int = G.MIXED, philox_seed: int = 1, log=None) -> dict:
    """Best (pad, block) by measured time per step; returns the choice and
    the whole table. With QA or PID on only the configured pad is tried."""
    if pad_is_physical(params):
        pads_A = (params["nl_pad"] / P.ANGSTROM,)
    table = []
    for pad in pads_A:
        p = dict(params)
        p["nl_pad"] = pad * P.ANGSTROM
        b = G.Batch(p, models, precision=precision, philox_seed=philox_seed,
                    steps_per_launch=trial_steps)
        b.run(trial_steps)                      # first rebuilds, steady state
        block, _ = b.autotune(trial_steps)
        b.run(trial_steps)
        ms, launches = b.last_ms()
        row = {"pad_A": pad, "block": block, "ms_per_step": ms / trial_steps}
        table.append(row)
        if log:
            log(f"tune: pad {pad} A block {block}: {row['ms_per_step']:.4f} ms/step")
        b.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    return {"pad_A": best["pad_A"], "block": best["block"],
            "ms_per_step": best["ms_per_step"], "table": table,
            "pad_fixed": pad_is_physical(params)}
