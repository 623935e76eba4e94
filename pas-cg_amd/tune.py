"""Autotuner of the batched form: Verlet skin (NL pad) and workgroup size.

The workgroup size is tuned in place by `pcg_batch_autotune` (trials on the
live state, restored after each). The skin changes the list capacities, so
each candidate pad gets its own batch: after a warm-up long enough to reach
the steady rebuild rate, the block size is tuned and the step timed. The pad
does not change the physics: every list keeps the pairs the force terms can
reach (nl_cutoff + pad bounds them from above), so only the rebuild rate and
the list lengths, hence the cost per step, depend on it.
"""
from __future__ import annotations

from . import pcg as G
from . import params as P


def tune(params: dict, models: list, pads_A=(4.0, 6.0, 8.0, 10.0, 12.0), trial_steps: int = 500,
         precision: int = G.MIXED, philox_seed: int = 1, log=None) -> dict:
    """Best (pad, block) by measured time per step; returns the choice and
    the whole table."""
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
            "ms_per_step": best["ms_per_step"], "table": table}
