"""GPU parity of the batched (persistent workgroup-per-system) and served
(continuous batching) forms against the FP64 oracle.

* noise switched off, every system of a ragged batch follows its own oracle
  trajectory (FP64 to 1e-8 of the box, through NL rebuilds and contact
  changes); mixed precision forces of the batch match the oracle's to 1e-5;
* with Philox noise the batch samples the same ensemble as the oracle with
  the reference's ran2 stream (kinetic temperature, radius of gyration,
  native-contact fraction) within statistical error;
* the server gives a job the same result alone and in a mix, retires a job
  that diverges (FAILED) while the others finish, cancels, and refuses a job
  larger than its HBM budget.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle
from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G

pytestmark = pytest.mark.gpu


def prm_full(**kw):
    p = P.default_params()
    p.update(chir_on=1, afm_on=1)
    p.update(kw)
    return p


def ragged_models(sizes, seed0=100, pulled=(1,)):
    out = []
    for k, n in enumerate(sizes):
        m = M.synthetic_protein(n, M.job_seed(seed0 + k))
        rng = np.random.default_rng(seed0 + k)
        m["pos"] = m["pos"] + rng.normal(scale=0.3 * P.ANGSTROM, size=m["pos"].shape)
        m["atype"][::6] = P.AA_INDEX["LYS"]
        m["atype"][2::9] = P.AA_INDEX["GLU"]
        if k in pulled:
            m = M.add_pulling(m)
        out.append(m)
    return out


@pytest.mark.parametrize("qa", [0, 1], ids=["nat_cont", "qa"])
def test_batch_noise_off_follows_oracle(qa):
    prm = prm_full(qa_on=qa, nat_cont_on=1 - qa)
    models = ragged_models([33, 120, 57, 200, 41])
    steps = 150
    # seed 1 everywhere: the initial velocities are drawn from ran2(seed)
    b = G.Batch(prm, models, seeds=[1] * len(models), precision=G.FP64, steps_per_launch=40)
    b.set_noise_scale(0.0)
    b.run(steps)
    for s, m in enumerate(models):
        o = Oracle(prm, m, seed=1)
        o.set_noise_scale(0.0)
        tr = o.step(steps)
        r = b.get(s, 0)
        err = np.abs(r - o.r()).max()
        assert err < 1e-8, (s, m["n"], err)
        V, _ = b.energy(s)
        assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (s, V, tr[-1])


def test_batch_mixed_forces_match_oracle():
    prm = prm_full(qa_on=0)
    models = ragged_models([64, 150, 47], seed0=7)
    b = G.Batch(prm, models, seeds=[1] * len(models), precision=G.MIXED)
    b.set_noise_scale(0.0)
    b.run(1)
    for s, m in enumerate(models):
        o = Oracle(prm, m, seed=1)
        o.fix_nl()
        o.eval_forces()
        F1 = o.F()
        F2 = b.get(s, 6)
        err = np.abs(F2 - F1).max() / np.abs(F1).max()
        assert err <= 1e-5, (s, err)
        V1 = o.V
        V2, _ = b.energy(s)
        assert abs(V2 - V1) <= 1e-6 * abs(V1), (s, V1, V2)


def test_batch_list_sizes_match_oracle():
    prm = prm_full(qa_on=0, pid_on=0)
    models = ragged_models([80, 35], seed0=3)
    b = G.Batch(prm, models, seeds=[1] * len(models), precision=G.FP64)
    b.set_noise_scale(0.0)
    b.run(1)
    for s, m in enumerate(models):
        o = Oracle(prm, m, seed=1)
        o.fix_nl()
        sz = b.sizes(s)
        assert sz["n"] == m["n"]
        assert sz["nl"] == len(o.list(0)[0])
        assert sz["pauli"] == len(o.list(2)[0])
        assert sz["const_dh"] == sz["rel_dh"] == len(o.list(3)[0])
        assert sz["tether"] == len(o.static_list(0)[0])
        assert sz["lrep"] == len(o.static_list(1)[0])


def _observables(r, m, nat_d):
    ii, jj = m["cont"][:, 0], m["cont"][:, 1]
    d = np.linalg.norm(r[ii] - r[jj], axis=1)
    q = float(np.mean(d < 1.2 * nat_d))
    rg = float(np.sqrt(((r - r.mean(0)) ** 2).sum(1).mean()))
    return rg, q


def test_philox_ensemble_matches_oracle_ran2():
    """Statistical parity: the GPU's Philox noise and the reference's ran2
    stream give the same kinetic temperature, radius of gyration and
    native-contact fraction (mean over copies and time, after 500 steps)."""
    prm = prm_full(qa_on=0, afm_on=0, temperature=P.f32(0.35))
    m = M.synthetic_protein(48, M.job_seed(5))
    nat_d = m["cont_len"]
    copies, chunks, chunk = 24, 12, 250
    b = G.Batch(prm, [m] * copies, seeds=list(range(copies)), precision=G.MIXED,
                philox_seed=2024, steps_per_launch=chunk)
    stats_g = []
    b.run(500)
    for _ in range(chunks):
        b.run(chunk)
        for s in range(copies):
            v = b.get(s, 7)
            rg, q = _observables(b.get(s, 0), m, nat_d)
            stats_g.append(((v * v).sum(1).mean() / 3.0, rg, q))
    stats_o = []
    for s in range(8):
        o = Oracle(prm, m, seed=1000 + s)
        o.step(500)
        for _ in range(chunks):
            o.step(chunk)
            v = o.y(1) / P.f32(5e-3)
            rg, q = _observables(o.r(), m, nat_d)
            stats_o.append(((v * v).sum(1).mean() / 3.0, rg, q))
    g = np.asarray(stats_g)
    o = np.asarray(stats_o)
    for k, name in enumerate(["T_kin", "Rg", "Q"]):
        mg, mo = g[:, k].mean(), o[:, k].mean()
        se = np.sqrt(g[:, k].var() / len(g) + o[:, k].var() / len(o))
        assert abs(mg - mo) < 5 * se + 0.02 * abs(mo), (name, mg, mo, se)
    assert abs(g[:, 0].mean() - 0.35) < 0.035


def _run_all(srv, cap=200):
    for _ in range(cap):
        if srv.round() == 0:
            return
    raise AssertionError("server did not drain")


def test_server_job_alone_equals_job_in_mix():
    prm = prm_full(qa_on=0)
    models = ragged_models([90, 40, 160, 70], seed0=40)
    one = G.Server(prm, precision=G.FP64, chunk_steps=64)
    j1 = one.submit(models[0], 300, seed=11)
    _run_all(one)
    mix = G.Server(prm, precision=G.FP64, chunk_steps=64, max_resident=2)
    ids = [mix.submit(m, 120 + 90 * k, seed=20 + k) for k, m in enumerate(models[1:])]
    j2 = mix.submit(models[0], 300, seed=11)
    _run_all(mix)
    assert one.status(j1) == ("done", 300)
    assert mix.status(j2) == ("done", 300)
    for k, j in enumerate(ids):
        assert mix.status(j) == ("done", 120 + 90 * k)
    r1, V1 = one.result(j1)
    r2, V2 = mix.result(j2)
    assert np.abs(r1 - r2).max() < 1e-8
    assert abs(V1 - V2) <= 1e-9 * abs(V1)
    t_sub, t_start, t_end = mix.times(j2)
    assert t_sub <= t_start <= t_end


def test_server_isolates_failed_job_and_cancels():
    prm = prm_full(qa_on=0)
    good = ragged_models([60, 90], seed0=60)
    bad = M.synthetic_protein(50, M.job_seed(61))
    bad["pos"] = bad["pos"].copy()
    bad["pos"][10] = bad["pos"][11]  # coincident bonded residues: NaN force
    srv = G.Server(prm, precision=G.MIXED, chunk_steps=50)
    g1 = srv.submit(good[0], 200, seed=1)
    jb = srv.submit(bad, 200, seed=2)
    g2 = srv.submit(good[1], 200, seed=3)
    jc = srv.submit(good[0], 200, seed=4)
    srv.cancel(jc)
    _run_all(srv)
    assert srv.status(jb)[0] == "failed"
    assert srv.status(jc)[0] == "cancelled"
    for j in (g1, g2):
        st, done = srv.status(j)
        assert (st, done) == ("done", 200)
        r, V = srv.result(j)
        assert np.isfinite(r).all() and np.isfinite(V)
    assert srv.stats()["resident"] == 0


def test_server_refuses_job_over_budget():
    prm = prm_full(qa_on=0)
    srv = G.Server(prm, precision=G.MIXED, hbm_budget_bytes=1 << 16)
    with pytest.raises(G.PcgError):
        srv.submit(M.synthetic_protein(300, M.job_seed(1)), 10)


def test_autotune_keeps_the_trajectory():
    """The autotuner's trials run on the live state and restore it: a tuned
    batch follows an untuned one (noise off, FP64; the workgroup size only
    changes the force summation order)."""
    prm = prm_full(qa_on=1, nat_cont_on=0)
    models = ragged_models([70, 150, 45], seed0=80)
    a = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=50)
    b = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=50)
    for x in (a, b):
        x.set_noise_scale(0.0)
        x.run(50)
    block, ms = b.autotune(60)
    assert block in (128, 256, 512, 1024) and ms > 0
    assert b.info()["block"] == block
    a.run(100)
    b.run(100)
    for s in range(3):
        assert np.abs(a.get(s, 0) - b.get(s, 0)).max() < 1e-9


def test_batch_mixed_forces_match_fp64_with_energies_off():
    """With energies off no step of the mixed batched kernel takes the FP64
    energy path, so its FP32 angles, dihedrals, contacts and local terms all
    act: the total force stays within 1e-5 of the same kernel in FP64
    (relative to the largest component) on evolved positions. In FP64 the
    switch leaves the trajectory unchanged."""
    prm = prm_full(qa_on=1)
    models = ragged_models([40, 150, 300], seed0=200)
    b0 = G.Batch(prm, models, precision=G.MIXED, steps_per_launch=200)
    b0.run(400)
    pos = [b0.get(s, 0) for s in range(3)]
    b0.close()
    F = {}
    for prec in (G.FP64, G.MIXED):
        b = G.Batch(prm, models, precision=prec, steps_per_launch=1)
        b.energies(False)
        b.set_noise_scale(0.0)
        for k, r in enumerate(pos):
            b.set(k, 0, r)
        b.run(1)
        F[prec] = [b.get(k, 6) for k in range(3)]
        b.close()
    for s, (a, c) in enumerate(zip(F[G.MIXED], F[G.FP64])):
        err = float(np.abs(a - c).max() / np.abs(c).max())
        assert err <= 1e-5, (s, err)
    runs = []
    for on in (True, False):
        b = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=25)
        b.set_noise_scale(0.0)
        b.energies(on)
        b.run(50)
        runs.append([b.get(s, 0) for s in range(3)])
        b.close()
    for s in range(3):
        assert np.abs(runs[0][s] - runs[1][s]).max() < 1e-9


def test_batch_checkpoint_resumes_the_trajectory():
    """A checkpoint restores a whole batch: resumed in place, or in a new batch
    of the same models, a run continues the trajectory of the uninterrupted
    batch (mixed precision, Philox noise on: the noise position is part of
    the state). A batch of other models refuses it."""
    prm = prm_full(qa_on=1)
    models = ragged_models([45, 130, 70], seed0=300)
    b = G.Batch(prm, models, precision=G.MIXED, philox_seed=9, steps_per_launch=100)
    b.run(200)
    ck = b.checkpoint()
    b.run(300)
    ref = [b.get(s, 0) for s in range(3)]
    b.restore(ck)
    b.run(300)
    again = [b.get(s, 0) for s in range(3)]
    b2 = G.Batch(prm, models, precision=G.MIXED, philox_seed=9, steps_per_launch=100)
    b2.restore(ck)
    b2.run(300)
    fresh = [b2.get(s, 0) for s in range(3)]
    for s in range(3):
        scale = np.abs(ref[s]).max()
        # the same trajectory up to the order of the LDS force additions
        assert np.abs(again[s] - ref[s]).max() <= 1e-7 * scale, s
        assert np.abs(fresh[s] - ref[s]).max() <= 1e-7 * scale, s
        assert abs(b2.report()[s]["TIME"] - b.report()[s]["TIME"]) == 0.0
    b3 = G.Batch(prm, models[:2], precision=G.MIXED, philox_seed=9, steps_per_launch=100)
    with pytest.raises(G.PcgError):
        b3.restore(ck)
    for x in (b, b2, b3):
        x.close()


def test_batch_graph_replays_match_direct_launches():
    """The hipGraph of one launch, replayed for each full launch of a run
    (a shorter last launch goes directly), with host reads in between,
    follows the directly launched batch; runs of 50 and 30 steps end in a
    partial launch."""
    prm = prm_full(qa_on=1, nat_cont_on=1)
    models = ragged_models([64, 130, 38], seed0=90)
    a = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=20)
    g = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=20,
                use_graph=True)
    for x in (a, g):
        x.set_noise_scale(0.0)
    for rep, n in enumerate((60, 50, 60, 30)):
        a.run(n)
        g.run(n)
        for s in range(3):
            err = np.abs(a.get(s, 0) - g.get(s, 0)).max()
            assert err < 1e-9, (rep, s, err)
            assert abs(a.energy(s)[0] - g.energy(s)[0]) <= 1e-9 * abs(a.energy(s)[0])


def test_batch_bonded_paths_follow_oracle():
    """The three ways the batched kernel evaluates the static bonded terms
    give the oracle's trajectory (noise off, FP64): chain-local lists staged
    in LDS (small protein), chain-local lists read from HBM (a protein too
    large to stage in FP64), and per-term loops for a system with a bonded
    item that is not chain-local (an extra tether)."""
    prm = prm_full(qa_on=0)
    small = ragged_models([70], seed0=120, pulled=())[0]
    large = ragged_models([1700], seed0=121, pulled=())[0]
    odd = ragged_models([60], seed0=122, pulled=())[0]
    odd["teth"] = np.concatenate([odd["teth"], np.array([[5, 40]], np.int32)])
    for m, steps in ((small, 60), (large, 20), (odd, 60)):
        b = G.Batch(prm, [m], seeds=[1], precision=G.FP64, steps_per_launch=steps)
        b.set_noise_scale(0.0)
        b.run(steps)
        o = Oracle(prm, m, seed=1)
        o.set_noise_scale(0.0)
        tr = o.step(steps)
        err = np.abs(b.get(0, 0) - o.r()).max()
        assert err < 1e-8, (m["n"], err)
        V, _ = b.energy(0)
        assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (m["n"], V, tr[-1])


def test_batch_tiny_and_empty_systems_follow_oracle():
    """Edge cases in one batch: systems of 1 to 5 residues (no pairs, no
    angles, no dihedrals, chain ends everywhere) and a protein without native
    contacts, beside an ordinary one, follow the oracle (noise off, FP64).
    One residue has NaN velocities from the start, in the reference as here:
    removing the mean velocity leaves zero, which is then scaled by
    sqrt(.../0) (cg/src/cg/simul/state.cpp:316-330)."""
    prm = prm_full(qa_on=1)
    models = [M.synthetic_protein(n, M.job_seed(500 + n)) for n in (1, 2, 3, 4, 5)]
    bare = M.synthetic_protein(40, M.job_seed(600))
    for key in ("cont", "cont_len", "cont_type"):
        bare[key] = bare[key][:0]
    models += [bare] + ragged_models([60], seed0=601, pulled=())
    steps = 80
    b = G.Batch(prm, models, seeds=[1] * len(models), precision=G.FP64, steps_per_launch=20)
    b.set_noise_scale(0.0)
    b.run(steps)
    for s, m in enumerate(models):
        o = Oracle(prm, m, seed=1)
        o.set_noise_scale(0.0)
        tr = o.step(steps)
        if m["n"] == 1:
            assert np.isnan(b.get(s, 0)).all() and np.isnan(o.r()).all()
            continue
        err = np.abs(b.get(s, 0) - o.r()).max()
        assert err < 1e-8, (s, m["n"], err)
        V, _ = b.energy(s)
        assert abs(V - tr[-1]) <= 1e-9 * max(abs(tr[-1]), 1e-12), (s, m["n"], V, tr[-1])


def test_batch_beyond_lds_takes_stage_kernels():
    """A system too large for the persistent workgroup's LDS (2800 residues)
    runs on the stage kernels, in a batch with a small one, and follows the
    oracle (noise off, FP64); the served form refuses such a job."""
    prm = prm_full(qa_on=1)
    models = ragged_models([2800, 50], seed0=700, pulled=())
    steps = 12
    b = G.Batch(prm, models, seeds=[1, 1], precision=G.FP64, steps_per_launch=6)
    b.set_noise_scale(0.0)
    b.run(steps)
    for s, m in enumerate(models):
        o = Oracle(prm, m, seed=1)
        o.set_noise_scale(0.0)
        tr = o.step(steps)
        err = np.abs(b.get(s, 0) - o.r()).max()
        assert err < 1e-8, (s, m["n"], err)
        V, _ = b.energy(s)
        assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (s, m["n"], V, tr[-1])
    b.close()
    srv = G.Server(prm, precision=G.MIXED)
    with pytest.raises(G.PcgError):
        srv.submit(models[0], 10, temperature=0.35, seed=1)
    srv.close()


def test_batch_set_positions_then_forces_match_oracle():
    """pcg_batch_set moves a system (and makes it rebuild its lists): the
    next step's forces match the oracle's at the new positions."""
    prm = prm_full(qa_on=0)
    models = ragged_models([90, 44], seed0=140)
    b = G.Batch(prm, models, seeds=[1, 1], precision=G.MIXED, steps_per_launch=1)
    b.set_noise_scale(0.0)
    b.run(3)
    rng = np.random.default_rng(3)
    for s, m in enumerate(models):
        r = m["pos"] + rng.normal(scale=0.4 * P.ANGSTROM, size=m["pos"].shape)
        b.set(s, 0, r)
    b.run(1)
    # the oracle evaluated at the same positions (and time, for the AFM tip)
    rng = np.random.default_rng(3)
    for s, m in enumerate(models):
        r = m["pos"] + rng.normal(scale=0.4 * P.ANGSTROM, size=m["pos"].shape)
        o = Oracle(prm, m, seed=1)
        o.set_r(r)
        o.set_t(3 * P.f32(5e-3))
        o.fix_nl()
        o.eval_forces()
        F1 = o.F()
        F2 = b.get(s, 6)
        err = np.abs(F2 - F1).max() / np.abs(F1).max()
        assert err <= 1e-5, (s, err)


def test_server_force_tip_job_alone_equals_job_in_mix():
    """A job pulled by a constant-force tip (the served form copies the force
    tips with the job's model) gives the same result alone and among jobs
    pulled by velocity tips or not pulled at all."""
    prm = prm_full(qa_on=0)
    models = ragged_models([70, 50, 110], seed0=60)
    pulled = M.add_force_pulling(dict(models[0]), force_eps_per_A=2.0)
    models[1] = M.add_pulling(models[1])
    one = G.Server(prm, precision=G.FP64, chunk_steps=50)
    j1 = one.submit(pulled, 250, seed=5)
    _run_all(one)
    mix = G.Server(prm, precision=G.FP64, chunk_steps=50, max_resident=2)
    ids = [mix.submit(m, 100 + 80 * k, seed=30 + k) for k, m in enumerate(models[1:])]
    j2 = mix.submit(pulled, 250, seed=5)
    _run_all(mix)
    assert one.status(j1) == ("done", 250) and mix.status(j2) == ("done", 250)
    r1, V1 = one.result(j1)
    r2, V2 = mix.result(j2)
    assert np.abs(r1 - r2).max() < 1e-8
    assert abs(V1 - V2) <= 1e-9 * abs(V1)
    # the force tip acts: the same job without it ends elsewhere
    free = M.add_force_pulling(dict(models[0]), force_eps_per_A=0.0)
    ref = G.Server(prm, precision=G.FP64, chunk_steps=50)
    j3 = ref.submit(free, 250, seed=5)
    _run_all(ref)
    r3, _ = ref.result(j3)
    assert np.abs(r3 - r1).max() > 1e-6


def test_batch_checkpoint_resume_follows_oracle():
    """A batch checkpointed mid-run and resumed in a new batch continues the
    oracle's noise-free FP64 trajectory (QA contacts, sync budgets, lists and
    native-contact state all come back with it); a batch with other
    parameters or other models of the same sizes refuses the checkpoint."""
    prm = prm_full(qa_on=1)
    models = ragged_models([45, 130, 70], seed0=310)
    b = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=50)
    b.set_noise_scale(0.0)
    b.run(100)
    ck = b.checkpoint()
    b.close()
    b2 = G.Batch(prm, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=50)
    b2.restore(ck)
    b2.run(100)
    for s, m in enumerate(models):
        o = Oracle(prm, m, seed=1)
        o.set_noise_scale(0.0)
        tr = o.step(200)
        assert np.abs(b2.get(s, 0) - o.r()).max() < 1e-8, s
        V, _ = b2.energy(s)
        assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (s, V, tr[-1])
    b2.close()
    other_p = prm_full(qa_on=1, temperature=0.4)
    b3 = G.Batch(other_p, models, seeds=[1] * 3, precision=G.FP64, steps_per_launch=50)
    with pytest.raises(G.PcgError):
        b3.restore(ck)
    b3.close()
    other_m = ragged_models([45, 130, 70], seed0=410)
    b4 = G.Batch(prm, other_m, seeds=[1] * 3, precision=G.FP64, steps_per_launch=50)
    with pytest.raises(G.PcgError):
        b4.restore(ck)
    b4.close()


def test_context_checkpoint_resume_follows_oracle():
    """The plain context with the reference's ran2 noise: checkpointed after
    60 steps and resumed in a new context, it continues the oracle's noisy
    FP64 trajectory step for step (the ran2 stream is part of the
    checkpoint); the same context without the restore would not."""
    prm = prm_full(qa_on=1, pid_on=1)
    m = ragged_models([80], seed0=330)[0]
    c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=7)
    c.step(60)
    ck = c.checkpoint()
    c.close()
    c2 = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=7)
    c2.restore(ck)
    o = Oracle(prm, m, seed=7)
    o.step(60)
    tr = o.step(60)
    for k in range(60):
        c2.step(1)
        assert abs(c2.energy()[0] - tr[k]) <= 1e-9 * abs(tr[k]), (k, c2.energy()[0], tr[k])
    assert np.abs(c2.r() - o.r()).max() < 1e-9
    assert c2.time()[1] == 120
    c3 = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=7)
    with pytest.raises(G.PcgError):
        c3.restore(ck)  # a checkpoint of the ran2 noise source
    c3.close()
    c2.close()
