"""Backward-Euler transients in the batched engine (the integrator of the
reference's SuperLU build: one implicit step per interval, RCutil.c:268-428
as compute_temp_grid calls it, temperature_grid.c:3632-3662), through the C
ABI.

* every job's sequence in the engine equals, bit for bit, what it gets alone
  in the engine and what the single-model API gives (steady_state_temp, then
  hs_transient_run with HS_TRANS_BE) on the same model: the symmetric jobs
  (FP64 Jacobi CG) and the microchannel jobs (GMRES with the shifted upwind
  multigrid preconditioner) share launches with others of their kind;
* the FP64 engine's intervals are within 1e-5 K of the oracle's restatement
  of the step (advance_be), as the single-model step is;
* jobs submitted while others are between intervals join at the next chunk
  (long-interval microchannel transients served beside conduction jobs).
"""
import tempfile

import numpy as np
import pytest

from hotspot_amd import cases, engine
from helpers import check_error, make, oracle_for, set_options
from test_gpu_engine import lib, transient_run  # noqa: F401  (module fixture)

pytestmark = pytest.mark.gpu

BE_FP64 = (engine.FP64, engine.BE)


class BeJob:
    """A model with its power rows: a steady solve first (from_init False)
    or given start temperatures, then `steps` intervals of `dt`."""

    def __init__(self, lib, name, case, pw, steps, dt, from_init=False, seed=7):
        self.name, self.steps, self.dt, self.from_init = name, steps, dt, from_init
        self.run = make(lib, case)
        self.P = self.run.power_vector(pw)
        self.seq = np.array(cases.jittered_powers(self.P, steps, seed=seed))
        self.init = np.full(self.run.n, 330.0) + cases.rng(seed).uniform(0, 5, self.run.n) if from_init else None

    def rows(self):
        return self.seq if self.from_init else np.vstack([self.P[None, :], self.seq])

    def submit(self, eng):
        return eng.submit(self.run.model, self.rows(), init=self.init, n_steps=self.steps, dt=self.dt)


def be_jobs(lib):
    c32, p32 = cases.stacked_case(32, seed=2, blocks_per_die=16)
    cdet, pdet = cases.stacked_case(16, seed=4, n_dies=2, blocks_per_die=8)
    cdet.lcf[0].flp = [u + (1.5e6 + 1e5 * i, 0.01 * (1 + 0.3 * (i % 3))) for i, u in enumerate(cdet.lcf[0].flp)]
    cdet.detailed_3D = True
    mc16, pmc16 = cases.microchannel_case(16, 5, tempfile.mkdtemp())
    mc32, pmc32 = cases.microchannel_case(32, 6, tempfile.mkdtemp())
    return [BeJob(lib, "default-64", cases.default_stack_case(64), [cases.rng(1).uniform(0.5, 5, 16)], 6, 1e-3),
            BeJob(lib, "stack-32-init", c32, p32, 5, 1e-2, from_init=True),
            BeJob(lib, "detailed-16", cdet, pdet, 4, 1e-3),
            BeJob(lib, "secondary-16", cases.default_stack_case(16, model_secondary=1),
                  [cases.rng(3).uniform(0.5, 5, 16)], 5, 3.333e-6),
            BeJob(lib, "microchannel-16", mc16, pmc16, 4, 1e-2),
            BeJob(lib, "microchannel-32-init", mc32, pmc32, 3, 1e-3, from_init=True),
            BeJob(lib, "default-32", cases.default_stack_case(32), [cases.rng(8).uniform(0.5, 5, 16)], 7, 1e-2)]


@pytest.fixture(scope="module")
def jobs(lib):
    js = be_jobs(lib)
    yield js
    for j in js:
        j.run.free()


def serve(lib, jobs, stagger=0, tick=(32, 128)):
    """Runs the jobs in one FP64 backward-Euler engine; with stagger > 0
    the k-th job is submitted after k x stagger ticks (while the earlier
    ones run).  Returns (sequence, info) per job and the most jobs in their
    transient phase at once."""
    eng = engine.Engine(lib, *BE_FP64)
    eng.set_tick(*tick)
    ids, trans_max = [None] * len(jobs), 0
    for t in range(10 ** 6):
        for k, j in enumerate(jobs):
            if ids[k] is None and t >= k * stagger:
                ids[k] = j.submit(eng)
        active = eng.tick()
        infos = [eng.info(i) for i in ids if i is not None]
        trans_max = max(trans_max, sum(1 for x in infos if x.state == engine.TRANSIENT))
        if active == 0 and all(i is not None for i in ids):
            break
    out = []
    for j, i in zip(jobs, ids):
        assert eng.poll(i) is not None, j.name
        out.append((eng.result_seq(i), eng.info(i)))
        eng.release(i)
    eng.close()
    return out, trans_max


def single_api(lib, j):
    """steady_state_temp (FP64 Jacobi CG, or GMRES for microchannel models),
    then hs_transient_run with the backward-Euler integrator"""
    set_options(lib, j.run, precision=0, steady_solver=0, transient=2)
    if j.from_init:
        T0 = j.init
    else:
        T0 = j.run.steady(j.P.copy())
        check_error(lib)
    s = transient_run(lib, j.run, j.seq, T0, dt=j.dt)
    set_options(lib, j.run, precision=0, steady_solver=0, transient=0)
    return s


def test_be_engine_equals_alone_and_single_api(lib, jobs):
    mix, trans_max = serve(lib, jobs)
    assert trans_max >= 3
    for j, (seq, info) in zip(jobs, mix):
        assert seq.shape == (j.steps, j.run.n)
        assert info.be_iters > 0, j.name
        alone = serve(lib, [j])[0][0][0]
        assert np.array_equal(alone, seq), j.name
        assert np.array_equal(single_api(lib, j), seq), j.name


def test_be_engine_matches_oracle(lib, jobs):
    mix, _ = serve(lib, jobs)
    for j, (seq, _) in zip(jobs, mix):
        om = oracle_for(j.run)
        if j.from_init:
            Tg = om.b2g(j.init, power=False)
        else:
            # the device starts from its own FP64 steady state (1e-12
            # relative residual), the oracle from its direct solve
            Tb, _ = om.steady(j.P)
            Tg = om.b2g(Tb, power=False)
        for k in range(j.steps):
            Tg = om.advance_be(Tg, om.power_grid(j.seq[k]), j.dt)
            assert np.abs(seq[k] - om.g2b(Tg)).max() <= 1e-5, (j.name, k)


def test_be_engine_staggered_arrivals(lib, jobs):
    """Jobs submitted one per tick while the earlier ones are in their
    intervals give the same sequences as all submitted together."""
    together, _ = serve(lib, jobs)
    staggered, trans_max = serve(lib, jobs, stagger=1, tick=(4, 64))
    assert trans_max >= 2
    for j, (a, _), (b, _) in zip(jobs, together, staggered):
        assert np.array_equal(a, b), j.name
