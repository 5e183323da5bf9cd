"""Continuous batching in the engine (the served form), through the C ABI.

Requests are submitted between ticks while other jobs are on the device:
steady solves and transient sequences (after a steady solve, or from given
temperatures) on conduction stacks from 64 x 64 to 256 x 256 and one
microchannel stack, with a device-memory budget that holds some of them in
the queue until earlier jobs finish.  Each job's result must equal, bit for
bit, its result when it is served alone (the north star's "same alone and in
the mix"); the FP64 form's results must be within 1e-4 K of the oracle
(steady: the converged reference operator; transient: the reference's
adaptive RK4 restated, temperature_grid.c:3603-3704 and RCutil.c:544-640,
from the same start state), and the mixed form's within 1e-2 K of the FP64
form's.  A steady job must be reported while transient jobs still run.
"""
import tempfile

import numpy as np
import pytest

from hotspot_amd import cases, engine
from helpers import make, oracle_for

pytestmark = pytest.mark.gpu

DT = 3.333e-6
FORMS = [(engine.FP64, engine.RK4), (engine.MIXED, engine.TAYLOR)]
FORM_IDS = ["fp64-rk4", "mixed-taylor"]


def serving_specs():
    """(name, case, block powers per power layer, intervals, from given
    temperatures?, arrival tick)"""
    c64, p64 = cases.stacked_case(64, seed=31, n_dies=2, blocks_per_die=16)
    c128, p128 = cases.chip_stack_case(128, seed=32, chip_layers=3, n_blocks=32, power=("density", 0.5),
                                       hotspots=4)
    c256, p256 = cases.chip_stack_case(256, seed=33, chip_layers=1, n_blocks=64, power=("density", 0.5),
                                       hotspots=4)
    cmc, pmc = cases.microchannel_case(32, 35, tempfile.mkdtemp(prefix="hsserve"))
    cs, ps = cases.stacked_case(128, seed=36, n_dies=2, blocks_per_die=16)
    return [
        ("stack-64 steady+200", c64, p64, 200, False, 0),
        ("default-128 steady", cases.default_stack_case(128), [cases.rng(1).uniform(0.5, 5, 16)], 0, False, 0),
        ("chip-256 steady+6", c256, p256, 6, False, 1),
        ("microchannel-32 steady+4", cmc, pmc, 4, False, 2),
        ("default-64 init+10", cases.default_stack_case(64), [cases.rng(2).uniform(0.5, 5, 16)], 10, True, 3),
        ("chip-128 steady", c128, p128, 0, False, 4),
        ("stack-128 steady+5", cs, ps, 5, False, 6),
        ("default-64 steady", cases.default_stack_case(64, 2, 2), [cases.rng(3).uniform(1, 3, 4)], 0, False, 8),
    ]


# later jobs on the models of earlier ones (the engine keeps a model's device
# state for its next job): (name, index of the job whose model it uses,
# power scale, intervals, arrival tick)
SHARED = [("default-128 steady again", 1, 1.3, 0, 30),
          ("microchannel-32 again steady+3", 3, 0.8, 3, 40),
          ("chip-128 again steady+4", 5, 1.1, 4, 60)]


class Job:
    def __init__(self, lib, spec, share=None):
        if share is None:
            self.name, case, pw, self.steps, self.from_init, self.arrival = spec
            self.run = make(lib, case)
            self.P = self.run.power_vector(pw)
        else:
            self.name, _, scale, self.steps, self.arrival = spec
            self.from_init = False
            self.run = share.run
            self.P = share.P * scale
        self.owner = share is None
        self.seq = np.array(cases.jittered_powers(self.P, self.steps, seed=11)) if self.steps else None
        self.init = (np.full(self.run.n, 325.0) + cases.rng(12).uniform(0, 10, self.run.n)) if self.from_init else None

    def rows(self):
        if not self.steps:
            return self.P[None, :]
        return self.seq if self.from_init else np.vstack([self.P[None, :], self.seq])

    def submit(self, eng):
        return eng.submit(self.run.model, self.rows(), init=self.init, n_steps=self.steps, dt=DT)


@pytest.fixture(scope="module")
def jobs(prod, gpu):
    js = [Job(prod, s) for s in serving_specs()]
    js += [Job(prod, s, share=js[s[1]]) for s in SHARED]
    yield js
    for j in js:
        if j.owner:
            j.run.free()


def footprints(lib, jobs, form):
    eng = engine.Engine(lib, *form)
    ids = [j.submit(eng) for j in jobs]
    need = [eng.info(i).device_bytes for i in ids]
    eng.close()
    return need


def serve_stream(lib, jobs, form, budget):
    """Submits each job at its arrival tick while the engine runs; returns
    the results, the events seen and the engine infos."""
    eng = engine.Engine(lib, *form)
    eng.set_budget(budget)
    eng.set_tick(8, 128)
    ids = [None] * len(jobs)
    seen = dict(submitted_while_resident=0, queued_then_admitted=set(), steady_done_while_transient=0)
    was_queued, steady_seen = set(), set()
    tick = 0
    while True:
        for k, j in enumerate(jobs):
            if ids[k] is None and j.arrival <= tick:
                resident = [i for i in ids if i is not None and
                            eng.info(i).state in (engine.STEADY, engine.TRANSIENT)]
                if resident:
                    seen["submitted_while_resident"] += 1
                ids[k] = j.submit(eng)
        active = eng.tick()
        tick += 1
        infos = {k: eng.info(i) for k, i in enumerate(ids) if i is not None}
        for k, x in infos.items():
            if x.state == engine.QUEUED:
                was_queued.add(k)
            elif k in was_queued:
                seen["queued_then_admitted"].add(k)
        running_trans = any(x.state == engine.TRANSIENT for x in infos.values())
        for k, x in infos.items():
            if x.state == engine.DONE and jobs[k].steps == 0 and k not in steady_seen:
                steady_seen.add(k)
                seen["steady_done_while_transient"] += int(running_trans)
        if all(i is not None for i in ids) and active == 0:
            break
        assert tick < 10 ** 6
    out = []
    for j, i in zip(jobs, ids):
        r = eng.poll(i)
        assert r is not None, j.name
        seq = eng.result_seq(i) if j.steps else None
        out.append((r, seq, eng.info(i)))
        eng.release(i)
    eng.close()
    return out, seen


def alone(lib, job, form):
    eng = engine.Engine(lib, *form)
    i = job.submit(eng)
    eng.run()
    r = eng.poll(i)
    seq = eng.result_seq(i) if job.steps else None
    eng.close()
    return r, seq


_results = {}


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_streamed_jobs_equal_alone(prod, jobs, form):
    need = footprints(prod, jobs, form)
    # room for the largest job and about half of the rest: later arrivals
    # wait in the queue until earlier jobs finish
    budget = max(1.05 * max(need), 0.5 * sum(need))
    res, seen = serve_stream(prod, jobs, form, budget)
    assert seen["submitted_while_resident"] >= 4, seen
    assert seen["queued_then_admitted"], "no job waited for memory and was admitted later"
    assert seen["steady_done_while_transient"] >= 1, "steady results waited for the transient jobs"
    for j, (r, seq, info) in zip(jobs, res):
        ra, sa = alone(prod, j, form)
        assert np.array_equal(r, ra), j.name
        if j.steps:
            assert np.array_equal(seq, sa), j.name
            assert np.array_equal(seq[-1], r), j.name
        assert info.latency_s >= info.wait_s >= 0
        assert info.device_bytes <= budget
    _results[form] = res


@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_resident_model_equals_fresh(prod, jobs, form):
    """A job admitted on the device state its model kept from an earlier job
    of the engine (no rebuild of coefficients, maps or multigrid levels)
    gives the result of a fresh engine, bit for bit."""
    first, again = jobs[1], next(j for j in jobs if j.name.startswith("chip-128 again"))
    share = next(j for j in jobs if j.name == "chip-128 steady")
    eng = engine.Engine(prod, *form)
    eng.set_tick(8, 128)
    a = first.submit(eng)
    b = share.submit(eng)
    eng.run()
    c = again.submit(eng)   # same model as b, which has finished
    eng.run()
    info = eng.info(c)
    r, seq = eng.poll(c), eng.result_seq(c)
    assert not eng.info(a).resident_model and not eng.info(b).resident_model
    eng.close()
    assert info.resident_model == 1
    ra, sa = alone(prod, again, form)
    assert np.array_equal(r, ra) and np.array_equal(seq, sa)


def test_streamed_fp64_matches_oracle(jobs):
    if FORMS[0] not in _results:
        pytest.skip("the FP64 stream did not run")
    for j, (r, seq, info) in zip(jobs, _results[FORMS[0]]):
        om = oracle_for(j.run)
        if not j.steps:
            Tb, _ = om.steady(j.P)
            assert np.abs(r - Tb).max() <= 1e-4, j.name
            continue
        if j.from_init:
            Tg = om.b2g(j.init, power=False)
        else:
            # the device starts from its own FP64 steady state (1e-12
            # relative residual), the oracle from the converged operator
            Tb, _ = om.steady(j.P)
            Tg = om.b2g(Tb, power=False)
        worst = 0.0
        for k in range(j.steps):
            Tg, _ = om.advance(Tg, om.power_grid(j.seq[k]), DT)
            worst = max(worst, float(np.abs(seq[k] - om.g2b(Tg)).max()))
        assert worst <= 1e-4, (j.name, worst)


def test_streamed_mixed_within_tolerance_of_fp64(jobs):
    if any(f not in _results for f in FORMS):
        pytest.skip("both streams must have run")
    for j, (a, sa, _), (b, sb, _) in zip(jobs, _results[FORMS[0]], _results[FORMS[1]]):
        assert np.abs(a - b).max() <= 1e-2, j.name
        if j.steps:
            assert np.abs(sa - sb).max() <= 1e-2, j.name
