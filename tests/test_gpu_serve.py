"""Serving loop: SCF jobs of different sizes admitted while others run,
their iterations on their own streams or, with cross-request batching, their
Fock builds in one batched contraction per round.  Each job's energy is the
same in the mix as alone (1e-8 Eh: both runs converge the density to an RMS
change of 1e-8, and the energy error is second order in it), the jobs really
overlap on the device, a budget that holds one job at a time serialises
them, and a job whose footprint exceeds the whole budget is refused."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _jobs(sizes, gap):
    from chronusq_public_amd.chem import water_cluster
    return [("6-31G(d)", water_cluster(n, 12.0 * (n / 16.0) ** (1 / 3), seed=40 + i), gap * i)
            for i, n in enumerate(sizes)]


@pytest.mark.parametrize("batch_fock", [0, 1])
def test_results_alone_equal_results_in_the_mix(batch_fock):
    from chronusq_public_amd import cqx
    jobs = _jobs([3, 5, 2, 4, 6, 3], 0.02)
    say(f"case: serve mix of 6 jobs, 4 slots, batch_fock {batch_fock}")
    mix = cqx.serve(jobs, max_inflight=4, batch_fock=batch_fock)
    assert all(r["status"] == 0 and r["converged"] for r in mix)
    # at least two jobs were on the device at the same time
    overlap = any(a["t_admit"] < b["t_done"] and b["t_admit"] < a["t_done"]
                  for i, a in enumerate(mix) for b in mix[i + 1:])
    assert overlap
    for (bas, mol, _), r in zip(jobs, mix):
        say(f"  case: job alone natom {mol.natom}")
        alone = cqx.serve([(bas, mol, 0.0)], max_inflight=1)[0]
        assert abs(alone["e_total"] - r["e_total"]) < 1e-8
        assert r["t_admit"] >= r["t_arrival"] - 1e-9 and r["t_done"] > r["t_admit"]


@pytest.mark.parametrize("batch_fock", [0, 1])
def test_hbm_budget_serialises_admission(batch_fock):
    from chronusq_public_amd import cqx
    jobs = _jobs([2, 3, 2], 0.0)
    say(f"case: serve with a budget that holds one job, batch_fock {batch_fock}")
    first = cqx.serve(jobs, max_inflight=4, batch_fock=batch_fock)
    foot = [r["footprint_bytes"] for r in first]
    budget = 1.2 * max(foot)
    assert 2 * min(foot) > budget
    res = cqx.serve(jobs, max_inflight=4, hbm_budget_bytes=budget, batch_fock=batch_fock)
    assert all(r["status"] == 0 for r in res)
    spans = sorted((r["t_admit"], r["t_done"]) for r in res)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1 - 1e-6  # one at a time


def test_job_beyond_the_budget_is_refused():
    """A job whose footprint alone exceeds the budget is refused (status
    CQX_ERR_MEM) at arrival instead of running; jobs that fit still run."""
    from chronusq_public_amd import cqx
    jobs = _jobs([2, 3], 0.0)
    say("case: serve refuses a job beyond the budget")
    foot = [r["footprint_bytes"] for r in cqx.serve(jobs, max_inflight=2)]
    small, big = sorted(range(2), key=lambda i: foot[i])
    assert foot[big] > foot[small]
    res = cqx.serve(jobs, max_inflight=2, hbm_budget_bytes=0.5 * (foot[small] + foot[big]),
                    allow_refused=True)
    assert res[small]["status"] == 0 and res[small]["converged"]
    assert res[big]["status"] == 3 and res[big]["iterations"] == 0
    with pytest.raises(cqx.CErr):
        cqx.serve(jobs, max_inflight=2, hbm_budget_bytes=1.0)


def test_recorded_fock_builds_give_the_same_energies():
    """graph=1: every job's later SCF iterations replay the Fock build its
    context recorded; energies and iteration counts as without it."""
    from chronusq_public_amd import cqx
    jobs = _jobs([3, 5, 4], 0.0)
    say("case: serve with recorded Fock builds")
    plain = cqx.serve(jobs, max_inflight=3)
    rec = cqx.serve(jobs, max_inflight=3, graph=1)
    for a, b in zip(plain, rec):
        assert a["status"] == 0 and b["status"] == 0 and a["converged"] and b["converged"]
        assert abs(a["e_total"] - b["e_total"]) < 1e-8
        assert abs(a["iterations"] - b["iterations"]) <= 1


def test_batched_rounds_with_recorded_builds():
    """batch_fock with graph=1: each merged context records its batched build
    and replays it while the set of jobs stays; energies as per-job runs."""
    from chronusq_public_amd import cqx
    jobs = _jobs([3, 5, 4], 0.0)
    say("case: batched serving with recorded Fock builds")
    plain = cqx.serve(jobs, max_inflight=3)
    rec = cqx.serve(jobs, max_inflight=3, graph=1, batch_fock=1)
    for a, b in zip(plain, rec):
        assert a["status"] == 0 and b["status"] == 0 and a["converged"] and b["converged"]
        assert abs(a["e_total"] - b["e_total"]) < 1e-8
