"""Serving loop: SCF jobs of different sizes admitted while others run.
Each job's energy is the same in the mix as alone (1e-8 Eh: both runs
converge the density to an RMS change of 1e-8, and the energy error is
second order in it), the jobs really overlap on the device, and an HBM
budget below every job's footprint serialises them."""
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


def test_results_alone_equal_results_in_the_mix():
    from chronusq_public_amd import cqx
    jobs = _jobs([3, 5, 2, 4, 6, 3], 0.02)
    say("case: serve mix of 6 jobs, 4 slots")
    mix = cqx.serve(jobs, max_inflight=4)
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


def test_hbm_budget_serialises_admission():
    from chronusq_public_amd import cqx
    jobs = _jobs([2, 3, 2], 0.0)
    say("case: serve with a budget below every footprint")
    res = cqx.serve(jobs, max_inflight=4, hbm_budget_bytes=1.0)
    assert all(r["status"] == 0 for r in res)
    spans = sorted((r["t_admit"], r["t_done"]) for r in res)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1 - 1e-6  # one at a time


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
