"""Recorded builds (option graph): a build records its launches into a HIP
graph and later builds with the same contraction signature replay it, with
every list and ERI kernel reading its count from the device.  Each replay
must equal an eager build of the same densities (1e-11), also with eager
work on other contexts between the replays, for real and complex,
single-molecule and batched mixed-precision contexts; a density with more
candidates than the recording build's capacities falls back to an eager
build and records again."""
import sys

import numpy as np
import pytest

from oracle.oracle import COULOMB, EXCHANGE

pytestmark = pytest.mark.gpu

TOL = 1e-11


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _run(gto, Xs, types, hers, screen=True):
    from chronusq_public_amd import cqx
    cplx = any(np.iscomplexobj(x) for x in Xs)
    dt = np.complex128 if cplx else np.float64
    AX = [np.zeros(x.shape, dtype=dt, order="F") for x in Xs]
    gto.twoBodyContract([cqx.TwoBodyContraction(np.asarray(x, dtype=dt), a, h, t)
                         for x, a, t, h in zip(Xs, AX, types, hers)], screen=screen)
    return AX


def _density(rng, nb, nocc, cplx=False, scale=1.0):
    C_ = rng.normal(0, 1 / np.sqrt(nb), size=(nb, nocc))
    if cplx:
        C_ = C_ + 1j * rng.normal(0, 1 / np.sqrt(nb), size=(nb, nocc))
    return np.asfortranarray(scale * (C_ @ C_.conj().T))


def _pair(basis, opts=()):
    """The same basis under a recording context and an eager one; the eager
    one builds quartet lists only (a recorded build never takes tiles), so
    that both count the same candidates."""
    from chronusq_public_amd import cqx
    out = []
    for graph in (1, 0):
        tpi = cqx.DirectTPI(basis, 1e-12)
        for k, v in opts:
            tpi.set_option(k, v)
        tpi.set_option("graph", graph)
        tpi.set_option("tiles", 0)
        out.append((tpi, cqx.GTODirectTPIContraction(tpi)))
    return out


def _maxdiff(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b))


@pytest.mark.parametrize("cplx", [False, True])
def test_replays_equal_eager_builds(cplx):
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    say("case: graph replays, (H2O)4 cc-pVDZ, cplx", cplx)
    b = cqx.BasisSet("cc-pVDZ", water_cluster(4, 8.0, seed=0))
    nb = b.nBasis
    (tg, gg), (te, ge) = _pair(b)
    if cplx:  # the GHF list: J on S, K on S, Z, Y, X
        types, hers = [COULOMB] + [EXCHANGE] * 4, [True] * 5
    else:
        types, hers = [COULOMB, EXCHANGE], [True, True]
    rng = np.random.Generator(np.random.PCG64(11))
    ob = cqx.BasisSet("6-31G(d)", water_cluster(2, 4.0, seed=3))
    other = cqx.GTODirectTPIContraction(cqx.DirectTPI(ob, 1e-12))
    modes = []
    for it in range(5):
        Xs = [_density(rng, nb, 20, cplx) for _ in types]
        got = _run(gg, Xs, types, hers)
        modes.append(tg.last_stats()["graph"])
        nq_g = tg.last_stats()["n_quartets"]
        ref = _run(ge, Xs, types, hers)
        assert te.last_stats()["graph"] == 0
        err = _maxdiff(got, ref)
        say(f"  build {it}: graph mode {modes[-1]} max|d| {err:.2e}")
        assert err < TOL
        assert nq_g == te.last_stats()["n_quartets"]
        # eager work on another context between the replays
        D2 = _density(rng, ob.nBasis, 5)
        _run(other, [D2, D2], [COULOMB, EXCHANGE], [True, True])
    assert modes[0] == 1 and all(m == 2 for m in modes[1:])


def test_batched_mixed_replays_equal_eager_builds():
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    say("case: graph replays, batched mixed 4 molecules")
    parts = [cqx.BasisSet(bas, mol) for mol, bas in
             [(water_cluster(2, 4.0, seed=5), "6-31G(d)"), (water_cluster(3, 6.0, seed=7), "cc-pVDZ"),
              (water_cluster(1, 3.0, seed=1), "cc-pVTZ"), (water_cluster(5, 9.0, seed=0), "cc-pVDZ")]]
    bb = cqx.BasisSet.batch(parts)
    (tg, gg), (te, ge) = _pair(bb, [("fp32_cut", 1e-4)])
    rng = np.random.Generator(np.random.PCG64(12))
    types, hers = [COULOMB, EXCHANGE] * len(parts), [True, True] * len(parts)
    for it in range(4):
        dens = [_density(rng, p.nBasis, 5) for p in parts]
        Xs = [D for D in dens for _ in (0, 1)]
        got = _run(gg, Xs, types, hers)
        st = tg.last_stats()
        ref = _run(ge, Xs, types, hers)
        err = _maxdiff(got, ref)
        say(f"  build {it}: graph mode {st['graph']} fp32 {st['n_quartets_f32']} max|d| {err:.2e}")
        assert st["graph"] == (1 if it == 0 else 2)
        assert st["n_quartets"] == te.last_stats()["n_quartets"]
        assert st["n_quartets_f32"] == te.last_stats()["n_quartets_f32"] > 0
        assert err < TOL


def test_more_candidates_than_recorded_falls_back_and_records_again():
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    say("case: graph capacity exceeded")
    b = cqx.BasisSet("cc-pVDZ", water_cluster(6, 10.0, seed=2))
    nb = b.nBasis
    (tg, gg), (te, ge) = _pair(b, [("graph_margin", 1.0)])
    types, hers = [COULOMB, EXCHANGE], [True, True]
    rng = np.random.Generator(np.random.PCG64(13))
    # a small density screens out many candidates; a large one keeps them
    small = _density(rng, nb, 20, scale=1e-5)
    large = _density(rng, nb, 20, scale=1.0)
    seq = [(small, 1), (small, 2), (large, 1), (large, 2)]
    for D, mode in seq:
        got = _run(gg, [D, D], types, hers)
        st = tg.last_stats()
        ref = _run(ge, [D, D], types, hers)
        err = _maxdiff(got, ref)
        say(f"  graph mode {st['graph']} (want {mode}) candidates {st['n_candidates']} max|d| {err:.2e}")
        assert st["graph"] == mode
        assert st["n_candidates"] == te.last_stats()["n_candidates"]
        assert err < TOL * max(1.0, np.abs(ref[0]).max())


def test_signature_change_and_options_build_eagerly():
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    say("case: graph signature change")
    b = cqx.BasisSet("cc-pVDZ", water_cluster(3, 6.0, seed=4))
    nb = b.nBasis
    (tg, gg), (te, ge) = _pair(b)
    rng = np.random.Generator(np.random.PCG64(14))
    D = _density(rng, nb, 15)
    _run(gg, [D, D], [COULOMB, EXCHANGE], [True, True])
    assert tg.last_stats()["graph"] == 1
    _run(gg, [D], [COULOMB], [True])            # another signature: eager, records anew
    assert tg.last_stats()["graph"] == 1
    got = _run(gg, [D], [COULOMB], [True])
    assert tg.last_stats()["graph"] == 2
    tg.set_option("fast_lanes", 4)               # any option drops the recording
    got2 = _run(gg, [D], [COULOMB], [True])
    assert tg.last_stats()["graph"] == 1
    ref = _run(ge, [D], [COULOMB], [True])
    assert _maxdiff(got, ref) < TOL and _maxdiff(got2, ref) < TOL


@pytest.mark.parametrize("case", ["nonhermitian", "unscreened"])
def test_replays_of_nonhermitian_and_unscreened_builds(case):
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    say("case: graph replays,", case)
    b = cqx.BasisSet("6-31G(d)", water_cluster(3, 6.0, seed=8))
    nb = b.nBasis
    (tg, gg), (te, ge) = _pair(b)
    rng = np.random.Generator(np.random.PCG64(15))
    her = case != "nonhermitian"
    screen = case != "unscreened"
    for it in range(3):
        if her:
            Xs = [_density(rng, nb, 12)] * 2
        else:
            Xs = [np.asfortranarray(rng.normal(0, 1 / nb, size=(nb, nb))) for _ in range(2)]
        got = _run(gg, Xs, [COULOMB, EXCHANGE], [her, her], screen)
        mode = tg.last_stats()["graph"]
        ref = _run(ge, Xs, [COULOMB, EXCHANGE], [her, her], screen)
        err = _maxdiff(got, ref)
        say(f"  build {it}: graph mode {mode} max|d| {err:.2e}")
        assert mode == (1 if it == 0 else 2)
        assert err < TOL


@pytest.mark.parametrize("cplx", [False, True])
def test_concurrent_streams_equal_one_stream(cplx):
    """Option streams (ERI launches dealt over 1-4 streams, also inside a
    recorded build): every stream count, eager and replayed, equals the
    one-stream build up to the order of the atomic additions (1e-11)."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    say("case: streams 1-4, eager and recorded, cplx", cplx)
    b = cqx.BasisSet("cc-pVDZ", water_cluster(4, 8.0, seed=0))
    rng = np.random.Generator(np.random.PCG64(71))
    D = _density(rng, b.nBasis, 20, cplx)
    Xs, types, hers = [D, D, D], [COULOMB, EXCHANGE, EXCHANGE], [True, True, False]
    tpi = cqx.DirectTPI(b, 1e-12)
    gto = cqx.GTODirectTPIContraction(tpi)
    tpi.set_option("streams", 1)
    ref = _run(gto, Xs, types, hers)
    for n in (2, 3, 4):
        tpi.set_option("streams", n)
        assert _maxdiff(_run(gto, Xs, types, hers), ref) <= TOL
    tpi.set_option("graph", 1)
    modes = []
    for _ in range(3):
        out = _run(gto, Xs, types, hers)
        modes.append(tpi.last_stats()["graph"])
        assert _maxdiff(out, ref) <= TOL
    assert modes == [1, 2, 2]
    with pytest.raises(cqx.CErr):
        tpi.set_option("streams", 5)
