"""Serving loop (continuous batching) through the C ABI.

Bars: a request's columns are bitwise equal to the same columns of the batch build (each
lane's arithmetic is its own), bitwise equal alone and in a mix of requests that arrive
while it is in flight (lanes are recycled as columns finish), forward solves within 1e-8
of the oracle's shear_traction, and K columns within 1e-8 of the oracle.
"""
import numpy as np
import pytest

import eqd
from eqd.serve import Server
from oracle.oracle import Geometry, OracleSystem

pytestmark = pytest.mark.gpu

RING = (9, 17, 400.0, 800.0, 32.04e9, 32.04e9, 800.0)
CUBE = (11, 11, 1.0, 1.0, 2.0, 1.0, 0.4)


def grid_args(case):
    n1, n23, Ln, Lf, lam, mu, lf = case
    return eqd.GridPair(n1, n23, Ln, Lf), lam, mu, lf


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ring_batch():
    g, lam, mu, lf = grid_args(RING)
    fe = eqd.FaultElasticity(g, lam, mu, l_f=lf, precision="mixed", tile_k=64)
    return eqd.fault_stiffness(fe)


def test_columns_match_batch_build_bitwise(ring_batch):
    print("[serve columns vs batch]", flush=True)
    g, lam, mu, lf = grid_args(RING)
    sv = Server(precision="mixed")
    gid = sv.add_grid(g, lam, mu, l_f=lf, max_columns=128)
    rng = np.random.default_rng(1)
    cols = rng.choice(ring_batch.shape[1], size=37, replace=False)
    rid = sv.submit_columns(gid, cols)
    done = sv.drain()
    assert done == [rid]
    out, iters, unc = sv.result(rid, gid, len(cols))
    assert unc == 0 and iters.max() > 0  # ring columns (b = 0) finish with 0 iterations
    assert np.array_equal(out, ring_batch[:, cols])


def test_same_result_alone_and_in_a_mix(ring_batch):
    print("[serve alone vs mix]", flush=True)
    g, lam, mu, lf = grid_args(RING)
    nf2 = ring_batch.shape[1]
    rng = np.random.default_rng(2)
    probe = rng.choice(nf2, size=20, replace=False)
    nf = nf2 // 2
    s2 = 1e-3 * rng.standard_normal(nf)
    s3 = 1e-3 * rng.standard_normal(nf)

    alone = Server(precision="mixed")
    ga = alone.add_grid(g, lam, mu, l_f=lf, max_columns=64)
    ra = alone.submit_columns(ga, probe)
    fa = alone.submit_forward(ga, s2, s3)
    alone.drain()
    A, _, _ = alone.result(ra, ga, len(probe))
    FA, _, _ = alone.result(fa, ga, 1)

    # a one-tile pool (64 lanes) and more work than fits: later requests wait for lanes,
    # and the probe arrives while others are in flight
    mix = Server(precision="mixed")
    gm = mix.add_grid(g, lam, mu, l_f=lf, max_columns=64)
    others = [mix.submit_columns(gm, rng.choice(nf2, size=k, replace=False)) for k in (30, 25)]
    for _ in range(3):
        mix.step()
    rb = mix.submit_columns(gm, probe)
    others.append(mix.submit_columns(gm, rng.choice(nf2, size=40, replace=False)))
    fb = mix.submit_forward(gm, s2, s3)
    mix.step()
    others.append(mix.submit_forward(gm, 2 * s2, -s3))
    done = mix.drain()
    assert set(done) == set(others) | {rb, fb}
    B, _, _ = mix.result(rb, gm, len(probe))
    FB, _, _ = mix.result(fb, gm, 1)
    assert np.array_equal(A, B)
    assert np.array_equal(FA, FB)
    assert np.array_equal(A, ring_batch[:, probe])


def test_forward_solve_matches_oracle():
    print("[serve forward vs oracle]", flush=True)
    g, lam, mu, lf = grid_args(RING)
    o = OracleSystem(Geometry(*RING))
    rng = np.random.default_rng(3)
    nf = o.nf
    x2 = np.linspace(-800, 800, 17)
    X, Y = np.meshgrid(x2, x2, indexing="ij")
    bump = 1e-3 * np.exp(-((X - 100) ** 2 + (Y + 200) ** 2) / (2 * 250.0 ** 2)).ravel(order="F")
    sv = Server(precision="mixed")
    gid = sv.add_grid(g, lam, mu, l_f=lf, max_columns=64)
    r1 = sv.submit_forward(gid, bump, np.zeros(nf))
    r2 = sv.submit_forward(gid, 0.3 * bump, rng.standard_normal(nf) * 1e-4)
    sv.drain()
    for rid, (s2, s3) in [(r1, (bump, np.zeros(nf))), (r2, (0.3 * bump, None))]:
        out, _, unc = sv.result(rid, gid, 1)
        assert unc == 0
        if s3 is None:
            continue
        d2, d3 = o.shear_traction(s2, s3)
        ref = np.concatenate([d2, d3])
        assert rel(out[:, 0], ref) <= 1e-8


def test_two_grids_in_one_server():
    print("[serve two grids]", flush=True)
    sv = Server(precision="mixed")
    gids = []
    for case in (RING, CUBE):
        g, lam, mu, lf = grid_args(case)
        gids.append(sv.add_grid(g, lam, mu, l_f=lf, max_columns=64))
    Ks = []
    for case in (RING, CUBE):
        Ks.append(OracleSystem(Geometry(*case)).fault_stiffness(nthreads=8))
    reqs = []
    rng = np.random.default_rng(4)
    for k in range(6):
        gi = k % 2
        cols = rng.choice(Ks[gi].shape[1], size=min(12, Ks[gi].shape[1]), replace=False)
        reqs.append((gi, cols, sv.submit_columns(gids[gi], cols)))
    sv.drain()
    for gi, cols, rid in reqs:
        out, _, unc = sv.result(rid, gids[gi], len(cols))
        assert unc == 0
        assert rel(out, Ks[gi][:, cols]) <= 1e-8


def test_bad_requests_rejected():
    g, lam, mu, lf = grid_args(CUBE)
    sv = Server()
    gid = sv.add_grid(g, lam, mu, l_f=lf, max_columns=64)
    nf2 = 2 * sv.grids[gid]["nf"]
    with pytest.raises(eqd.EqdError):
        sv.submit_columns(gid, [nf2])
    with pytest.raises(eqd.EqdError):
        sv.submit_columns(gid + 5, [0])
    with pytest.raises(eqd.EqdError):
        sv.submit_columns(gid, np.arange(65) % nf2)  # larger than the pool


def test_lanes_solved_at_admission_leave_neighbours_alone():
    """Columns whose right-hand side vanishes (unit slip on the locked ring, l_f = L_fault)
    finish at admission; re-initialising their lanes as idle must not switch off the tile
    that still holds an active column (regression: a divergent tile-activity reduction)."""
    print("[serve solved-at-admission neighbours]", flush=True)
    g, lam, mu, lf = grid_args(RING)
    nf = 17 * 17
    rng = np.random.default_rng(5)
    s2 = 1e-3 * rng.standard_normal(nf)
    s3 = 1e-3 * rng.standard_normal(nf)
    ring = [i for i in range(nf) if i % 17 in (0, 16) or i // 17 in (0, 16)][:30]
    results = []
    for before in (None, ring):
        sv = Server(precision="mixed")
        gid = sv.add_grid(g, lam, mu, l_f=lf, max_columns=64)
        if before is not None:
            sv.submit_columns(gid, before)
        f = sv.submit_forward(gid, s2, s3)
        sv.drain()
        out, iters, unc = sv.result(f, gid, 1)
        assert unc == 0 and iters[0] > 40
        results.append(out)
    assert np.array_equal(results[0], results[1])
