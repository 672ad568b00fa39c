"""Serving engine (rf_server_*): continuous batching of independent
suspensions on one GPU.

  * each result is bitwise the same served alone and served in a mix;
  * a served suspension follows the single-simulation path (Simulation) to
    the GMRES tolerance;
  * admission respects the memory budget and the slot count, and a request
    whose footprint exceeds the budget is refused;
  * a request that turns non-finite fails alone, its neighbours unchanged.
"""
import numpy as np
import pytest

from helpers import ref_config, rel_err

pytestmark = pytest.mark.gpu


def _params(tol=1e-8):
    p = ref_config(validate=False)
    p.GMRES_TOLERANCE = tol
    return p


def _requests(sizes, steps, seed=7):
    from rigidfibers_amd.scenes import cube

    return [(M, *cube(M, (10.0 * M) ** (1 / 3), min_dist=0.4, seed=seed + k), steps) for k, M in enumerate(sizes)]


def _serve(server, reqs):
    ids = [(server.submit(pos, ori, steps), M, steps) for M, pos, ori, steps in reqs]
    out = []
    for rid, M, steps in ids:
        assert server.wait(rid, timeout=120.0)
        out.append(server.collect(rid, M, steps))
    return out


def test_result_same_alone_and_in_mix():
    from rigidfibers_amd.serve import Server

    reqs = _requests([23, 64, 17, 120, 45, 90], steps=4)
    with Server(_params(), max_active=8) as srv:
        alone = [_serve(srv, [r])[0] for r in reqs]
        mixed = _serve(srv, reqs)
        st = srv.stats()
    for a, b in zip(alone, mixed):
        assert a["info"]["status"] == 0 and b["info"]["status"] == 0
        assert np.array_equal(a["positions"], b["positions"]) and np.array_equal(a["orientations"], b["orientations"])
        assert a["info"]["gmres_iterations"] == b["info"]["gmres_iterations"]
    assert st["completed"] == 2 * len(reqs) and st["queued"] == 0 and st["active"] == 0


def test_graph_replay_matches_eager_launches():
    """Each tick replays its captured graph of a round of GMRES iterations
    (kernels and device-side control steps) several times, with a device to
    host copy and a synchronisation between replays; the results must be
    bitwise those of launching the same kernels one by one."""
    from rigidfibers_amd.serve import Server

    reqs = _requests([23, 64, 17, 120, 45], steps=4, seed=5)
    out = {}
    for eager in (False, True):
        with Server(_params(), max_active=8, eager=eager) as srv:
            out[eager] = _serve(srv, reqs)
            assert srv.stats()["ticks"] >= 4
    for g, e in zip(out[False], out[True]):
        assert g["info"]["status"] == 0 and e["info"]["status"] == 0
        assert g["info"]["gmres_iterations"] == e["info"]["gmres_iterations"] > 4
        assert np.array_equal(g["positions"], e["positions"]) and np.array_equal(g["orientations"], e["orientations"])


@pytest.mark.parametrize("integration", ["numerical", "analytical"])
def test_served_matches_simulation(integration):
    from rigidfibers_amd.serve import Server
    from rigidfibers_amd.simulation import Simulation

    params = _params(tol=1e-10)
    reqs = _requests([37, 80], steps=5, seed=11)
    with Server(params, integration=integration, max_active=4) as srv:
        served = _serve(srv, reqs)
    for (M, pos, ori, steps), r in zip(reqs, served):
        p = _params(tol=1e-10)
        p.NUMBER_OF_FIBERS = M
        with Simulation(p, pos, ori, solver="gmres", integration=integration) as sim:
            for t in range(steps):
                sim.step(t)
            ps, os_ = sim.state()
        ep, eo = rel_err(r["positions"], ps), rel_err(r["orientations"], os_)
        print(f"M={M} {integration}: pos {ep:.1e} ori {eo:.1e}")
        assert ep < 1e-8 and eo < 1e-8


def test_served_confirm_residual_same_result():
    """With confirm_residual the device-side control computes the true
    residual after a converged cycle before it stops; the result and the
    Arnoldi step counts are those of stopping on the cycle's estimate."""
    from rigidfibers_amd.serve import Server

    reqs = _requests([23, 64, 45], steps=3, seed=9)
    out = {}
    for confirm in (False, True):
        with Server(_params(), max_active=4, confirm_residual=confirm) as srv:
            out[confirm] = _serve(srv, reqs)
    for a, b in zip(out[False], out[True]):
        assert a["info"]["status"] == 0 and b["info"]["status"] == 0
        assert a["info"]["gmres_iterations"] == b["info"]["gmres_iterations"] > 3
        assert np.array_equal(a["positions"], b["positions"]) and np.array_equal(a["orientations"], b["orientations"])


def test_bad_request_fails_alone():
    """A request whose first two fibers coincide fails with RF_ERR_NONFINITE;
    the requests served beside it give bitwise the results they give alone."""
    from rigidfibers_amd import _lib
    from rigidfibers_amd.serve import Server

    reqs = _requests([23, 40, 31], steps=3, seed=13)
    M, pos, ori, steps = reqs[1]
    pos, ori = pos.copy(), ori.copy()
    pos[3:6], ori[3:6] = pos[0:3], ori[0:3]
    bad = (M, pos, ori, steps)
    with Server(_params(), max_active=8) as srv:
        alone = [_serve(srv, [r])[0] for r in (reqs[0], reqs[2])]
        mixed = _serve(srv, [reqs[0], bad, reqs[2]])
        st = srv.stats()
    assert mixed[1]["info"]["status"] == _lib.RF_ERR_NONFINITE
    for a, b in zip(alone, (mixed[0], mixed[2])):
        assert a["info"]["status"] == 0 and b["info"]["status"] == 0
        assert a["info"]["gmres_iterations"] == b["info"]["gmres_iterations"]
        assert np.array_equal(a["positions"], b["positions"]) and np.array_equal(a["orientations"], b["orientations"])
    assert st["active"] == 0 and st["queued"] == 0


def test_memory_budget_and_slots():
    from rigidfibers_amd.serve import Server

    reqs = _requests([60, 60, 60, 60], steps=3, seed=3)
    with Server(_params(), max_active=8) as srv:  # learn the footprint
        fp = _serve(srv, reqs[:1])[0]["info"]["footprint"]
    budget = int(1.5 * fp)  # room for one suspension at a time
    with Server(_params(), memory_budget=budget, max_active=8) as srv:
        out = _serve(srv, reqs)
        st = srv.stats()
    assert st["peak_bytes"] <= budget and st["memory_budget"] == budget
    spans = sorted((r["info"]["admitted"], r["info"]["finished"]) for r in out)
    for (a0, f0), (a1, f1) in zip(spans, spans[1:]):
        assert a1 >= f0 - 1e-6  # never two admitted together
    with Server(_params(), max_active=2) as srv:
        out = _serve(srv, _requests([30, 31, 32, 33, 34], steps=2, seed=5))
    assert all(r["info"]["status"] == 0 for r in out)


def test_oversized_request_refused():
    from rigidfibers_amd import _lib
    from rigidfibers_amd.serve import Server

    (M, pos, ori, steps), = _requests([50], steps=1)
    with Server(_params(), memory_budget=1 << 20, max_active=4) as srv:
        with pytest.raises(_lib.RfError) as e:
            srv.submit(pos, ori, steps)
    assert e.value.code == _lib.RF_ERR_OUT_OF_MEMORY


def test_served_mix_matches_oracle(oracle64):
    """A mix spanning the serving stream's range of sizes, M = 16 .. 400 (cube
    at number density 0.1, GMRES tolerance 1e-8, restart 10, each solve
    starting from the previous solution), served together for 10 steps:
    every suspension's state against the FP64 oracle's own trajectory
    (fortran/fibers.f95:182-349 per suspension, as the reference runs one
    suspension per process, src/main.cu:83-99) within 1e-5."""
    from oracle.pyoracle import make_params
    from rigidfibers_amd.serve import Server

    steps = 10
    reqs = _requests([16, 400, 57, 230, 121, 333], steps=steps, seed=31)
    with Server(_params(tol=1e-8), max_active=8) as srv:
        served = _serve(srv, reqs)
    for (M, pos, ori, _), r in zip(reqs, served):
        assert r["info"]["status"] == 0
        p_ref, o_ref = oracle64.simulate(make_params(M, threads=8), pos, ori, steps, solver="gmres",
                                         gmres=(10, 1000, 1e-8))
        ep, eo = rel_err(r["positions"], p_ref), rel_err(r["orientations"], o_ref)
        print(f"M={M}: positions {ep:.1e}, orientations {eo:.1e} against the FP64 oracle")
        assert ep <= 1e-5 and eo <= 1e-5


def test_full_request_stream():
    """The whole serving stream -- 512 suspensions, M ~ U[16, 400], 20 steps
    each, GMRES tolerance 1e-6 -- with arrivals ten times faster than its 50
    per second: every request finishes with status OK after its 20 steps, and
    eight of them, run again alone, give bitwise the same final state."""
    from rigidfibers_amd.serve import Server, request_stream, serve_stream

    reqs = request_stream(count=512, m_range=(16, 400), density=0.1, rate=50.0, steps=20, seed=2026)
    with Server(_params(tol=1e-6), max_active=64) as srv:
        res, wall = serve_stream(srv, reqs, time_scale=0.1)
        st = srv.stats()
        print(f"512 requests in {wall:.2f} s, {st['ticks']} ticks, peak {st['peak_bytes'] / 1e6:.0f} MB")
        assert len(res) == 512 and all(r["status"] == 0 for r in res)
        assert all(len(r["step_ms"]) == 20 for r in res)
        assert st["completed"] == 512 and st["queued"] == 0 and st["active"] == 0
        assert st["fiber_steps"] == 20 * sum(M for _, M, _, _, _ in reqs)
        for k in (0, 1, 100, 255, 256, 300, 450, 511):
            _, M, pos, ori, steps = reqs[k]
            alone = _serve(srv, [(M, pos, ori, steps)])[0]
            assert alone["info"]["status"] == 0
            assert np.array_equal(alone["positions"], res[k]["positions"])
            assert np.array_equal(alone["orientations"], res[k]["orientations"])
            assert alone["info"]["gmres_iterations"] == res[k]["gmres_iterations"]


def test_iteration_cap_with_confirmation_fails_no_neighbour():
    """GMRES(2) capped at 3 Arnoldi steps per solve, confirming each cycle's
    estimate with the true residual, against an unreachable tolerance: every
    suspension ends each step unconverged (status RF_WARN_NOT_CONVERGED, not
    a device error), its result is the same alone and in the mix, and the
    capped steps count."""
    from rigidfibers_amd import _lib
    from rigidfibers_amd.serve import Server

    params = _params(tol=1e-14)
    params.GMRES_RESTART = 2
    params.GMRES_MAX_ITERATIONS = 3
    reqs = _requests([23, 64, 17], steps=3, seed=13)
    with Server(params, max_active=4, confirm_residual=True) as srv:
        alone = [_serve(srv, [r])[0] for r in reqs]
        mixed = _serve(srv, reqs)
    for a, b in zip(alone, mixed):
        assert a["info"]["status"] == b["info"]["status"] == _lib.RF_WARN_NOT_CONVERGED
        assert a["info"]["gmres_iterations"] == b["info"]["gmres_iterations"] == 3 * 3
        assert np.array_equal(a["positions"], b["positions"]) and np.array_equal(a["orientations"], b["orientations"])
