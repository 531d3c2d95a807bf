"""Serving loop on the GPU: requests of mixed kind (2D diode, 2D MOSFET, 3D
tri-gate) arriving over time, admitted into a pool too small to hold them
all at once, must each give exactly the result they give alone (same
bits: the kernels never mix devices and per-device reductions run in a
fixed order), and the run must actually overlap requests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mods():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from tcadhip import serve
    from tcadhip.device import NewtonOptions
    return serve, NewtonOptions


@pytest.mark.parametrize("mode,predictor", [("bicgstab", False), ("bicgstab_mixed", False),
                                            ("bicgstab", True)])
def test_serving_alone_equals_mix(mode, predictor):
    """Each request gives the bits alone that it gives in the mix, with the
    reference's warm start and with the secant predictor between its
    continuation steps."""
    serve, NewtonOptions = _mods()
    reqs = serve.request_stream(7, rate=40.0, seed=23, small=True, max_points=8)
    kinds = sorted({r.kind for r in reqs})
    print(f"[serve] {len(reqs)} requests {kinds}, {mode}, predictor {predictor}", flush=True)
    big = sorted(r.spec.N for r in reqs)[-3:]
    opts = NewtonOptions(linsolve=mode)
    S = serve.Server(capacity_nodes=sum(big), opts=opts, predictor=predictor)
    served = S.run(reqs, realtime=False)
    print(f"  rounds {S.rounds} rebuilds {S.rebuilds} max resident {S.max_resident}",
          flush=True)
    assert S.max_resident >= 2 and S.rebuilds >= 2
    for r, sv in zip(reqs, served):
        alone = serve.serve_alone(r.spec, opts, predictor=predictor)
        mix = sv.result
        assert mix.failed == alone.failed, (r.spec.name, mix.failed, alone.failed)
        assert mix.newton == alone.newton, r.spec.name
        assert len(mix.currents) == len(alone.currents) == len(r.spec.sweep) or mix.failed
        for a, b in zip(mix.currents, alone.currents):
            assert a.keys() == b.keys()
            assert all(a[k] == b[k] for k in a), (r.spec.name, a, b)
        assert sv.done >= sv.admitted >= sv.arrival


def test_serving_lanes_alone_equals_mix():
    """Two lanes (own HIP stream and host thread each), real-time arrivals."""
    serve, NewtonOptions = _mods()
    reqs = serve.request_stream(6, rate=30.0, seed=29, small=True, max_points=6)
    print(f"[serve-lanes] {len(reqs)} requests, 2 lanes", flush=True)
    opts = NewtonOptions()
    S = serve.Server(capacity_nodes=2 * max(r.spec.N for r in reqs) + 100, opts=opts, lanes=2)
    served = S.run(reqs, realtime=True)
    assert all(L.rounds > 0 for L in S.lanes)
    for r, sv in zip(reqs, served):
        alone = serve.serve_alone(r.spec, opts)
        assert sv.result.failed == alone.failed
        assert sv.result.newton == alone.newton
        for a, b in zip(sv.result.currents, alone.currents):
            assert all(a[k] == b[k] for k in a), (r.spec.name, a, b)


def test_serving_at_workload_sizes():
    """Serving at its stated sizes: a seeded stream with a 2D diode
    (50 x 40 = 2,000 nodes), a 2D MOSFET (100 x 60 = 6,000 nodes, 0.2 nm
    first cell) and a 3D tri-gate (32 x 24 x 24 = 18,432 nodes), plus a
    second diode, 5 points each, in a pool that cannot hold all four at
    once.  Every request gives the same bits alone and in the mix; the 2D
    requests match the oracle (Newton +-1, currents to the bar of
    tests/current_bar.py)."""
    serve, NewtonOptions = _mods()
    from current_bar import check, oracle_sweep
    reqs = _size_stream(serve)
    cap = reqs[2].spec.N + reqs[1].spec.N + 100
    print(f"[serve-sizes] {[r.spec.N for r in reqs]} nodes, pool {cap}", flush=True)
    opts = NewtonOptions()
    S = serve.Server(capacity_nodes=cap, opts=opts)
    served = S.run(reqs, realtime=False)
    print(f"  rounds {S.rounds} rebuilds {S.rebuilds} max resident {S.max_resident}",
          flush=True)
    assert S.max_resident >= 2 and S.rebuilds >= 2
    for r, sv in zip(reqs, served):
        mix = sv.result
        assert not mix.failed, r.spec.name
        alone = serve.serve_alone(r.spec, opts)
        assert mix.newton == alone.newton, r.spec.name
        for a, b in zip(mix.currents, alone.currents):
            assert all(a[k] == b[k] for k in a), (r.spec.name, a, b)
        if r.spec.dim == 2:
            oits, ocur = oracle_sweep(r.spec)
            names = [c.name for c in r.spec.contacts]
            got = np.array([[c[nm] for nm in names] for c in mix.currents])
            print(f"  {r.spec.name}: Newton {mix.newton} oracle {oits.tolist()}", flush=True)
            assert np.all(np.abs(np.array(mix.newton) - oits) <= 1)
            check(got, ocur, r.spec.name, say=lambda m: print(m, flush=True))


def test_serving_with_predicted_forcing():
    """The stream of test_serving_at_workload_sizes served as bench.py
    serves: the secant predictor, and quadratic Newton forcing terms
    (0.1, 1e-3) on the devices whose start it predicted.  Every request
    gives the same bits alone (with the same options) and in the mix; the
    2D requests' currents match the oracle to the bar of
    tests/current_bar.py (their Newton counts follow the predicted starts
    and the forcing terms, so no +-1 bar applies)."""
    serve, NewtonOptions = _mods()
    from current_bar import check, oracle_sweep
    reqs = _size_stream(serve)
    cap = reqs[2].spec.N + reqs[1].spec.N + 100
    opts = NewtonOptions(forcing=True, forcing_pow=2, forcing_slope=0.1, forcing_max=1e-3,
                         forcing_predicted=True)
    S = serve.Server(capacity_nodes=cap, opts=opts, predictor=True)
    served = S.run(reqs, realtime=False)
    for r, sv in zip(reqs, served):
        mix = sv.result
        assert not mix.failed, r.spec.name
        alone = serve.serve_alone(r.spec, opts, predictor=True)
        assert mix.newton == alone.newton, r.spec.name
        for a, b in zip(mix.currents, alone.currents):
            assert all(a[k] == b[k] for k in a), (r.spec.name, a, b)
        if r.spec.dim == 2:
            oits, ocur = oracle_sweep(r.spec)
            names = [c.name for c in r.spec.contacts]
            got = np.array([[c[nm] for nm in names] for c in mix.currents])
            print(f"  {r.spec.name}: Newton {mix.newton} oracle {oits.tolist()}", flush=True)
            check(got, ocur, r.spec.name, say=lambda m: print(m, flush=True))


def _size_stream(serve):
    """Four seeded requests at the server's stated sizes (see
    test_serving_at_workload_sizes)."""
    from tcadhip import configs
    rng = np.random.default_rng(20261204)
    arrivals = np.cumsum(rng.exponential(0.05, size=4))
    specs = [
        ("diode", configs.diode_2d(50, 40, Na=3e17, Nd=2e16, xj_frac=0.45,
                                   biases=[-0.4 * (s + 1) for s in range(5)],
                                   models=dict(doping_mobility=True, srh=True, auger=True,
                                               bgn=True, surface_mobility=False),
                                   material={}, name="req0_diode_50x40")),
        ("mosfet", configs.mosfet_2d(100, 60, Vd=0.5, vg_list=np.linspace(0.0, 1.0, 5),
                                     name="req1_mosfet_100x60")),
        ("trigate", configs.trigate_3d(32, 24, 24, Vd=0.7, vg_list=np.linspace(0.0, 0.8, 5),
                                       name="req2_trigate_32x24x24")),
        ("diode", configs.diode_2d(60, 40, Na=1e16, Nd=1e18, xj_frac=0.6,
                                   biases=[0.12 * (s + 1) for s in range(5)],
                                   models=dict(doping_mobility=True, srh=True, auger=True,
                                               bgn=True, surface_mobility=False),
                                   material={}, name="req3_diode_60x40")),
    ]
    return [serve.Request(i, sp, float(t), kind) for i, ((kind, sp), t)
            in enumerate(zip(specs, arrivals))]
