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


@pytest.mark.parametrize("mode", ["bicgstab", "bicgstab_mixed"])
def test_serving_alone_equals_mix(mode):
    serve, NewtonOptions = _mods()
    reqs = serve.request_stream(7, rate=40.0, seed=23, small=True, max_points=8)
    kinds = sorted({r.kind for r in reqs})
    print(f"[serve] {len(reqs)} requests {kinds}, {mode}", flush=True)
    big = sorted(r.spec.N for r in reqs)[-3:]
    opts = NewtonOptions(linsolve=mode)
    S = serve.Server(capacity_nodes=sum(big), opts=opts)
    served = S.run(reqs, realtime=False)
    print(f"  rounds {S.rounds} rebuilds {S.rebuilds} max resident {S.max_resident}",
          flush=True)
    assert S.max_resident >= 2 and S.rebuilds >= 2
    for r, sv in zip(reqs, served):
        alone = serve.serve_alone(r.spec, opts)
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
