"""Serving loop on the GPU: every result is the same alone and in a concurrent mix.

Alone: one slot, each request waits for the previous result.  Mix: three slots,
all requests submitted at once, so different systems overlap on the device.
Frames of one system keep their order in both runs (same warm-start chain) and
every reduction in the library has a fixed order, so the results agree to
rounding; the tolerance written here is 1e-10 relative.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _stream():
    from dmff_amd.admp.serving import water_spec
    rng = np.random.default_rng(7)
    systems = [water_spec(f"s{k}", n, 100 + k) for k, n in enumerate((64, 110, 180))]
    reqs = []
    order = [0, 1, 2, 0, 1, 0, 2]
    for k in order:
        spec, pos0, box = systems[k]
        pos = pos0 + rng.normal(0.0, 0.02, pos0.shape) if reqs and any(s is spec for s, _, _ in reqs) else pos0
        reqs.append((spec, pos, box))
    return reqs


def _run(slots, sequential):
    from dmff_amd.admp.serving import ADMPServer, Request
    srv = ADMPServer(device=DEV, precision="fp64", slots=slots, tol=1e-7)
    out = {}
    futs = []
    try:
        for rid, (spec, pos, box) in enumerate(_stream()):
            f = srv.submit(Request(rid, spec, pos, box))
            if sequential:
                r = f.result(timeout=120)
                out[rid] = r
            futs.append((rid, f))
        for rid, f in futs:
            out[rid] = f.result(timeout=120)
    finally:
        srv.close()
    return out


def test_results_alone_equal_results_in_mix():
    alone = _run(1, True)
    mix = _run(3, False)
    assert set(alone) == set(mix)
    for rid in alone:
        a, m = alone[rid], mix[rid]
        assert a.warm == m.warm
        assert abs(a.energy - m.energy) <= 1e-10 * abs(a.energy), (rid, a.energy, m.energy)
        d = float(torch.linalg.norm(a.dpos - m.dpos) / torch.linalg.norm(a.dpos))
        assert d <= 1e-10, (rid, d)
    # follow-on frames started from the previous frame's induced dipoles
    assert any(r.warm for r in alone.values())
