"""Query scratch reuse: a context keeps one scratch buffer per stream for its query and
backward calls, grows it when a call needs more, and recycles it for the next call on that
stream without a host wait.  Answers must not depend on which buffer a call gets: calls of
growing and shrinking size on one stream, and calls on two streams issued back to back
with no host synchronisation, all equal the brute force."""
import numpy as np
import pytest
import torch

import gpurt
from gpurt import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_calls_share_per_stream_scratch(ctx, dev):
    w = W.rays_scene(n_rays=20000, grid=101, n_spheres=6)
    V, F, R = w["verts"], w["tris"], w["rays"]
    ep, et, euv = O.ray_closest_brute_f32(V, F, R)
    eo = O.ray_any_brute_f32(V, F, R)
    c = W.cpq_small(n_queries=3000)
    cp = O.cpq_brute_f32(c["verts"], c["tris"], c["queries"])[0]
    bvh = gpurt.Bvh(ctx, _t(V, dev), _t(F, dev))
    cb = gpurt.Bvh(ctx, _t(c["verts"], dev), _t(c["tris"], dev))
    tr, tq = _t(R, dev), _t(c["queries"], dev)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(dev), torch.cuda.Stream(dev)
    plan = [(s1, "closest", 64), (s1, "closest", 20000), (s1, "closest", 1000), (s2, "any", 20000),
            (s1, "cpq", 3000), (s2, "closest", 20000), (s1, "any", 500), (s2, "cpq", 2000), (s1, "closest", 20000)]
    outs = []
    for s, kind, k in plan:  # no host synchronisation between the calls
        with torch.cuda.stream(s):
            if kind == "closest":
                outs.append((kind, k, bvh.trace_closest(tr[:k])))
            elif kind == "any":
                outs.append((kind, k, (bvh.trace_any(tr[:k]),)))
            else:
                outs.append((kind, k, (cb.closest_point(tq[:k])[0],)))
    torch.cuda.synchronize()
    for kind, k, res in outs:
        print("call", kind, k, flush=True)
        if kind == "closest":
            p, t, uv = (x.cpu().numpy() for x in res)
            assert np.array_equal(p, ep[:k]) and np.array_equal(t, et[:k]) and np.array_equal(uv, euv[:k])
        elif kind == "any":
            assert np.array_equal(res[0].cpu().numpy(), eo[:k])
        else:
            assert np.array_equal(res[0].cpu().numpy(), cp[:k])
