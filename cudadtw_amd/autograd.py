"""torch.autograd wrappers of the cDTW query gradient (mdtw_cdtw_backward).

Forward passes run the ordinary search / distance entry points; backward
recomputes the banded DP of the matched windows, walks the optimal warping
path back and carries dDist through the query's z-normalisation
(population sigma, stats.cuh:8-28) to the raw query.  The gradient is that of
the piecewise-quadratic cDTW along its optimal path: exact wherever the path
is unique; at a distance of 0 the sqrt is not differentiable and its
contribution is 0.  Window positions (the argmin / top-k selection) are
piecewise constant in the query and carry no gradient.
"""
from __future__ import annotations

import numpy as np
import torch

from .api import METRIC_L2, Engine, znormalize


def _prepare(engine: Engine, query: torch.Tensor, w: int):
    q = query.detach().to(dtype=torch.float32).reshape(-1)
    zq = znormalize(q.cpu().numpy())
    engine.set_query(zq, w)
    return q.to(engine._dev).contiguous()


class CDTWAt(torch.autograd.Function):
    """dist[p] = cDTW(znormalize(query), window indices[p]) for fixed windows
    (sqrt of the squared distance for L2, as search distances are reported)."""

    @staticmethod
    def forward(ctx, query, engine, series, avg, sd, indices, w, metric=METRIC_L2):
        q_dev = _prepare(engine, query, w)
        d2 = engine.gpu_cdtw(series, avg, sd, w, indices=indices, metric=metric)
        dist = torch.sqrt(d2) if metric == METRIC_L2 else d2.clone()
        ctx.engine, ctx.series, ctx.avg, ctx.sd = engine, series, avg, sd
        ctx.w, ctx.metric = w, metric
        ctx.save_for_backward(q_dev, indices)
        ctx.query_dtype, ctx.query_device = query.dtype, query.device
        return dist

    @staticmethod
    def backward(ctx, gdist):
        q_dev, indices = ctx.saved_tensors
        eng = ctx.engine
        _prepare(eng, q_dev, ctx.w)
        g, _ = eng.cdtw_backward(ctx.series, ctx.avg, ctx.sd, indices,
                                 gdist.detach().to(eng._dev, torch.float32).contiguous(), ctx.w,
                                 ctx.metric, sqrt_dist=ctx.metric == METRIC_L2,
                                 query_raw=q_dev)
        g = g.to(device=ctx.query_device, dtype=ctx.query_dtype)
        return g, None, None, None, None, None, None, None


class TopKCDTW(torch.autograd.Function):
    """Distances of the k best non-overlapping matches of ``query`` in
    ``series`` (mdtw_search, exclusion zone excl); returns (dist[k'], pos[k'])
    with k' <= k the matches found.  Differentiable in ``query`` through
    dist."""

    @staticmethod
    def forward(ctx, query, engine, series, avg, sd, w, k, excl, metric=METRIC_L2):
        q_dev = _prepare(engine, query, w)
        r = engine.search(series, w, metric=metric, k=k, excl=excl, avg=avg, sd=sd)
        found = int(np.sum(r.positions >= 0))
        pos = torch.from_numpy(r.positions[:found].astype(np.int32)).to(engine._dev)
        dist = torch.from_numpy(r.distances[:found].copy()).to(engine._dev)
        ctx.engine, ctx.series, ctx.avg, ctx.sd = engine, series, avg, sd
        ctx.w, ctx.metric = w, metric
        ctx.save_for_backward(q_dev, pos)
        ctx.query_dtype, ctx.query_device = query.dtype, query.device
        ctx.mark_non_differentiable(pos)
        return dist, pos

    @staticmethod
    def backward(ctx, gdist, _gpos):
        q_dev, pos = ctx.saved_tensors
        eng = ctx.engine
        if pos.numel() == 0:
            g = torch.zeros_like(q_dev)
        else:
            _prepare(eng, q_dev, ctx.w)
            g, _ = eng.cdtw_backward(ctx.series, ctx.avg, ctx.sd, pos,
                                     gdist.detach().to(eng._dev, torch.float32).contiguous(),
                                     ctx.w, ctx.metric, sqrt_dist=ctx.metric == METRIC_L2,
                                     query_raw=q_dev)
        g = g.to(device=ctx.query_device, dtype=ctx.query_dtype)
        return g, None, None, None, None, None, None, None, None


def cdtw_at(query, engine, series, avg, sd, indices, w, metric=METRIC_L2):
    return CDTWAt.apply(query, engine, series, avg, sd, indices, w, metric)


def topk_cdtw(query, engine, series, avg, sd, w, k, excl, metric=METRIC_L2):
    return TopKCDTW.apply(query, engine, series, avg, sd, w, k, excl, metric)
