"""Batched relaxation: many cases of uneven size and boundary conditions in one
launch per sweep (optionally replayed from a hipGraph) give every case exactly
its own single-case result."""
import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from test_gpu_parity import box, rnd, same

pytestmark = pytest.mark.gpu

SPECS = [((17, 13, 11), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5)),
         ((70, 9, 40), (1, 1, -1, -1, 1, 1), (2.0, 1.0, 1.0), (0.0, 0.0, 1.5)),
         ((67, 10, 37), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.3, 0.0)),
         ((33, 33, 33), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))]


# no case periodic in x: the batch kernels run the plane loop without the run-time periodicity branch
WALL_X_SPECS = [SPECS[0], SPECS[1], SPECS[3]]


def _cases(seed, specs=SPECS):
    cs, data = [], []
    for q, (N, bc, L, st) in enumerate(specs):
        c = ih.Case(N, bc, L, st)
        p, b = rnd(c.shape, seed + 2 * q), rnd(c.shape, seed + 2 * q + 1)
        c.upload(ih.PRE, p)
        c.upload(ih.RES, b)
        cs.append(c)
        data.append((p, b))
    return cs, data


@pytest.mark.parametrize("shape", [(256, 32, 0), (128, 8, 1), (512, 16, 1)])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("sweeps", [1, 6])
def test_batch_jor_equals_single_case(sweeps, graph, shape):
    cs, data = _cases(300)
    ref = []
    for c, (p, b) in zip(cs, data):
        s = ih.Case(c.N, c.bc, c.L, c.stretch)
        s.upload(ih.PRE, p)
        s.upload(ih.RES, b)
        s.jor(sweeps, 0.8, s.Sp, s.Np)
        ref.append(s.download(ih.PRE))
        s.close()
    bt = ih.Batch(cs)
    bt.set_shape(*shape)
    bt.jor(sweeps, 0.8, graph=graph)
    if graph:  # a second replay of the same graph continues from the new state
        bt.jor(sweeps, 0.8, graph=True)
        ref = []
        for c, (p, b) in zip(cs, data):
            s = ih.Case(c.N, c.bc, c.L, c.stretch)
            s.upload(ih.PRE, p)
            s.upload(ih.RES, b)
            s.jor(sweeps, 0.8, s.Sp, s.Np)
            s.jor(sweeps, 0.8, s.Sp, s.Np)
            ref.append(s.download(ih.PRE))
            s.close()
    for q, c in enumerate(cs):
        same(c.download(ih.PRE)[box(c.Sp, c.Np)], ref[q][box(c.Sp, c.Np)], f"case {q}")
    bt.close()


def test_batch_fp32_equals_oracle():
    cs, data = _cases(320)
    bt = ih.Batch(cs)
    tpb, zc, lay, ms = bt.autotune(3)
    assert tpb in (128, 256, 512) and zc in (8, 12, 16, 24, 32) and lay in (0, 1) and ms > 0
    for c, (p, b) in zip(cs, data):   # autotuning leaves PRE untouched
        same(c.download(ih.PRE)[box(c.Sp, c.Np)], p[box(c.Sp, c.Np)], "autotune")
    bt.jor(5, 0.8, fp32=True, graph=True)
    for q, (c, (p, b)) in enumerate(zip(cs, data)):
        oc = O.OracleCase(c.N, c.bc, c.L, c.stretch)
        oc.jor_mixed(5, 0.8, oc.Sp, oc.Np, p, b)
        same(c.download(ih.PRE)[box(c.Sp, c.Np)], p[box(c.Sp, c.Np)], f"fp32 case {q}")
    bt.close()


@pytest.mark.parametrize("shape", [(256, 12, 0), (128, 8, 1)])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_without_periodic_x(graph, shape):
    """The form for batches in which no case is periodic in x, FP64 against
    single-case sweeps and FP32 against the oracle."""
    cs, data = _cases(340, WALL_X_SPECS)
    ref = []
    for c, (p, b) in zip(cs, data):
        s = ih.Case(c.N, c.bc, c.L, c.stretch)
        s.upload(ih.PRE, p)
        s.upload(ih.RES, b)
        s.jor(4, 0.8, s.Sp, s.Np)
        ref.append(s.download(ih.PRE))
        s.close()
    bt = ih.Batch(cs)
    bt.set_shape(*shape)
    bt.jor(4, 0.8, graph=graph)
    for q, c in enumerate(cs):
        same(c.download(ih.PRE)[box(c.Sp, c.Np)], ref[q][box(c.Sp, c.Np)], f"case {q}")
    for c, (p, b) in zip(cs, data):
        c.upload(ih.PRE, p)
    bt.jor(3, 0.8, fp32=True, graph=graph)
    for q, (c, (p, b)) in enumerate(zip(cs, data)):
        oc = O.OracleCase(c.N, c.bc, c.L, c.stretch)
        oc.jor_mixed(3, 0.8, oc.Sp, oc.Np, p, b)
        same(c.download(ih.PRE)[box(c.Sp, c.Np)], p[box(c.Sp, c.Np)], f"fp32 case {q}")
    bt.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_groups(graph):
    """Cases relaxed group after group (a budget of one case's fields: every
    case its own group, and one of two cases) give the same results as one group."""
    cs, data = _cases(360)
    ref = []
    for c, (p, b) in zip(cs, data):
        s = ih.Case(c.N, c.bc, c.L, c.stretch)
        s.upload(ih.PRE, p)
        s.upload(ih.RES, b)
        s.jor(5, 0.8, s.Sp, s.Np)
        ref.append(s.download(ih.PRE))
        s.close()
    bt = ih.Batch(cs)
    for layout in (0, 1):
        for budget, expect in ((1, len(cs)), (2 << 20, None)):
            for c, (p, b) in zip(cs, data):
                c.upload(ih.PRE, p)
            bt.set_shape(256, 12, layout)
            bt.set_groups(budget)
            nb, cnt = bt.groups()
            assert nb == budget and cnt >= 1
            if expect:
                assert cnt == expect
            bt.jor(5, 0.8, graph=graph)
            for q, c in enumerate(cs):
                same(c.download(ih.PRE)[box(c.Sp, c.Np)], ref[q][box(c.Sp, c.Np)], f"case {q}, budget {budget}")
    bt.close()
