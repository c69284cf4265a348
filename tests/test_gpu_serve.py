"""Serving loop on a pool of resident slots: requests give the same result in
the mix as alone; the pool respects the HBM budget; a size whose slot cannot
fit is refused; the HBM estimate of a case includes its multigrid levels."""
import pytest

import impact_hip as ih

pytestmark = pytest.mark.gpu

REQS = [dict(arrival_s=0.0, N=(17, 13, 13), steps=3, radius=0.30, stenosis=0.5),
        dict(arrival_s=0.0, N=(21, 17, 17), steps=2, radius=0.25, stenosis=0.3),
        dict(arrival_s=0.01, N=(33, 17, 17), steps=2, radius=0.35, stenosis=0.7),
        dict(arrival_s=0.02, N=(25, 13, 13), steps=3, radius=0.20, stenosis=0.6)]


def _step():
    return ih.StepParams(Re=500.0, dtime=5e-4, jor_sweeps=10, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=2000,
                         level_pressure=1, inflow_kind=2, inflow_amp=1.0, use_mask=1, jor_fp32=0, precond_sweeps=0)


def test_results_same_alone_and_in_the_mix():
    alone = []
    for r in REQS:
        res, summ, rc = ih.serve([dict(r, arrival_s=0.0)], _step(), hbm_budget=1 << 34, max_active=1)
        assert rc == 0 and res[0].status == 0
        alone.append(res[0])
    mixed, summ, rc = ih.serve(REQS, _step(), hbm_budget=1 << 34, max_active=4)
    assert rc == 0
    assert summ.peak_active >= 2   # the requests did overlap
    for a, m in zip(alone, mixed):
        assert m.status == 0
        assert (m.checksum_vel, m.checksum_pre, m.bicg_iterations) == (a.checksum_vel, a.checksum_pre,
                                                                        a.bicg_iterations)
        assert m.latency_s >= m.finish_s - m.admit_s > 0


def test_budget_limits_the_pool():
    sizes = [ih.load().ih_grid_bytes(ih.grid_desc(r["N"], (1,) * 6)) for r in REQS]
    budget = sum(sizes)                       # one slot of each size, never two
    reqs = REQS + [dict(r, arrival_s=0.03) for r in REQS]
    res, summ, rc = ih.serve(reqs, _step(), hbm_budget=budget, max_active=8)
    assert rc == 0 and all(r.status == 0 for r in res)
    assert summ.pool_slots == 4 and summ.pool_bytes <= budget
    assert summ.peak_bytes <= budget and summ.peak_active <= 4
    assert summ.setup_s > 0 and summ.admit_wait_max_s >= summ.admit_wait_mean_s >= 0
    big = dict(REQS[0], N=(65, 65, 65))
    res, summ, rc = ih.serve([big], _step(), hbm_budget=min(sizes), max_active=2)
    assert res[0].status == -12 and rc == -12      # IH_ENOMEM: larger than the whole budget


def test_hbm_estimate_includes_the_multigrid_levels():
    """ih_grid_bytes_step adds what building the multigrid levels adds to the
    case's HBM (ih_case_bytes), for a step with preconditioner 1."""
    N = (33, 17, 17)
    g = ih.grid_desc(N, (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    L = ih.load()
    plain = ih.StepParams(precond_poisson=0)
    mg = ih.StepParams(precond_poisson=1)
    base, with_mg = L.ih_grid_bytes_step(g, plain), L.ih_grid_bytes_step(g, mg)
    assert base == L.ih_grid_bytes(g) and with_mg > base
    c = ih.Case(N, (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    b0 = c.bytes()
    assert b0 <= base
    c.multigrid_v(ih.TMP1, ih.TMP2)      # builds the levels
    assert c.bytes() - b0 == with_mg - base
    c.close()
