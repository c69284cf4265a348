"""Pins of the oracle's Windkessel models, the label-driven Dirichlet pressure
and the device_jacobi_solver stopping rule (src/usr_func.f90:654-855,
src/set_pressure_BC_kernel.cu:1-42, src/jacobi_solver_host_v4.f90:283-420).

The systolic pressure of get_pressure_and_flow_rate_WKM is checked against an
independent numerical solution of the two-element Windkessel ODE
C dp/dt + p/R = Q sin(pi t/Ts), p(0) = 80 mmHg.  The reference's closed form
differs from that solution by one term: its sine term is Ts sin(pi t/Ts)/D
where the particular solution has Ts^2 Q R sin(pi t/Ts)/D (D = Ts^2 + C^2 pi^2
R^2).  The test states that difference and checks the rest of the formula
against the ODE, so a wrong exponent, cosine term or constant fails."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle_lib as O
from ct_image import GRID, LENGTHS, OFFSETS, label_image, read_waveform, write_vtk

RHO = 1060.0
MMHG = 133.322
# (C, R, Q_peak) of the three branches, src/usr_func.f90:786-850
BRANCH = {1: (1.5496e-6, 0.9758e6, 0.0004), 2: (1.5e-6, 1.12e6, 0.0004 * 0.85), 3: (0.059e-6, 19.68e6, 0.0004 * 0.05)}
TS = 0.4


def to_mmhg(p_nd, U):
    return p_nd * RHO * U * U / MMHG


@pytest.mark.parametrize("branch", [1, 2, 3])
def test_systole_solves_the_windkessel_ode_up_to_the_stated_sine_term(branch):
    C, R, Q = BRANCH[branch]
    ref = np.array([1.0, 1.0, 1.0])  # reference time 1 s: time is physiological time
    ts = np.linspace(0.0, TS, 41)
    sol = solve_ivp(lambda t, p: (Q * np.sin(np.pi * t / TS) - p / R) / C, (0.0, TS), [80.0], t_eval=ts,
                    rtol=1e-12, atol=1e-10, method="DOP853")
    D = TS ** 2 + C ** 2 * np.pi ** 2 * R ** 2
    defect = TS * np.sin(np.pi * ts / TS) * (1.0 - TS * Q * R) / D
    got = np.array([to_mmhg(O.wk_pressure_flow(branch, t, ref, np.zeros(3))[0], 1.0) for t in ts])
    assert abs(got[0] - 80.0) < 1e-9
    np.testing.assert_allclose(got, sol.y[0] + defect, rtol=0, atol=1e-7 * 80.0)
    # the flow is the half sine in systole, nondimensionalised by L^3/T (= 1 here)
    q = np.array([O.wk_pressure_flow(branch, t, ref, np.zeros(3))[1] for t in ts])
    np.testing.assert_allclose(q, Q * np.sin(np.pi * ts / TS), rtol=1e-14, atol=1e-20)


@pytest.mark.parametrize("branch", [1, 2, 3])
def test_diastole_decays_from_the_end_systolic_pressure(branch):
    C, R, _ = BRANCH[branch]
    ref = np.array([2.7, 0.2, 0.2 / 2.7])
    state = np.zeros(3)
    t_end = TS / ref[2]  # reference units
    p_end, _ = O.wk_pressure_flow(branch, t_end, ref, state)
    assert state[branch - 1] == pytest.approx(to_mmhg(p_end, 2.7), rel=1e-14)
    # diastole reads the stored end-systolic pressure, whatever time it was stored at
    for dt_phys in (1e-3, 0.05, 0.3):
        p, q = O.wk_pressure_flow(branch, (TS + dt_phys) / ref[2], ref, state)
        assert q == 0.0
        assert p / p_end == pytest.approx(np.exp(-dt_phys / (R * C)), rel=1e-12)
    state2 = state.copy()
    state2[branch - 1] = 100.0
    p, _ = O.wk_pressure_flow(branch, 0.7 / ref[2], ref, state2)
    assert to_mmhg(p, 2.7) == pytest.approx(100.0 * np.exp(-(0.7 - TS) / (R * C)), rel=1e-12)


def test_units_of_the_idealized_scales():
    ref = np.array([2.7, 0.2, 0.2 / 2.7])
    p, q = O.wk_pressure_flow(1, 0.0, ref, np.zeros(3))
    assert p == pytest.approx(80.0 * MMHG / (RHO * 2.7 ** 2), rel=1e-14)
    t = 0.1 / ref[2]
    _, q = O.wk_pressure_flow(1, t, ref, np.zeros(3))
    assert q == pytest.approx(0.0004 * np.sin(np.pi * 0.1 / TS) / (0.2 ** 3 / ref[2]), rel=1e-13)


def test_convert_waveform_branches():
    t, q = read_waveform("waveform_idealized.csv")
    tc, qc, ref = O.waveform_convert_branch(t, q, 65, 0)
    tc0, qc0, ref0 = O.waveform_convert(t, q, 65)
    assert np.array_equal(tc, tc0) and np.array_equal(qc, qc0) and np.array_equal(ref, ref0)
    tc, qc, ref = O.waveform_convert_branch(t, q, 65, 1)
    assert list(ref) == [2.7, 0.2, 0.2 / 2.7]
    np.testing.assert_allclose(tc, t / 1000.0 / (0.2 / 2.7), rtol=1e-15)
    np.testing.assert_allclose(qc, q / 2.7, rtol=1e-15)


def test_echo_windkessel_is_one_euler_step_per_holding_interval():
    """get_pressure_waveform_echo_WKM: from 80 mmHg, one explicit step of
    C dp/dt = -p/R - q over each waveform interval that holds the pulse time
    (two at an inner breakpoint, applied in order), q the interval's start
    velocity scaled to the peak flow; piecewise constant in time."""
    t, q = read_waveform("waveform_tavi0.csv")
    tc, qc, ref = O.waveform_convert_branch(t, q, 65, 0)
    U = ref[0]

    def step(p, i, R, C, qmax):
        dt = (t[i + 1] - t[i]) / 1000.0
        return p * (1.0 - dt / (C * R)) - dt / C * (q[i] * qmax / q.max())

    models = [(0.78e6, 2.166e-6, 0.0004), (0.898e6, 2.066e-6, 0.0004 * 0.85)]
    for i in (0, 7, 30, len(t) - 2):
        tm = 0.5 * (tc[i] + tc[i + 1])
        got = O.wk_echo(t, q, tc, tm, U)
        for side, (R, C, qmax) in enumerate(models):
            assert to_mmhg(got[side], U) == pytest.approx(step(80.0, i, R, C, qmax), rel=1e-13)
        # the same value anywhere inside the interval, and one pulse later
        assert O.wk_echo(t, q, tc, tc[i] + 0.3 * (tc[i + 1] - tc[i]), U) == got
    i = 12  # on an inner breakpoint both neighbouring intervals apply
    got = O.wk_echo(t, q, tc, tc[i], U)
    for side, (R, C, qmax) in enumerate(models):
        assert to_mmhg(got[side], U) == pytest.approx(step(step(80.0, i - 1, R, C, qmax), i, R, C, qmax), rel=1e-13)


def _ct_case(tmp_path, grid=GRID):
    path = tmp_path / "labels.vtk"
    write_vtk(path, label_image())
    spec = (grid, (1, 1, 1, 1, 1, 1), LENGTHS, (0.0, 1.2, 0.0))
    oc = O.OracleCase(*spec)
    t, q = read_waveform("waveform_idealized.csv")
    tc, qc, ref = O.waveform_convert_branch(t, q, grid[0], 1)
    dims, vals = O.vtk_read(path)
    lab, info, _ = O.ct_build(oc, dims, vals, OFFSETS, ref[0], ref[1])
    return oc, lab, info, (t, q, tc, qc, ref)


def test_pressure_bc_writes_the_labelled_points_of_one_plane(tmp_path):
    oc, lab, info, _ = _ct_case(tmp_path)
    ks = info[4]
    rel = np.zeros(oc.shape)
    O.set_pressure_bc(oc, lab, (1, 1, 1), ks, 1.5, -2.5, 7.0, rel)
    plane = ks - 4  # Fortran index; array offset plane + 3
    assert ks - 10 <= plane <= ks - 1  # inside the 10-plane inlet / outlet zone
    n5 = int((lab[plane + 3] == 5).sum())
    n9 = int((lab[plane + 3] == 9).sum())
    n11 = int((lab == 11).sum())
    assert n5 > 0 and n9 > 0 and n11 > 0
    assert int((rel == 1.5).sum()) == n5 and np.all(rel[plane + 3][lab[plane + 3] == 5] == 1.5)
    assert int((rel == -2.5).sum()) == n9
    assert int((rel == 7.0).sum()) == n11
    assert int((rel != 0.0).sum()) == n5 + n9 + n11
    # type flags off: nothing written
    rel2 = np.zeros(oc.shape)
    O.set_pressure_bc(oc, lab, (0, 0, 0), ks, 1.5, -2.5, 7.0, rel2)
    assert not rel2.any()


@pytest.mark.parametrize("grid,covered", [((33, 25, 33), (32, 24, 32)), ((22, 25, 33), (20, 24, 32)),
                                          ((16, 16, 16), None)])
def test_pressure_bc_covers_what_the_launch_reaches(grid, covered):
    """blocks (N-1)//8 and threads (N-1)//blocks per direction: the grid reaches
    1..blocks*threads; with a block over 1024 threads (15^3) nothing runs."""
    oc = O.OracleCase(grid, (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    lab = np.full(oc.shape, 11, dtype=np.uint8)
    rel = np.zeros(oc.shape)
    O.set_pressure_bc(oc, lab, (0, 0, 1), 15, 0.0, 0.0, 1.0, rel)
    if covered is None:
        assert not rel.any()
        return
    want = np.zeros(oc.shape)
    want[4:covered[2] + 4, 4:covered[1] + 4, 4:covered[0] + 4] = 1.0
    assert np.array_equal(rel, want)


def _jacobi_by_hand(oc, omega, ss, nn, rel, bb, max_iter=4000):
    """the stopping rule driven from Python over the oracle's single sweeps"""
    M = oc.N
    thresh = 20.0 * 20 * 8.0 * 1.0 / 3.0 * 0.8 * 3.1415 * 3.1415 * sum(1.0 / (m * m) for m in M)
    comp = oc.zeros()
    rel_old = oc.zeros()
    err_old = 0.0
    sweeps = 0
    sl = tuple(slice(ss[d] + 3, nn[d] + 4) for d in (2, 1, 0))
    for it in range(1, max_iter + 1):
        if it % 20 == 1:
            err = float(np.sum(np.abs(rel[sl] - rel_old[sl])))
            with np.errstate(divide="ignore"):
                if it > 1 and (err == 0.0 or abs(np.log(err_old / err)) <= thresh):
                    break
            rel_old[...] = rel
            err_old = err
        oc.jor(1, omega, ss, nn, rel, bb, comp)
        sweeps += 1
    return sweeps


@pytest.mark.parametrize("mode", [0, 1])
def test_device_jacobi_solver_stopping_rule(mode):
    spec = ((17, 13, 11), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 1.2, 0.0))
    oc = O.OracleCase(*spec)
    oc.set_reduction_mode(mode)
    rng = np.random.default_rng(3)
    bb = np.ascontiguousarray(rng.standard_normal(oc.shape))
    rel0 = np.ascontiguousarray(0.1 * rng.standard_normal(oc.shape))
    ss, nn = oc.Sp, oc.Np
    rel = rel0.copy()
    rc, sweeps = O.device_jacobi_solver(oc, 0.8, ss, nn, rel, bb)
    assert rc == 0 and sweeps % 20 == 0 and sweeps > 0
    by_hand = rel0.copy()
    assert _jacobi_by_hand(oc, 0.8, ss, nn, by_hand, bb) == sweeps
    sl = tuple(slice(ss[d] + 3, nn[d] + 4) for d in (2, 1, 0))
    assert np.array_equal(rel[sl], by_hand[sl])


def test_device_jacobi_solver_order_of_the_sum_does_not_move_the_stop():
    spec = ((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    counts = []
    for mode in (0, 1):
        oc = O.OracleCase(*spec)
        oc.set_reduction_mode(mode)
        rng = np.random.default_rng(11)
        bb = np.ascontiguousarray(rng.standard_normal(oc.shape))
        rel = oc.zeros()
        counts.append(O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, rel, bb)[1])
    assert counts[0] == counts[1]


def test_device_jacobi_solver_fixed_point_and_limit():
    oc = O.OracleCase((9, 9, 9), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    rel, bb = oc.zeros(), oc.zeros()
    # the first test only records; err_tot 0 at the second (a fixed point) stops
    assert O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, rel, bb) == (0, 20)
    rng = np.random.default_rng(2)
    bb = np.ascontiguousarray(rng.standard_normal(oc.shape))
    rc, sweeps = O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, rel, bb, max_iter=7)
    assert (rc, sweeps) == (1, 7)


def test_device_jacobi_solver_pressure_on_substep_3(tmp_path):
    """substep 3 writes the scaled pressures into the labelled plane before every
    sweep and before the exit, so the result holds them there; other substeps
    leave the plane to the relaxation."""
    oc, lab, info, _ = _ct_case(tmp_path)
    ks = info[4]
    rng = np.random.default_rng(4)
    bb = np.ascontiguousarray(rng.standard_normal(oc.shape))
    pres = np.array([3.0, -1.0, 5.0])
    dt = 0.01
    f = (3.0 / 4.0 + -5.0 / 12.0)
    rel3 = oc.zeros()
    rc, sw = O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, rel3, bb, lab, ks, 3, dt, pres, max_iter=400)
    plane = ks - 4
    m5 = lab[plane + 3] == 5
    m9 = lab[plane + 3] == 9
    assert np.all(rel3[plane + 3][m5] == pres[0] * f * dt)
    assert np.all(rel3[plane + 3][m9] == pres[1] * f * dt)
    rel1 = oc.zeros()
    O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, rel1, bb, lab, ks, 1, dt, pres, max_iter=400)
    assert not np.any(rel1[plane + 3][m5] == pres[0] * f * dt)
