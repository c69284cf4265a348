"""GPU parity of the CT geometry path (read_vtk_3D, convert_waveform and the CT
branch of apply_CT_forcing; src/usr_func.f90:187-651, 909-945): the label
volume byte for byte, the detection results and reference scales exactly, and
the forced velocity bit for bit against the oracle at several times (inside a
pulse, on a waveform breakpoint, in a later pulse)."""
import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from ct_image import DIMS, GRID, LENGTHS, OFFSETS, label_image, read_waveform, write_vtk
from test_gpu_parity import same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wave", ["waveform_idealized.csv", "waveform_tavi0.csv"])
def test_ct_labels_and_forcing(tmp_path, wave):
    path = tmp_path / "labels.vtk"
    write_vtk(path, label_image())
    t, q = read_waveform(wave)
    spec = (GRID, (1, 1, 1, 1, 1, 1), LENGTHS, (0.0, 1.2, 0.0))
    oc = O.OracleCase(*spec)
    gc = ih.Case(*spec)
    info = gc.ct_load_vtk(path, OFFSETS, t, q)
    dims, vals = O.vtk_read(path)
    tc, qc, ref = O.waveform_convert(t, q, GRID[0])
    lab, oinfo, centre = O.ct_build(oc, dims, vals, OFFSETS, ref[0], ref[1])
    assert [info.inlet_cells, info.outlet_cells, info.outlet2_cells, info.inlet_jet_voxels, info.k_start,
            info.k_start2] == oinfo
    assert list(info.inlet_centre) + list(info.outlet_centre) == list(centre)
    assert [info.reference_velocity, info.reference_length, info.reference_time] == list(ref)
    assert np.array_equal(gc.ct_labels(), lab)
    vel = np.ascontiguousarray(np.random.default_rng(5).standard_normal((3,) + oc.shape))
    for q_ in range(3):
        gc.upload(ih.VEL1 + q_, vel[q_])
    for time in (0.37 * tc[-1], tc[5], 2.6 * tc[-1]):
        v, found = gc.ct_inflow(time)
        assert (v, found) == O.ct_inflow(tc, qc, time)
        O.ct_forcing(oc, lab, tc, qc, time, vel)
        gc.ct_forcing(time)
        for q_ in range(3):
            same(gc.download(ih.VEL1 + q_), vel[q_], f"CT forcing vel{q_ + 1} at t={time}")


def test_steps_with_ct_forcing(tmp_path):
    """Two projection steps on the CT labels: the forcing at each step's end
    time (after the time update, as src/mod_timeint.f90:405-416 orders it),
    bit-exact with the oracle."""
    import ctypes
    from test_gpu_parity import compare_state
    path = tmp_path / "labels.vtk"
    write_vtk(path, label_image())
    t, q = read_waveform("waveform_idealized.csv")
    spec = (GRID, (1, 1, 1, 1, 1, 1), LENGTHS, (0.0, 1.2, 0.0))
    oc = O.OracleCase(*spec)
    gc = ih.Case(*spec)
    gc.ct_load_vtk(path, OFFSETS, t, q)
    dims, vals = O.vtk_read(path)
    tc, qc, ref = O.waveform_convert(t, q, GRID[0])
    lab, _, _ = O.ct_build(oc, dims, vals, OFFSETS, ref[0], ref[1])
    vel = np.ascontiguousarray(0.01 * np.random.default_rng(6).standard_normal((3,) + oc.shape))
    pre, nl = oc.zeros(), oc.zeros(3)
    for q_ in range(3):
        gc.upload(ih.VEL1 + q_, vel[q_])
        gc.upload(ih.NL1 + q_, nl[q_])
    gc.upload(ih.PRE, pre)
    dt = 0.05 * tc[-1]
    kw = dict(Re=100.0, dtime=dt, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=3000,
              level_pressure=1, inflow_kind=0, inflow_amp=0.0)
    po = O.StepParams(ct_lab=lab.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ct_npts=len(tc),
                      ct_t=tc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      ct_q=qc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), **kw)
    pg = ih.StepParams(use_ct=1, **kw)
    time = 0.0
    for step in range(2):
        so = oc.timestep(po, step, time, vel, pre, nl)
        sg = gc.timestep(pg, time)
        time += dt
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        compare_state(oc, gc, vel, pre, nl, f"CT step {step}")
    # the jet carries the waveform's inflow
    v, found = O.ct_inflow(tc, qc, time)
    assert found and np.all(vel[2][lab == 5] == v)


def _loaded(tmp_path, ct_based, idealized, wave="waveform_idealized.csv"):
    path = tmp_path / "labels.vtk"
    write_vtk(path, label_image())
    t, q = read_waveform(wave)
    spec = (GRID, (1, 1, 1, 1, 1, 1), LENGTHS, (0.0, 1.2, 0.0))
    oc = O.OracleCase(*spec)
    gc = ih.Case(*spec)
    gc.ct_configure(ct_based, idealized)
    info = gc.ct_load_vtk(path, OFFSETS, t, q)
    dims, vals = O.vtk_read(path)
    tc, qc, ref = O.waveform_convert_branch(t, q, GRID[0], idealized)
    lab, oinfo, _ = O.ct_build(oc, dims, vals, OFFSETS, ref[0], ref[1])
    assert [info.inlet_cells, info.outlet_cells, info.outlet2_cells, info.inlet_jet_voxels, info.k_start,
            info.k_start2] == oinfo
    assert [info.reference_velocity, info.reference_length, info.reference_time] == list(ref)
    assert np.array_equal(gc.ct_labels(), lab)
    st = O.ct_state(ct_based, idealized, t, q, tc, qc, ref, oinfo[3], oinfo[4])
    return oc, gc, lab, st, tc


@pytest.mark.parametrize("branches", [(1, 0), (0, 1), (1, 1)])
def test_forcing_branches_and_windkessel_state(tmp_path, branches):
    """Both branches of apply_CT_forcing (src/usr_func.f90:909-1048): forced
    velocity bit for bit, and the pressures and end-systolic state the
    Windkessel models leave, over systole, diastole and a later pulse."""
    oc, gc, lab, st, tc = _loaded(tmp_path, *branches)
    vel = np.ascontiguousarray(np.random.default_rng(7).standard_normal((3,) + oc.shape))
    for q_ in range(3):
        gc.upload(ih.VEL1 + q_, vel[q_])
    T = st.ref[2]
    for time in (0.3 * tc[-1], 0.1 / T, 0.4 / T, 0.55 / T, 2.6 * tc[-1]):
        O.ct_apply_forcing(oc, lab, st, time, vel)
        gc.ct_forcing(time)
        for q_ in range(3):
            same(gc.download(ih.VEL1 + q_), vel[q_], f"forcing {branches} vel{q_ + 1} at t={time}")
        pres, wk = gc.ct_state()
        assert list(pres) == list(st.pres), (branches, time)
        assert list(wk) == list(st.wk_state), (branches, time)
    if branches[1]:
        assert st.wk_state[0] != 0.0 and st.wk_state[1] != 0.0 and st.wk_state[2] != 0.0


def test_set_pressure_bc_kernel(tmp_path):
    oc, gc, lab, st, _ = _loaded(tmp_path, 0, 1)
    rel = np.ascontiguousarray(np.random.default_rng(8).standard_normal(oc.shape))
    for types in ((1, 1, 0), (1, 1, 1), (0, 0, 1)):
        gc.upload(ih.PRE, rel)
        want = rel.copy()
        O.set_pressure_bc(oc, lab, types, st.k_start, 0.25, -0.5, 1.75, want)
        gc.set_pressure_bc(ih.PRE, types, 0.25, -0.5, 1.75)
        same(gc.download(ih.PRE), want, f"set_pressure_BC {types}")
        assert int((want != rel).sum()) > 0


@pytest.mark.parametrize("substep", [1, 3])
def test_device_jacobi_solver_with_pressure_bc(tmp_path, substep):
    """device_jacobi_solver with the substep-3 Dirichlet pressure of the last
    forcing (src/jacobi_solver_host_v4.f90:302-305) and the 20-sweep stopping
    test: the result bit for bit against the oracle (fixed-tree err_tot)."""
    oc, gc, lab, st, _ = _loaded(tmp_path, 0, 1)
    vel = np.ascontiguousarray(np.zeros((3,) + oc.shape))
    O.ct_apply_forcing(oc, lab, st, 0.2 / st.ref[2], vel)
    gc.ct_forcing(0.2 / st.ref[2])
    pres, _ = gc.ct_state()
    assert list(pres) == list(st.pres) and pres[0] != 0.0
    dt = 0.01
    gc.set_substep(substep, dt)
    rng = np.random.default_rng(9)
    bb = np.ascontiguousarray(rng.standard_normal(oc.shape))
    rel = np.ascontiguousarray(0.1 * rng.standard_normal(oc.shape))
    want = rel.copy()
    rc, sweeps = O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, want, bb, lab, st.k_start, substep, dt, st.pres)
    assert rc == 0 and sweeps > 0
    got = rel.copy()
    assert gc.device_jacobi_solver(0.8, oc.Sp, oc.Np, got, bb.copy()) == 0
    s = tuple(slice(oc.Sp[d] + 3, oc.Np[d] + 4) for d in (2, 1, 0))
    same(got[s], want[s], f"device_jacobi_solver substep {substep} ({sweeps} sweeps)")


def test_steps_idealized_aorta_with_pressure_bc(tmp_path):
    """Two projection steps of the idealized-aorta branch (the reference's
    default): Windkessel inflow at the jet, the Dirichlet pressure of the
    third substep's Jacobi sweeps from the previous forcing, bit-exact."""
    import ctypes
    from test_gpu_parity import compare_state
    oc, gc, lab, st, tc = _loaded(tmp_path, 0, 1)
    vel = np.ascontiguousarray(0.01 * np.random.default_rng(10).standard_normal((3,) + oc.shape))
    pre, nl = oc.zeros(), oc.zeros(3)
    for q_ in range(3):
        gc.upload(ih.VEL1 + q_, vel[q_])
        gc.upload(ih.NL1 + q_, nl[q_])
    gc.upload(ih.PRE, pre)
    dt = 0.02 / st.ref[2]
    kw = dict(Re=100.0, dtime=dt, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=3000,
              level_pressure=1, inflow_kind=0, inflow_amp=0.0)
    po = O.StepParams(ct_lab=lab.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ct_npts=len(tc),
                      ct_state=ctypes.pointer(st), pressure_bc=1, **kw)
    pg = ih.StepParams(use_ct=1, pressure_bc=1, **kw)
    time = 0.0
    for step in range(3):
        so = oc.timestep(po, step, time, vel, pre, nl)
        sg = gc.timestep(pg, time)
        time += dt
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        compare_state(oc, gc, vel, pre, nl, f"idealized step {step}")
        pres, wk = gc.ct_state()
        assert list(pres) == list(st.pres) and list(wk) == list(st.wk_state)
    assert st.pres[0] != 0.0
