"""Pins of the oracle's CT geometry path: read_vtk_3D, convert_waveform and
the CT branch of apply_CT_forcing (src/usr_func.f90:187-651, 909-945).

The reference's images are not in its tree; the waveforms are
(preprocessing/*/waveform.csv, copied to tests/golden/).  Each step is checked
against an independent statement of what it must produce: the image read back
value for value, the placement, the inlet / outlet / branch cells counted from
the disks' geometry, their centres and jet cells from numpy, the reference
scales from their physical definitions, the inflow lookup by searchsorted, and
the forcing cell by cell."""
import numpy as np
import pytest

import oracle_lib as O
from ct_image import DIMS, GRID, LENGTHS, OFFSETS, label_image, read_waveform, write_vtk


@pytest.fixture(scope="module")
def case_and_image(tmp_path_factory):
    c = O.OracleCase(GRID, (1, 1, 1, 1, 1, 1), LENGTHS, (0.0, 1.2, 0.0))
    lab = label_image()
    path = tmp_path_factory.mktemp("ct") / "labels.vtk"
    write_vtk(path, lab)
    return c, lab, path


def test_vtk_round_trip(case_and_image):
    c, lab, path = case_and_image
    dims, vals = O.vtk_read(path)
    assert dims == DIMS
    assert np.array_equal(vals, lab.transpose(2, 1, 0).reshape(-1).astype(float))


@pytest.mark.parametrize("name", ["waveform_idealized.csv", "waveform_tavi0.csv"])
def test_waveform_reference_scales(name):
    t, q = read_waveform(name)
    tc, qc, ref = O.waveform_convert(t, q, GRID[0])
    assert ref[0] == q.max()                          # peak inflow is 1 in reference units
    assert ref[1] == pytest.approx(0.0004 * GRID[0], rel=1e-15)  # CT pixel 0.4 mm is 1/M1
    assert ref[2] == pytest.approx(ref[1] / ref[0], rel=1e-15)
    assert qc.max() == 1.0
    np.testing.assert_allclose(tc * ref[2] * 1000.0, t, rtol=1e-14)  # back to milliseconds


def test_inflow_lookup_is_the_last_holding_interval():
    t, q = read_waveform("waveform_idealized.csv")
    tc, qc, _ = O.waveform_convert(t, q, GRID[0])
    pulse = tc[-1]
    rng = np.random.default_rng(3)
    for s in rng.uniform(0.0, 3.0 * pulse, 200):
        v, found = O.ct_inflow(tc, qc, s)
        shifted = s - np.floor(s / pulse) * pulse
        i = np.searchsorted(tc, shifted, side="right") - 1  # last t_i <= shifted
        if i >= len(tc) - 1:
            i = len(tc) - 2
        if shifted < tc[0]:
            assert not found
        else:
            assert found and v == qc[i]


def test_detection_counts_centres_and_jet(case_and_image):
    c, lab, path = case_and_image
    dims, vals = O.vtk_read(path)
    t, q = read_waveform("waveform_idealized.csv")
    _, _, ref = O.waveform_convert(t, q, GRID[0])
    L, info, centre = O.ct_build(c, dims, vals, OFFSETS, ref[0], ref[1])
    xo, yo, zo = OFFSETS
    # the image lands transposed: grid (j + y_off, i + x_off, k + z_off)
    grid = np.full((GRID[2] + 7, GRID[1] + 7, GRID[0] + 7), 2)
    for k in range(DIMS[2]):
        grid[k + 1 + zo + 3, xo + 4:xo + 4 + DIMS[0], yo + 4:yo + 4 + DIMS[1]] = lab[:, :, k]
    ks, ks2 = zo + 1, zo + DIMS[2]
    assert info[4:] == [ks, ks2]
    plane, last = grid[ks + 3, 4:GRID[1] + 4, 4:GRID[0] + 4], grid[ks2 + 3, 4:GRID[1] + 4, 4:GRID[0] + 4]
    half = GRID[0] // 2
    xs, ys = c.coords(0, 0)[4:GRID[0] + 4], c.coords(1, 0)[4:GRID[1] + 4]
    inlet = np.argwhere(plane[:, :half] == 1)
    outlet = np.argwhere(plane[:, half - 1:] == 1) + [0, half - 1]
    assert info[0] == len(inlet) and info[1] == len(outlet) and info[2] == int((last == 1).sum())
    np.testing.assert_allclose(centre[:2], [xs[inlet[:, 1]].mean(), ys[inlet[:, 0]].mean()], rtol=1e-14)
    np.testing.assert_allclose(centre[2:], [xs[outlet[:, 1]].mean(), ys[outlet[:, 0]].mean()], rtol=1e-14)
    R = np.sqrt(0.0004 / (np.pi * ref[0])) / ref[1]
    r = np.hypot(xs[inlet[:, 1]] - centre[0], ys[inlet[:, 0]] - centre[1])
    assert 0 < info[3] == int((r <= R).sum()) < len(inlet)  # the jet is part of the inlet
    # labels below the first plane: 5 (jet), 7 (rest of the inlet), 9 (outlet), 11 above the last
    lab3 = L[:, 3:GRID[1] + 4, 3:GRID[0] + 4]  # z, y, x with index 0 = coordinate 0
    for kk in range(ks - 10, ks):
        got = lab3[kk + 3]
        assert int((got == 5).sum()) == info[3]
        assert int((got == 9).sum()) == len(outlet)
        assert int(((got == 5) | (got == 7) | (got == 9)).sum()) == len(np.unique(np.vstack([inlet, outlet]), axis=0))
    for kk in range(ks2 + 1, ks2 + 11):
        assert int((lab3[kk + 3] == 11).sum()) == info[2]


def test_forcing_cell_by_cell(case_and_image):
    c, lab, path = case_and_image
    dims, vals = O.vtk_read(path)
    t, q = read_waveform("waveform_idealized.csv")
    tc, qc, ref = O.waveform_convert(t, q, GRID[0])
    L, info, centre = O.ct_build(c, dims, vals, OFFSETS, ref[0], ref[1])
    rng = np.random.default_rng(4)
    vel = np.ascontiguousarray(rng.standard_normal((3,) + c.shape))
    time = 0.37 * tc[-1]
    before = vel.copy()
    O.ct_forcing(c, L, tc, qc, time, vel)
    v, found = O.ct_inflow(tc, qc, time)
    assert found
    s = tuple(slice(c.SB[0][d] + 3, c.NB[0][d] + 4) for d in (2, 1, 0))
    Le = np.full(L.shape, 2, dtype=L.dtype)  # positions outside 1..N read label 2
    Le[4:GRID[2] + 4, 4:GRID[1] + 4, 4:GRID[0] + 4] = L[4:GRID[2] + 4, 4:GRID[1] + 4, 4:GRID[0] + 4]
    lb = Le[s]
    zero = (lb == 0) | (lb == 7)
    for m in range(3):
        assert np.all(vel[m][s][zero] == 0.0)
    assert np.all(vel[2][s][lb == 5] == v) and np.all(vel[0][s][lb == 5] == 0.0)
    keep = ~zero & (lb != 5)
    for m in range(3):
        assert np.array_equal(vel[m][s][keep], before[m][s][keep])
