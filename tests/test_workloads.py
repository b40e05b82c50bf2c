"""The synthetic workloads of the benchmark lines: grids, potentials and
initial states as the reference's inputs define them (gen_psi_0.cpp
normalisation, misc.cpp rect, rt_solver.cpp potentials)."""
import numpy as np

from atus2_amd.workloads import (BraggWorkload, DoubleSlitWorkload, HarmonicLineWorkload, axis_coords,
                                 particle_number)


def test_axis_coords_follow_the_header_convention():
    # x_i = (i - n/2) dx, dx = L/n: the grid holds -L/2 and not +L/2
    x = axis_coords(8, 4.0)
    assert np.array_equal(x, np.array([-2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5]))


def test_harmonic_line_workload():
    wl = HarmonicLineWorkload()
    dx, dk = wl.spacings
    assert wl.shape == (4096,) and dx == [80.0 / 4096] and np.isclose(dk[0], 2 * np.pi / 80.0)
    psi = wl.initial_state()
    assert abs(particle_number(psi, dx) - 1.0) < 1e-14
    # the seeded noise: the same state on every call, 1e-3 around the
    # oscillator ground state
    assert np.array_equal(psi, wl.initial_state())
    x = axis_coords(4096, 80.0)
    ground = np.pi ** -0.25 * np.exp(-0.5 * x ** 2)
    assert 1e-4 < np.max(np.abs(psi - ground)) < 1e-2
    assert np.array_equal(wl.potential(), 0.5 * x ** 2)


def test_double_slit_wall_is_half_open():
    # rect(x, xl, xr) is 1 on [xl, xr) (misc.cpp:40-53): with dx = 1/16 the
    # grid holds x = -0.5 and x = +0.5, so the wall covers 16 columns, from
    # x = -0.5 up to x = 0.5 - dx
    wl = DoubleSlitWorkload()
    x = axis_coords(wl.n, wl.L)
    V = wl.potential()
    assert V.shape == (2048, 2048)
    cols = np.nonzero(V.max(axis=1) > 0)[0]
    assert x[cols[0]] == -0.5 and x[cols[-1]] == 0.5 - 1.0 / 16 and len(cols) == 16
    # the slits: rows with y in [-5, -3) and [3, 5) are open in the wall
    wall = V[cols[0]]
    open_rows = x[wall == 0.0]
    assert np.array_equal(open_rows, x[((x >= -5) & (x < -3)) | ((x >= 3) & (x < 5))])
    assert set(np.unique(V)) == {0.0, 1e3}


def test_double_slit_initial_state():
    wl = DoubleSlitWorkload(n=256, L=128.0)
    psi = wl.initial_state()
    dx, _ = wl.spacings
    assert abs(particle_number(psi, dx) - 1.0) < 1e-13
    x = axis_coords(256, 128.0)
    i, j = np.unravel_index(np.argmax(np.abs(psi)), psi.shape)
    assert x[i] == -20.0 and x[j] == 0.0
    assert np.all(psi.imag == 0.0)   # the momentum is set on the device (atus_set_momentum)


def test_bragg_workload_is_normalised_and_empty_in_the_second_component():
    wl = BraggWorkload.proxy(32)
    psi1, psi2 = wl.initial_state()
    dx, _ = wl.spacings
    assert abs(particle_number(psi1, dx) - 1.0) < 1e-13
    assert not psi2.any()
    # the proxy keeps the grid spacing of the 256^3 case
    assert np.allclose(dx, [0.25] * 3)
