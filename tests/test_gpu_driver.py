"""The ePIC++ driver (epic2d_run: reference main loop over the C++ host
classes) on an input.ini in the reference's format: its kinetic/potential
energy history matches the oracle's replay of the same loop from the same
mt19937(0) load within 1e-12 relative."""
import os
import subprocess

import numpy as np
import pytest

from harness import small_setup
from oracle_api import OracleSim

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "epic2d_mcc_amd", "epic2d_run")

INI = """[file]
output = {out}
[time]
NUM_TS = 10
DT_coeff = 0.05
[diagnostics]
write_interval = 1
write_interval_phase = 100
write_diagnostics = 5
write_flag = 0
save_fig = 0
sub_cycle_interval = 1
precision = 6
[PlotFlags]
phase_space = 0
[domain]
Nx = 16
Ny = 16
x0 = 0
y0 = 0
[Grids]
grid_count = 0
[Emitters]
count = 0
[normalization]
norm_scheme = 1
vel_norm_scheme = 1
lenght_scale = 0.005
time_scale = omegape
[solver]
solvertype = cg
tolerance = 1e-10
max_iteration = 10000
[simulation]
shapefunction = CIC
push_parallal = false
deposit_parallal = false
density = 1E13
bc = pbc
ionfixed = 0
[collision]
elastic = false
excitation = false
ionization = false
GAS_DENSITY = 0
[species]
electron,9.10938215E-31,2000,1,-1,1,0,0,random
ion,1.6724823e-27,2000,0.026,1,0,0,0,random
"""


def test_driver_energy_history_matches_oracle(tmp_path):
    out = tmp_path / "data"
    ini = tmp_path / "input.ini"
    ini.write_text(INI.format(out=str(out)))
    r = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    hist = np.loadtxt(out / "kinetic_energy.csv", delimiter=",")
    assert hist.shape == (11, 8)
    # oracle: same setup (main.cpp normalisation), mt19937(0) load, same loop
    st = small_setup(N=2000, Nx=16, Ny=16, tol=1e-10, ion_mass=1.6724823e-27)
    o = OracleSim(st.domain, st.species)
    for s in range(2):
        o.init_species(s, 0)
    o.initial_field()
    for ts in range(11):
        o.step()
        row = [ts * st.domain.DT]
        for s in range(2):
            row += list(o.ke(s, 0))
        row.append(o.pe(0))
        row = np.array(row)
        assert np.allclose(hist[ts], row, rtol=1e-12, atol=1e-300), (ts, hist[ts], row)


def test_driver_rejects_missing_key(tmp_path):
    ini = tmp_path / "input.ini"
    ini.write_text(INI.format(out=str(tmp_path / "d")).replace("max_iteration = 10000\n", ""))
    r = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "epic2d_run" in r.stderr


GRIDS = """[Grids]
grid_count = 2
grid_0.type = reactconduct
grid_0.min_x = 3
grid_0.max_x = 6
grid_0.min_y = 9
grid_0.max_y = 13
grid_0.grid_voltage = 2
grid_1.type = circular
grid_1.x_center = 10
grid_1.y_center = 6
grid_1.electrode_number = 4
grid_1.grid_radius = 3
grid_1.electrod_radius = 1.5
grid_1.electrode_voltage = -1
"""


def test_driver_with_conductors_matches_oracle(tmp_path):
    """[Grids] as main.cpp:287-323 reads it: the loader avoids the rectangle,
    the push absorbs on it (the circular grid has no electrodes under cg),
    and the energy history follows the oracle's replay."""
    from epic2d_mcc_amd import _lib
    out = tmp_path / "data"
    ini = tmp_path / "input.ini"
    text = INI.format(out=str(out)).replace("[Grids]\ngrid_count = 0\n", GRIDS).replace("bc = pbc", "bc = open")
    ini.write_text(text)
    r = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    hist = np.loadtxt(out / "kinetic_energy.csv", delimiter=",")
    st = small_setup(N=2000, Nx=16, Ny=16, bc="open", tol=1e-10, ion_mass=1.6724823e-27)
    o = OracleSim(st.domain, st.species)
    o.set_conductors(_lib.conductors([
        {"kind": _lib.COND_RECT, "min_x": 3, "max_x": 6, "min_y": 9, "max_y": 13, "voltage": 2.0},
        {"kind": _lib.COND_CIRCULAR, "x_center": 10, "y_center": 6, "electrode_number": 4, "grid_radius": 3.0,
         "electrode_radius": 1.5, "voltage": -1.0}]))
    for s in range(2):
        assert o.init_species(s, 0) == 0
    o.initial_field()
    for ts in range(11):
        o.step()
        row = [ts * st.domain.DT]
        for s in range(2):
            row += list(o.ke(s, 0))
        row.append(o.pe(0))
        assert np.allclose(hist[ts], np.array(row), rtol=1e-12, atol=1e-300), (ts, hist[ts], row)
