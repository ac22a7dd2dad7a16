"""The owt_solver CLI, the drop-in for the reference driver (src/main.cpp:7-93),
run on the reference's own case files.

The case texts below carry the key values of test-cases/Retau-180/case.toml and
test-cases/Retau-550/case.toml (nu, dt, dkx, dkz, NX, NY, NZ, alpha_mesh,
initial_condition, PX, PZ, save_interval), with NT cut to a few steps so that
a run takes seconds; every other value is the reference's. The driver must
print one friction line and one timing line per step (src/solver.cpp:314,
src/main.cpp:79-82), save before stepping when id_step % save_interval == 0
(src/main.cpp:72-75), and, with initial_condition = 3 (renewal), continue from
the full-state checkpoint of an earlier run exactly as an uninterrupted run.
"""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOLVER = os.path.join(ROOT, "open-3d-wallturbulence_amd", "bin", "owt_solver")

RETAU180 = """# physics
nu = 3.571428571428571e-4 # kinetic viscosity
# CFD
dt = 2.0e-3 # time step size
dkx = 0.25 # Lx=2*pi/dkx
dkz = 0.5 # Lz=2*pi/dkx
NX = {NX} # number of cells in x direction
NY = {NY} # number of cells in y direction
NZ = {NZ} # number of cells in z direction
alpha_mesh = 2.0 # the de-sigularity size of mesh y=tanh(alpha eta)/tanh(alpha)
NT = {NT} # number of time steps
initial_condition = {IC} # 1 - "laminar only", 2 - "laminar perturbation", 3 - "renewal"
# MPI
PX = 16  # number of processors in x direction
PZ = 8 # number of processors in z direction
# IO
save_interval = {SI} # save results every $(save_interval) time steps
"""

RETAU550 = """# physics
nu = 1.0e-4 # kinetic viscosity
# CFD
dt = 1.0e-3 # time step size
dkx = 0.5 # Lx=2*pi/dkx
dkz = 1.0 # Lz=2*pi/dkx
NX = {NX} # number of cells in x direction
NY = {NY} # number of cells in y direction
NZ = {NZ} # number of cells in z direction
alpha_mesh = 2.4 # the de-sigularity size of mesh y=tanh(alpha eta)/tanh(alpha)
NT = {NT} # number of time steps
initial_condition = {IC} # 1 - "laminar only", 2 - "laminar perturbation", 3 - "renewal"
# MPI
PX = 32  # number of processors in x direction
PZ = 16 # number of processors in z direction
# IO
save_interval = {SI} # save results every $(save_interval) time steps
"""


def run_case(tmp, text, timeout=600):
    path = os.path.join(tmp, "case.toml")
    with open(path, "w") as f:
        f.write(text)
    print(f"[gpu] owt_solver {text.splitlines()[6].split('#')[0].strip()} ...", flush=True)
    r = subprocess.run([SOLVER, path], cwd=tmp, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def steps_of(out):
    fric = re.findall(r"fx_fric, ([-+0-9.eE]+), bottom, ([-+0-9.eE]+), upper, ([-+0-9.eE]+)", out)
    ids = [int(m) for m in re.findall(r"time step, (\d+), using time", out)]
    return fric, ids


def test_retau180_case_file_full_size(tmp_path):
    """test-cases/Retau-180/case.toml at its own size, 320x160x320 (480-point padded
    transforms, even NY), NT = 2: three steps, snapshots at 0, 1, 2 before stepping."""
    out = run_case(str(tmp_path), RETAU180.format(NX=320, NY=160, NZ=320, NT=2, IC=1, SI=1))
    assert "NX, NY, NZ = 320 160 320" in out and "NXH, NZH, NX2, NZ2 = 160 160 480 480" in out
    fric, ids = steps_of(out)
    assert ids == [0, 1, 2] and len(fric) == 3
    # the laminar profile's friction force is 3 nu, and the wave barely moves it
    nu = 3.571428571428571e-4
    assert all(abs(float(f[0]) - 3 * nu) < 1e-3 * 3 * nu for f in fric), fric
    for s in range(3):
        assert os.path.exists(tmp_path / "DATA" / f"F-{s:08d}.H5")
    assert os.path.exists(tmp_path / "DATA" / "restart.ckpt")


def test_renewal_continues_exactly(tmp_path):
    """initial_condition = 3 (renewal): a run of 2 steps, then a renewed run of 2
    more, ends with the same snapshot as 4 uninterrupted steps."""
    import h5read

    if h5read.lib() is None:
        pytest.skip("HDF5 C library not available to read the snapshots back")
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    shape = dict(NX=32, NY=33, NZ=32)
    run_case(str(a), RETAU180.format(**shape, NT=3, IC=1, SI=1))   # steps 0..3, snapshots 0..3
    run_case(str(b), RETAU180.format(**shape, NT=1, IC=1, SI=1))   # steps 0, 1 -> checkpoint at 2
    out = run_case(str(b), RETAU180.format(**shape, NT=1, IC=3, SI=1))  # restart at 2: steps 2, 3
    assert "restart from DATA/restart.ckpt at step 2" in out
    _, ids = steps_of(out)
    assert ids == [2, 3]
    for s in (2, 3):
        for key in ("velx_real", "vely_imag", "velz_real"):
            x = h5read.read(str(a / "DATA" / f"F-{s:08d}.H5"), key)
            y = h5read.read(str(b / "DATA" / f"F-{s:08d}.H5"), key)
            assert np.array_equal(x, y), (s, key)


def test_owt_table_selects_the_seeded_field(tmp_path, owt_mod):
    """[owt] init = "seeded" with seed and amp: the saved initial field equals
    owt_init_fields(mode 2, seed, amp) through the C ABI."""
    import h5read

    if h5read.lib() is None:
        pytest.skip("HDF5 C library not available to read the snapshots back")
    text = RETAU180.format(NX=32, NY=33, NZ=32, NT=0, IC=1, SI=1) + '\n[owt]\ninit = "seeded"\nseed = 7\namp = 0.2\n'
    run_case(str(tmp_path), text)
    p = dict(nu=3.571428571428571e-4, dt=2e-3, dkx=0.25, dkz=0.5, alpha_mesh=2.0, NX=32, NY=33, NZ=32)
    with owt_mod.Channel(params=p) as ch:
        ch.init_fields(2, seed=7, amp=0.2)
        u = ch.field("u")
    assert np.array_equal(h5read.read(str(tmp_path / "DATA" / "F-00000000.H5"), "velx_real"), u.real)


def test_run_params_of_the_case_file(tmp_path, owt_mod):
    path = tmp_path / "case.toml"
    path.write_text(RETAU550.format(NX=16, NY=288, NZ=16, NT=7, IC=3, SI=500))
    with owt_mod.Channel(str(path)) as ch:
        rp = ch.run_params()
        assert (rp["NT"], rp["save_interval"], rp["init_mode"], rp["restart"]) == (7, 500, 4, "DATA/restart.ckpt")
    path.write_text(RETAU550.format(NX=16, NY=288, NZ=16, NT=7, IC=1, SI=500) + '\n[owt]\ninit = "laminar"\n')
    with owt_mod.Channel(str(path)) as ch:
        assert ch.run_params()["init_mode"] == 3
        ch.init_fields(3)
        assert np.abs(ch.field("v")).max() == 0 and np.abs(ch.field("u")[1:]).max() == 0
