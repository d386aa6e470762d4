"""The C++ host (ISimulation / makeSimulation / run loop / plotfile /
checkpoint) driving the gfx950 kernels, against the oracle and the
reference's regression criteria."""
import os
import subprocess

import numpy as np
import pytest

from muphfasa_amd import host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INP = os.path.join(ROOT, "tests", "golden", "inputs")


def _read(name):
    return open(os.path.join(INP, name)).read()


def _oracle_fab(L, n, valid):
    f = O.new_fab(L, n)
    O.valid(L, f)[...] = valid
    return f


def test_lid_cavity_128_through_host_matches_oracle():
    """Configuration 1: D2Q9 lid cavity 128^2, Re=100, 10 000 steps."""
    sim = host.Simulation(_read("lid_cavity_128"))
    sim.initialize()
    f0 = sim.f()
    print("host: 10000 steps", flush=True)
    sim.advance(10000)
    sim.update_macroscopic(9999)
    got = sim.f()
    rho_g, vel_g = sim.macro()
    L = O.Lattice("D2Q9")
    n = [128, 128]
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    res = O.resolve_bc(2, [1, 1, 1, 0], np.ones(6), fu)
    fab = _oracle_fab(L, n, f0)
    O.apply_bc(L, fab, n, [0, 0, 0], res)
    print("oracle: 10000 steps", flush=True)
    fe = O.run(L, fab, n, sim.omega, None, [0, 0], res, 10000)
    assert np.array_equal(got, O.valid(L, fe))
    rho_e, vel_e = O.moments(L, fe, n)
    assert np.array_equal(rho_g, rho_e) and np.array_equal(vel_g, vel_e)
    assert abs(sim.time - 10000.0) == 0.0


def test_periodic_noise_init_and_steps_match_oracle():
    """Configuration 2 keys at 32^3: init (noise + Guo shift) then 100 steps."""
    text = _read("periodic_d3q19_256").replace("256 256 256", "32 32 32")
    sim = host.Simulation(text)
    sim.initialize()
    f0 = sim.f()
    L = O.Lattice("D3Q19")
    n = [32, 32, 32]
    # the initial populations are feq(rho, u - F/2rho) of the seeded noise
    rho0, vel0 = O.moments(L, _oracle_fab(L, n, f0), n, sim.force, streamed=False)
    fe0, _, _ = O.init_equilibrium(L, np.ones((32, 32, 32)), vel0 - 0.5 * np.array(sim.force)[:, None, None, None],
                                   n, None)
    assert np.allclose(O.valid(L, fe0), f0, rtol=0, atol=1e-15)
    assert np.abs(vel0).max() <= 1e-3 + 1e-12 and np.abs(vel0).max() > 9e-4
    sim.advance(100)
    res = O.resolve_bc(3, [5] * 6)
    fab = _oracle_fab(L, n, f0)
    O.fill_periodic(L, fab, n, [1, 1, 1])
    fe = O.run(L, fab, n, sim.omega, sim.force, [1, 1, 1], res, 100)
    assert np.array_equal(sim.f(), O.valid(L, fe))


def test_uniform_force_every_step_through_executable(tmp_path):
    """tests/regression/uniform_force, checked at every step (plot_int=1):
    u_k = k F / rho exactly and rho = 1, in the plotfiles the executable
    writes (physical units: u * dx/dt)."""
    inp = tmp_path / "inputs"
    inp.write_text(_read("uniform_force_2d"))
    exe = host.EXE_PATH
    r = subprocess.run([exe, str(inp), "amr.plot_int=1", f"amr.plot_file={tmp_path}/plt",
                        "amr.chk_int=0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    u = host.inputs_units(_read("uniform_force_2d"))
    scale = u["dx"] / u["dt"]
    F = np.array([u["F0"], u["F1"]])
    for step in range(30):
        names, t, s, data = host.read_plotfile(str(tmp_path / f"plt{step:07d}"))
        assert names == ["density", "vx", "vy"] and s == step
        k = step + 1
        assert abs(t - k * u["dt"]) < 1e-12
        assert np.max(np.abs(data[0] - 1.0)) < 1e-12
        for d in range(2):
            assert np.max(np.abs(data[1 + d] / scale - k * F[d])) < 1e-12
    # the run ends with the stage table: 30 steps, 30 moment passes and plotfiles
    table = r.stdout[r.stdout.index("Stage timing"):].splitlines()[2:]
    rows = {l.split()[0]: (float(l.split()[1]), int(l.split()[2])) for l in table if l.strip()}
    assert rows["advance"][1] == 30 and rows["advance"][0] > 0
    assert rows["updateMacroscopic"][1] == 30 and rows["writePlotfile"][1] == 30
    assert rows["writeCheckpoint"][1] == 0 and rows["initialize"][1] == 1


def test_stage_times_count_every_stage(tmp_path):
    """Per-stage timing: device time of the steps (single and graph-replayed
    alike), of the moment passes that ran, host time of the writes."""
    sim = host.Simulation(_read("uniform_force_2d"),
                          ["amr.plot_int=10", f"amr.plot_file={tmp_path}/plt", "amr.chk_int=0"])
    sim.initialize()
    sim.advance(1)
    sim.advance(130)
    sim.update_macroscopic(130)
    sim.update_macroscopic(130)           # already up to date: no pass
    sim.write_output(130)                 # a plotfile (its moments are current)
    t = sim.stage_times()
    print(t)
    assert t["advance"][1] == 131 and t["advance"][0] > 0
    assert t["updateMacroscopic"][1] == 1 and t["updateMacroscopic"][0] > 0
    assert t["writePlotfile"][1] == 1 and t["writeCheckpoint"][1] == 0
    assert t["initialize"][1] == 1


def test_restart_is_bit_identical(tmp_path):
    """tests/unit/test_restart.cpp:137-209: run to step 35 writing a
    checkpoint every 10 steps; restart from chk0000020 and finish; the final
    populations are identical."""
    base = _read("lid_cavity_128").replace("128 128", "32 32") + \
        f"\namr.check_file = {tmp_path}/chk\namr.chk_int = 10\nsim.max_step = 35\n"
    a = host.Simulation(base)
    a.initialize()
    for step in range(0, 36):
        a.advance()
        a.write_output(step)
    fa = a.f()
    assert os.path.exists(tmp_path / "chk0000020" / "Header")
    b = host.Simulation(base, [f"amr.restart={tmp_path}/chk0000020", "amr.chk_int=0"])
    b.initialize()
    assert b.start_step == 21 and b.time == 21.0
    for step in range(b.start_step, 36):
        b.advance()
    assert np.array_equal(b.f(), fa)
    with pytest.raises(host.HostError, match="domain mismatch"):
        c = host.Simulation(base.replace("32 32", "16 16"),
                            [f"amr.restart={tmp_path}/chk0000020"])
        c.initialize()


def test_restart_from_a_reference_checkpoint_of_several_boxes(tmp_path):
    """A checkpoint in the reference's layout (tests/unit/test_restart.cpp:
    219-226 restarts across box layouts): the populations of 4 boxes of
    max_grid_size 16, each FAB grown by one ghost cell, in two data files.
    The engine restarts its one box from it, bit for bit, and advances it
    like the state written by this engine itself."""
    from tests.test_host_cpu import write_reference_checkpoint
    base = _read("lid_cavity_128").replace("128 128", "32 32") + \
        f"\namr.check_file = {tmp_path}/chk\namr.chk_int = 10\nsim.max_step = 30\n"
    a = host.Simulation(base)
    a.initialize()
    for step in range(0, 21):
        a.advance()
        a.write_output(step)
    f20 = a.f()
    for step in range(21, 31):
        a.advance()
    boxes = [([0, 0], [15, 15]), ([16, 0], [31, 15]), ([0, 16], [15, 31]), ([16, 16], [31, 31])]
    d = write_reference_checkpoint(str(tmp_path / "ref_chk"), [32, 32], 9, f20.reshape(9, 32, 32),
                                   boxes, ngrow=1, step=20, time=21.0)
    b = host.Simulation(base, [f"amr.restart={d}", "amr.chk_int=0"])
    b.initialize()
    assert b.start_step == 21 and b.time == 21.0
    assert np.array_equal(b.f(), f20)
    for step in range(21, 31):
        b.advance()
    assert np.array_equal(b.f(), a.f())


def test_closed_box_mass_conservation_with_refreshed_moments():
    """tests/unit/test_mass_conservation.cpp with the moments refreshed (the
    reference compares two stale rho fields): NoSlip box, tau=1, 50 steps."""
    text = ("grid.ncells = 12 6\nsim.max_step = 49\nlbm.lattice_type = D2Q9\nlbm.tau = 1.0\n"
            "units.Re = 10\ngeometry.prob_lo = 0 0\ngeometry.prob_hi = 2 1\n"
            "bc.lo_type = NoSlip NoSlip\nbc.hi_type = NoSlip NoSlip\nlbm.init = noise\n"
            "lbm.init_amp = 0.01\namr.plot_int = 0\namr.chk_int = 0\n")
    sim = host.Simulation(text)
    sim.initialize()
    r0 = O.moments(O.Lattice("D2Q9"), _oracle_fab(O.Lattice("D2Q9"), [12, 6], sim.f()), [12, 6],
                   streamed=False)[0].sum()
    sim.advance(50)
    sim.update_macroscopic(49)
    r1, v1 = sim.macro()
    assert abs(r1.sum() - r0) / r0 < 1e-12
    assert np.abs(v1).max() > 0


@pytest.mark.parametrize("case,window,tol,trans,spread", [
    ("poiseuille_periodic", (0.0, 1.0), 1.5e-3, 1e-6, 1e-8),
    ("poiseuille_inflow", (0.5, 0.8), 5e-3, 4e-5, 2e-5),
])
def test_poiseuille_regression(case, window, tol, trans, spread):
    """tests/regression/poiseuille/main.cpp:130-229 criteria: profile error and
    mean velocity vs F H^2/(8 nu), transverse velocity and density spread."""
    text = _read(case)
    u = host.inputs_units(text)
    sim = host.Simulation(text)
    sim.initialize()
    max_step = int([l.split("=")[1] for l in text.splitlines() if l.startswith("sim.max_step")][0])
    sim.advance(max_step + 1)
    sim.update_macroscopic(max_step)
    rho, vel = sim.macro()
    Nx, Ny = sim.n[0], sim.n[1]
    H = float(Ny)
    nu = (u["tau"] - 0.5) / 3.0
    F = u["F0"]
    U_max = F * H * H / (8.0 * nu)
    i_lo, i_hi = int(round(window[0] * Nx)), int(round(window[1] * Nx))
    y = np.arange(Ny) + 0.5
    ua = F / (2 * nu) * y * (H - y)
    ux = vel[0, 0, :, i_lo:i_hi]
    err = np.abs(ux - ua[:, None]).max() / U_max
    mean_err = abs(ux.mean() - 2.0 / 3.0 * U_max) / U_max
    tr = np.abs(vel[1, 0, :, i_lo:i_hi]).max() / U_max
    rs = rho[0, :, i_lo:i_hi].max() - rho[0, :, i_lo:i_hi].min()
    print(case, err, mean_err, tr, rs)
    assert err < tol and mean_err < tol and tr < trans and rs < spread


@pytest.mark.parametrize("lattice", ["D3Q15", "D3Q19", "D3Q27"])
def test_graph_replayed_steps_equal_single_steps(lattice):
    """advance(n) replays captured graphs of 64 steps (taken in pairs through
    mfa_step2 on these periodic boxes) from either buffer; the state, time and
    moments equal those of n single advance() calls, for counts that start
    from both buffers and leave a remainder."""
    text = _read("periodic_d3q19_256").replace("256 256 256", "24 20 16")
    text = text.replace("lbm.lattice_type   = D3Q19", f"lbm.lattice_type   = {lattice}")
    assert f"= {lattice}" in text
    a = host.Simulation(text)
    b = host.Simulation(text)
    a.initialize()
    b.initialize()
    for n in (1, 130, 64, 77, 200):
        a.advance(n)
        for _ in range(n):
            b.advance(1)
        assert np.array_equal(a.f(), b.f()), n
        assert a.time == b.time
    a.update_macroscopic(0)
    b.update_macroscopic(0)
    ra, va = a.macro()
    rb, vb = b.macro()
    assert np.array_equal(ra, rb) and np.array_equal(va, vb)
