"""The C++ driver (turbulent_lbm_multigpu_amd/lbm_main) on the GPU: the
reference's validation mode (decomposed grid == one domain of N - 2(k-1)
cells, src/main.cpp:309-408) and a normal run from the reference's
configuration layout."""
import os
import re
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "turbulent_lbm_multigpu_amd", "lbm_main")
DATA = os.path.join(ROOT, "tests", "data")


def run(*args, cwd=None, timeout=120):
    print("case: lbm_main", *args, flush=True)
    sys.stdout.flush()
    if not os.path.exists(EXE):
        raise FileNotFoundError(f"{EXE} missing: build with __graft_entry__.build()")
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)


@pytest.mark.parametrize("fused_x,alpha_wg", [(0, 0), (1, 0), (1, 512)])
def test_validation_mode_all_subdomains_match(tmp_path, fused_x, alpha_wg):
    # x faces inside the step kernel (the default) or by the copy launch, alpha at the group's
    # work-group size or at its own: the same cells
    text = open(os.path.join(DATA, "conf_validate_les.xml")).read().replace(
        "<exchange-mode>1</exchange-mode>",
        f"<exchange-mode>1</exchange-mode><fused-x>{fused_x}</fused-x><alpha-work-group>{alpha_wg}</alpha-work-group>")
    conf = str(tmp_path / "conf_fused.xml")
    open(conf, "w").write(text)
    r = run("-c", conf, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = re.findall(r"NUMBER OF FAILED CELLS/TOTAL NUMBER OF CELLS: (\d+)/(\d+)", r.stdout)
    assert len(lines) == 8
    for failed, total in lines:
        assert failed == "0" and int(total) == 30 * 22 * 22


def test_normal_run_reference_layout(tmp_path):
    r = run("-c", os.path.join(DATA, "conf_reference_layout.xml"), cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "limiting timestep" in r.stdout
    m = re.search(r"MLUPS: ([0-9.eE+-]+)", r.stdout)
    assert m and float(m.group(1)) > 0
    assert "Cube: [96, 32, 32]" in r.stdout


def test_cli_flags_and_vtk_output(tmp_path):
    r = run("-x", "24", "-y", "16", "-z", "16", "-X", "2", "-l", "3", "-g", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted(os.listdir(tmp_path / "output" / "vtk"))
    assert files[0] == "OUTPUT.0.0.vtk" and len(files) == 6
    head = (tmp_path / "output" / "vtk" / "OUTPUT.1.2.vtk").read_text().splitlines()[:8]
    assert head[0].startswith("# vtk DataFile") and "STRUCTURED_GRID" in head[4]
    assert head[5].split() == ["DIMENSIONS", "13", "17", "17"]


# --------------------------------------------------------------------------
# The caller surface anchored to the oracle: conf.xml -> CConfiguration ->
# CLbmSkeleton -> CManager/CController -> CLbmSolver -> kernels, then the
# reference's outputs (debug checksum, VTK files, benchmark ini) compared
# with the CPU oracle run of the same case.

def _oracle_checksum(vel, flags):
    """getVelocityChecksum (src/CLbmSolver.hpp:1103-1123): float sum over
    FLUID cells of vx + vy + vz, in cell order, times 1000.0f."""
    import numpy as np

    fl = flags == 2
    per = (vel[0][fl] + vel[1][fl]) + vel[2][fl]
    return np.float32(np.cumsum(per.astype(np.float32), dtype=np.float32)[-1]) * np.float32(1000.0), \
        float(np.abs(vel[:, fl]).sum()) * 1000.0


def _conf(path, *, size, nums, loops, physics, vtk=0, benchmark=0, lattice=""):
    visc, g, cav = physics
    open(path, "w").write(f"""<?xml version="1.0"?>
<lbm-configuration>
  <physics>
    <viscosity>{visc}</viscosity>
    <gravitation><x>{g[0]}</x><y>{g[1]}</y><z>{g[2]}</z></gravitation>
    <cavity-velocity><x>{cav}</x><y>0</y><z>0</z><w>1</w></cavity-velocity>
  </physics>
  <grid>
    <domain-size><x>{size[0]}</x><y>{size[1]}</y><z>{size[2]}</z></domain-size>
    <subdomain-num><x>{nums[0]}</x><y>{nums[1]}</y><z>{nums[2]}</z></subdomain-num>
    <domian-length><x>0.1</x><y>0.1</y><z>0.1</z></domian-length>
  </grid>
  <simulation>
    <loops>{loops}</loops>
    <timestep>-1.0</timestep>
    <visualization><VTK>{vtk}</VTK></visualization>
    <validate>0</validate>
    <benchmark>{benchmark}</benchmark>
  </simulation>
  <device>
    <kernel-count>256</kernel-count>
    <device-number>0</device-number>
  </device>{lattice}
</lbm-configuration>
""")


def test_config1_through_physical_surface_matches_oracle(tmp_path):
    """Config 1 (64^3 BGK cavity, tau 0.6, lid 0.05, 1000 steps) expressed in
    the reference's physical units (g_y = -9.81 turns on the time-step limiter:
    dt = sqrt(1e-4 dx / 9.81), nu = 6.44825e-4 gives tau = 0.6, U = 396.18
    gives lid = 0.05).  The driver's debug checksum equals the oracle's run
    with the float32 parametrization to 1e-5 of sum |v|; the benchmark ini
    holds the reference's keys and formulas (src/CController.hpp:445-489)."""
    import numpy as np

    from oracle import oracle as orc
    from tests.test_host_cpp import skeleton_float32
    from turbulent_lbm_multigpu_amd import decomposition as dec

    n, steps = 64, 1000
    nu, U = 6.44825e-4, 396.18
    conf = str(tmp_path / "conf_config1.xml")
    _conf(conf, size=(n, n, n), nums=(1, 1, 1), loops=steps, physics=(nu, (0, -9.81, 0), U), benchmark=1)
    r = run("-c", conf, "-v", cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "limiting timestep" in r.stdout
    m = re.search(r"Checksum: ([-0-9.]+)\s", r.stdout)
    assert m, r.stdout[-600:]
    got = float(m.group(1))

    dx, dt, tau, inv_tau, g, lid = skeleton_float32(n, 0.1, -9.81, nu, U)
    assert abs(float(tau) - 0.6) < 1e-5 and abs(float(lid) - 0.05) < 1e-5
    o = orc.OracleDomain((n, n, n), (1,) * 6, inv_tau=float(inv_tau), tau=float(tau), lid_velocity=float(lid),
                         gravitation=tuple(float(v) for v in g))
    origin, size = dec.lid_region((n, n, n))
    o.set_flags(origin, size, orc.FLAG_INJECTION)
    o.run(steps)
    want, scale = _oracle_checksum(o.vel, o.flags)
    print(f"checksum: driver {got:.8f}, oracle {float(want):.8f}, scale {scale:.3e}", flush=True)
    assert scale > 1.0
    assert abs(got - float(want)) <= 1e-5 * scale

    # the report lines and the benchmark ini (floats per cell: 19*2 + 1 + 3 in debug mode)
    secs = float(re.search(r"Seconds: ([0-9.eE+-]+)", r.stdout).group(1))
    assert f"Cube: [{n}, {n}, {n}]" in r.stdout
    ini = (tmp_path / "output" / "benchmark" / "benchmark_1.ini").read_text()
    kv = dict((k.strip(), float(v)) for k, v in (ln.split(":") for ln in ini.splitlines() if ":" in ln))
    assert set(kv) == {"CUBE_X", "CUBE_Y", "CUBE_Z", "SECONDS", "FPS", "MLUPS", "BANDWIDTH"}
    assert kv["CUBE_X"] == kv["CUBE_Y"] == kv["CUBE_Z"] == n
    # every value is printed with the stream's default 6 significant digits:
    # two roundings of <= 5e-6 each separate a value from the one derived from it
    assert kv["SECONDS"] == pytest.approx(secs, rel=1e-5)
    assert kv["FPS"] == pytest.approx(steps / kv["SECONDS"], rel=1.1e-5)
    assert kv["MLUPS"] == pytest.approx(kv["FPS"] * n ** 3 * 1e-6, rel=1.1e-5)
    assert kv["BANDWIDTH"] == pytest.approx(kv["MLUPS"] * 42 * 4, rel=1.1e-5)
    mlups = float(re.search(r"MLUPS: ([0-9.eE+-]+)", r.stdout).group(1))
    assert mlups == pytest.approx(kv["MLUPS"], rel=1.1e-5)


def _read_vtk(path):
    """Legacy-ASCII STRUCTURED_GRID of CLbmVisualizationVTK: dims, points,
    cell density, flag and velocity."""
    import numpy as np

    lines = open(path).read().splitlines()
    i = next(k for k, ln in enumerate(lines) if ln.startswith("DIMENSIONS"))
    dims = tuple(int(v) for v in lines[i].split()[1:4])
    npts = int(lines[i + 1].split()[1])
    pts = np.array([ln.split() for ln in lines[i + 3:i + 3 + npts]], dtype=np.float64)
    k = next(k for k, ln in enumerate(lines) if ln.startswith("CELL_DATA"))
    ncell = int(lines[k].split()[1])
    dens = np.array(lines[k + 3:k + 3 + ncell], dtype=np.float64)
    k2 = k + 3 + ncell
    assert lines[k2 + 1].startswith("SCALARS flag INT")
    flags = np.array(lines[k2 + 3:k2 + 3 + ncell], dtype=np.int64)
    k3 = k2 + 3 + ncell
    assert lines[k3 + 1].startswith("VECTORS velocity float")
    vel = np.array([ln.split() for ln in lines[k3 + 2:k3 + 2 + ncell]], dtype=np.float64).T
    return dims, pts, dens, flags, vel


@pytest.mark.parametrize("exchange_mode", [0, 1])
def test_decomposed_les_vtk_output_matches_oracle(tmp_path, exchange_mode):
    """A 2x1x1 decomposed LES cavity with the seeded perturbation, configured
    by the <lattice> block: each subdomain's VTK file after 11 steps holds the
    oracle decomposition's density output, flags and velocity output
    (src/libvis/CLbmVisualizationVTK.hpp:27-99), to 1e-5 relative plus the
    %f print resolution (5e-7).

    The oracle exchanges all 19 populations after alpha, as the reference
    does (exchange mode 0): then every cell agrees.  Exchange mode 1 copies
    only the 5 populations a fluid cell reads from a ghost; the others stay
    stale in the ghost layer.  The OBSTACLE cells of the outer x walls read
    such slots through the linear wrap (their x - 1 or x + 1 neighbour is the
    ghost cell of the adjacent row), so their density output -- never a fluid
    value -- differs there; every other cell and every velocity agree."""
    import numpy as np

    from tests.test_oracle import build_oracle_decomposition

    size, nums, steps = (34, 26, 22), (2, 1, 1), 11
    tau, lid, cs, amp, seed = 0.5005, 0.1, 0.17, 1e-4, 20260101
    conf = str(tmp_path / "conf_les_vtk.xml")
    lattice = f"""
  <lattice>
    <tau>{tau}</tau>
    <lid-velocity>{lid}</lid-velocity>
    <smagorinsky-cs>{cs}</smagorinsky-cs>
    <exchange-mode>{exchange_mode}</exchange-mode>
    <perturbation><amplitude>{amp}</amplitude><seed>{seed}</seed></perturbation>
  </lattice>"""
    _conf(conf, size=size, nums=nums, loops=steps, physics=(0.001308, (0, 0, 0), 100), vtk=1, lattice=lattice)
    r = run("-c", conf, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr

    # the C++ layer's float parameters: tau = (float)0.5005, inv_tau = 1.0f / tau
    t32 = np.float32(tau)
    kw = dict(inv_tau=float(np.float32(1.0) / t32), tau=float(t32), smagorinsky_cs=cs)
    multi, _ = build_oracle_decomposition(size, nums, lid=float(np.float32(lid)), amp=amp, seed=seed, **kw)
    multi.sync(after_alpha=True)   # CManager's initial ghost-layer copy (all 19 populations)
    multi.run(steps)
    dx = np.float32(0.1) / np.float32(size[0])
    for uid, sub in enumerate(multi.subs):
        d = sub["domain"]
        dims, pts, dens, flags, vel = _read_vtk(tmp_path / "output" / "vtk" / f"OUTPUT.{uid}.{steps - 1}.vtk")
        assert dims == tuple(v + 1 for v in d.dims)
        ox = sub["sd"].domain.origin[0] * dx
        assert pts[0] == pytest.approx([ox, 0, 0], abs=1e-6)
        assert pts[-1] == pytest.approx([ox + d.dims[0] * dx, d.dims[1] * dx, d.dims[2] * dx], abs=1e-6)
        np.testing.assert_array_equal(flags, d.flags)
        assert (flags == 2).sum() > 0.5 * flags.size
        keep = np.ones(flags.size, bool)
        if exchange_mode == 1:
            x = np.arange(flags.size) % d.dims[0]
            wall = (flags == 1) & (((x == 0) & (uid == 0)) | ((x == d.dims[0] - 1) & (uid == len(multi.subs) - 1)))
            keep = ~wall
        np.testing.assert_allclose(dens[keep], d.dens[keep], rtol=1e-5, atol=5e-7)
        np.testing.assert_allclose(vel, d.vel, rtol=1e-5, atol=5e-7)
        assert np.abs(d.vel).max() > 1e-3 and np.abs(d.dens).max() > 0.5   # fields that can fail the check
