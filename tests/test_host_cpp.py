"""The C++ host surface on the CPU: configuration parsing (conf.xml tags,
including the `domian-length` spelling) and the physical -> lattice
parametrization of CLbmSkeleton (src/CLbmSkeleton.hpp:81-190), compiled with
g++ from the headers in turbulent_lbm_multigpu_amd/csrc/host."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "turbulent_lbm_multigpu_amd", "csrc")
DATA = os.path.join(ROOT, "tests", "data")


@pytest.fixture(scope="module")
def selftest(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("host") / "host_selftest")
    subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "cpp", "host_selftest.cpp"), "-o", exe], check=True)
    return exe


def run(exe, *args):
    out = subprocess.run([exe, *args], capture_output=True, text=True)
    return out.returncode, out.stdout


def test_parses_reference_config_layout(selftest):
    rc, out = run(selftest, "config", os.path.join(DATA, "conf_reference_layout.xml"))
    assert rc == 0, out
    lines = dict(ln.split(" ", 1) for ln in out.strip().splitlines())
    assert float(lines["viscosity"]) == pytest.approx(0.001308)
    assert [float(v) for v in lines["gravitation"].split()] == pytest.approx([0, -9.81, 0])
    assert [float(v) for v in lines["cavity"].split()] == [100, 0, 0, 1]
    assert lines["domain"] == "96 32 32" and lines["subdomains"] == "3 1 1"
    assert [float(v) for v in lines["length"].split()] == pytest.approx([0.1, 0.1, 0.1])
    assert lines["loops"].startswith("40 ")
    assert lines["kernel_count"] == "128 device 0"
    assert lines["lattice"].startswith("0 ")


def test_parses_lattice_block(selftest):
    rc, out = run(selftest, "config", os.path.join(DATA, "conf_validate_les.xml"))
    assert rc == 0, out
    lat = out.strip().splitlines()[-1].split()
    assert lat[:2] == ["lattice", "1"]
    assert float(lat[3]) == pytest.approx(0.5005) and float(lat[7]) == pytest.approx(0.17)
    assert lat[-1] == "1"  # exchange mode


def test_rejects_malformed_config(selftest, tmp_path):
    bad = tmp_path / "bad.xml"
    bad.write_text("<lbm-configuration><physics><viscosity>1</physics></lbm-configuration>")
    rc, out = run(selftest, "config", str(bad))
    assert rc == 3 and "Loading XML file failed" in out
    missing = tmp_path / "missing.xml"
    missing.write_text("<lbm-configuration><device><kernel-count>1</kernel-count></device></lbm-configuration>")
    rc, out = run(selftest, "config", str(missing))
    assert rc == 3


def skeleton_float32(n, L, gy, nu, U, tau0=0.953575, gmax=1e-4):
    """src/CLbmSkeleton.hpp:81-116,132-165 evaluated in float32 with numpy."""
    f = np.float32
    dx = f(L) / f(n)
    tau = f(tau0)
    dt = (dx * dx) * (f(2.0) * tau - f(1.0)) / (f(6.0) * f(nu) * f(np.sqrt(f(1.0))))
    g = np.array([0, gy, 0], dtype=np.float32) * ((dt * dt) / dx)
    glen = f(np.sqrt(f(g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]))
    if glen >= f(gmax):
        gp = np.array([0, gy, 0], dtype=np.float32)
        gplen = f(np.sqrt(f(gp[0] * gp[0] + gp[1] * gp[1]) + gp[2] * gp[2]))
        dt = f(np.sqrt((f(gmax) * dx) / gplen))
        g = gp * ((dt * dt) / dx)
        tau = f(0.5) * (dt * f(nu) * f(1.0) * f(6.0)) / (dx * dx) + f(0.5)
    return dx, dt, tau, f(1.0) / tau, g, f(U) * dt


@pytest.mark.parametrize("n,L,gy,nu,U", [
    (64, 0.1, -9.81, 6.44825e-4, 396.18),     # limiter on: tau = 0.6, lid = 0.05
    (32, 0.1, -9.81, 0.001308, 100.0),        # the sample configuration's physics
    (64, 0.1, 0.0, 0.001308, 100.0),          # no gravity: tau stays 0.953575
])
def test_parametrization_matches_float32_closed_form(selftest, n, L, gy, nu, U):
    rc, out = run(selftest, "skeleton", str(n), str(L), str(gy), str(nu), str(U))
    assert rc == 0
    v = out.split()
    got = {v[i]: v[i + 1:i + 4] if v[i] == "g" else v[i + 1] for i in range(0, len(v), 2) if not v[i][0].isdigit()
           and not v[i][0] == "-"}
    dx, dt, tau, inv_tau, g, lid = skeleton_float32(n, L, gy, nu, U)
    assert np.float32(got["dx"]) == dx
    assert np.float32(got["dt"]) == dt
    assert np.float32(got["tau"]) == tau
    assert np.float32(got["inv_tau"]) == inv_tau
    assert np.float32(got["lid"]) == np.float32(lid)


def test_parametrization_worked_example():
    """64^3, L = 0.1, g_y = -9.81, nu = 6.44825e-4, U = 396.18 gives tau = 0.6 and
    lid = 0.05 (dt = sqrt(1e-4 dx / 9.81) = 1.26205e-4)."""
    dx, dt, tau, inv_tau, g, lid = skeleton_float32(64, 0.1, -9.81, 6.44825e-4, 396.18)
    assert abs(float(dt) - 1.26205e-4) < 1e-8
    assert abs(float(tau) - 0.6) < 1e-5
    assert abs(float(lid) - 0.05) < 1e-5


@pytest.mark.parametrize("n,L,gy,nu,U", [
    (64, 0.1, -9.81, 6.44825e-4, 396.18),     # config 1 in physical units: lid 0.05
    (32, 0.1, -9.81, 0.001308, 100.0),
])
def test_solver_hands_the_lattice_lid_to_the_library(tmp_path, n, L, gy, nu, U):
    """CLbmSolver passes the skeleton's lattice values (lid = U dt, not the
    physical U it also keeps, src/CLbmSolver.hpp:567) to lbm_solver_create.
    Linked against the built library; no device call succeeds or is needed."""
    lib_dir = os.path.join(ROOT, "turbulent_lbm_multigpu_amd")
    if not os.path.exists(os.path.join(lib_dir, "liblbm_hip.so")):
        pytest.skip("liblbm_hip.so not built")
    exe = str(tmp_path / "host_solver_selftest")
    subprocess.run([shutil.which("g++") or "g++", "-O1", "-std=c++17", "-Wall", f"-I{CSRC}",
                    f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "cpp", "host_solver_selftest.cpp"),
                    "-o", exe, f"-L{lib_dir}", "-llbm_hip", f"-Wl,-rpath,{lib_dir}"], check=True)
    out = subprocess.run([exe, str(n), str(L), str(gy), str(nu), str(U)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    v = out.stdout.split()
    got = dict(zip(v[0::2], v[1::2]))
    dx, dt, tau, inv_tau, g, lid = skeleton_float32(n, L, gy, nu, U)
    assert np.float32(got["lid"]) == np.float32(lid) and abs(float(lid)) < 0.2
    assert np.float32(got["tau"]) == tau and np.float32(got["inv_tau"]) == inv_tau
