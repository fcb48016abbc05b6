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


@pytest.mark.parametrize("fused_x", [0, 1])
def test_validation_mode_all_subdomains_match(tmp_path, fused_x):
    # x faces inside the step kernel (the default) or by the copy launch: the same cells
    text = open(os.path.join(DATA, "conf_validate_les.xml")).read().replace(
        "<exchange-mode>1</exchange-mode>", f"<exchange-mode>1</exchange-mode><fused-x>{fused_x}</fused-x>")
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
