"""The Fortran iso_c_binding layer (chidg_amd/fortran/mod_chidg_hip.f90)
compiles, and on the GPU reproduces the reference known answers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "fortran", "known_answers")


def _flang():
    return shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


def test_fortran_binding_compiles(tmp_path):
    fc = _flang()
    if fc is None:
        pytest.skip("amdflang not available")
    src = os.path.join(ROOT, "chidg_amd", "fortran", "mod_chidg_hip.f90")
    subprocess.run([fc, "-c", src, "-J", str(tmp_path), "-o", str(tmp_path / "m.o")], check=True)
    assert (tmp_path / "mod_chidg_hip.mod").exists()


@pytest.mark.gpu
def test_fortran_known_answers_on_gpu():
    assert os.path.exists(BIN), "build the Fortran test program first (chidg_amd/fortran/build.sh)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASS" in r.stdout
    for name in ("ilu0_apply", "ilu0_apply_folded", "ilu0_apply_tuned", "ilu0_apply_stream", "mat_fetch",
                 "mem_used", "bridge_pattern", "bridge_pack", "bridge_get_packed", "bridge_mv", "bridge_set_packed",
                 "controller_refactor_rule", "controller_solution", "fgmres_3x3", "fgmres_2x2_skew", "mv"):
        assert f"PASS {name}" in r.stdout


@pytest.mark.gpu
def test_fortran_newton_with_device_vectors():
    """Newton steps with q, dq and rhs resident on the device through the
    Fortran binding (hip_vector_t, hip_euler_residual, solve_dev) equal the
    same steps with host arrays."""
    exe = os.path.join(ROOT, "tests", "fortran", "newton_device")
    assert os.path.exists(exe), "build the Fortran test programs first (chidg_amd/fortran/build.sh)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASS" in r.stdout
    for name in ("newton_device_residual_norms", "newton_device_residual_falls", "newton_device_state",
                 "device_vector_dot"):
        assert f"PASS {name}" in r.stdout
