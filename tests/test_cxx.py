"""Runs the C++ drop-in tests (tests/cxx/test_main.cxx) and the `atrip` CLI."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cxx", "_build", "test_main")
CLI = os.path.join(ROOT, "atrip_amd", "atrip")


def _need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build()")


def test_cxx_host_cases():
    _need(BIN)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host cases passed" in r.stdout


@pytest.mark.gpu
def test_cxx_run_cases():
    _need(BIN)
    r = subprocess.run([BIN, "--gpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all cases passed" in r.stdout


@pytest.mark.gpu
def test_cli_energy_matches_oracle(oracle, tmp_path):
    _need(CLI)
    no, nv, seed = 7, 15, 3
    r = subprocess.run([CLI, "--no", str(no), "--nv", str(nv), "--seed", str(seed),
                        "--nocheckpoint"], capture_output=True, text=True, timeout=300,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    e = float(re.search(r"^Energy: (\S+)", r.stdout, re.M).group(1))
    e_ref, _ = oracle.run(no, nv, oracle.synthetic(no, nv, seed))
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)


@pytest.mark.gpu
def test_cli_complex_matches_oracle(oracle, tmp_path):
    _need(CLI)
    no, nv, seed = 5, 9, 4
    r = subprocess.run([CLI, "--no", str(no), "--nv", str(nv), "--seed", str(seed),
                        "--complex", "--cT", "--nocheckpoint"], capture_output=True,
                       text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    e = float(re.search(r"^Energy: (\S+)", r.stdout, re.M).group(1))
    ect = float(re.search(r"^Energy \(cT\): (\S+)", r.stdout, re.M).group(1))
    e_ref, ect_ref = oracle.run_z(no, nv, oracle.synthetic_z(no, nv, seed, with_j=True))
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(ect - ect_ref) <= 1e-10 * abs(ect_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["split_fp32", "fp32", "bf16x3"])
def test_cli_mixed_forms(oracle, tmp_path, form):
    """`--mixed-form` selects the mixed form of the drop-in layer; each stays
    within the mixed tolerance (1e-7 relative) of the FP64 oracle."""
    _need(CLI)
    no, nv, seed = 7, 15, 3
    r = subprocess.run([CLI, "--no", str(no), "--nv", str(nv), "--seed", str(seed),
                        "--mixed-form", form, "--nocheckpoint"], capture_output=True,
                       text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    e = float(re.search(r"^Energy: (\S+)", r.stdout, re.M).group(1))
    e_ref, _ = oracle.run(no, nv, oracle.synthetic(no, nv, seed))
    assert abs(e - e_ref) <= 1e-7 * abs(e_ref)


@pytest.mark.gpu
def test_cli_unknown_mixed_form(tmp_path):
    _need(CLI)
    r = subprocess.run([CLI, "--no", "4", "--nv", "6", "--mixed-form", "fp16",
                        "--nocheckpoint"], capture_output=True, text=True, timeout=300,
                       cwd=tmp_path)
    assert r.returncode != 0
    assert "unknown mixed_form" in r.stdout + r.stderr
