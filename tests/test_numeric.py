"""Host build of the shared numeric header: the branch-free log of the SSA's waiting time
gives log_f's bits on its whole domain (y = 1/r1 >= 1, and +inf for r1 = 0)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_log_f_ge1_matches_log_f(tmp_path):
    exe = str(tmp_path / "log_check")
    src = os.path.join(HERE, "cpp", "log_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", src, "-o", exe], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
