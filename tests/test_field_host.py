"""Host build of the device field arithmetic (field.h) against 128-bit
remainders, including Goldilocks multiplication by powers of 2."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_field_header_host_build():
    exe = os.path.join(HERE, "cpp", "test_field")
    src = os.path.join(HERE, "cpp", "test_field.cpp")
    if not os.path.exists(exe) or os.path.getmtime(exe) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "test_field"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
