"""The remap divides each pair's N m v by the node mass through the mass's
reciprocal (div_rcp in code_mpm_amd/csrc/mpm_device.h: a product and two FMA
residual corrections).  Its bitwise parity with the reference's per-pair
division (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:560)
rests on div_rcp returning the IEEE quotient over the range it is used in.
The header's own function, compiled for the host, is checked here against
a / b on random, near-midpoint, badly-rounded-reciprocal, exact and signed
zero cases and on the kernels' own quotients (tests/host/div_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "host", "_build", "div_check")


def test_div_rcp_equals_ieee_division():
    if not os.path.exists(CHECK):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "host"), "_build/div_check"], check=True)
    r = subprocess.run([CHECK, "2000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "mismatches 0", r.stdout
    assert sum(1 for ln in lines if ln.endswith(" 0 mismatches")) == 6, r.stdout
