"""The boundary from compiled code: tests/capi/mpf_caller.cpp, a C++ program
built with the host compiler against include/mpf.h and linked to libmpf.so,
makes the calls of the integration snippets.  Here (no GPU): the narrow cones
from info_narrow_cone.txt (rays.cpp:186-198), bit for bit equal to the
reference's own Miscellaneous::GenerateNarrowCones output
(tests/golden/targets.npz), and the usage error."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "capi", "_bin", "mpf_caller")
TARGETS = np.load(os.path.join(HERE, "golden", "targets.npz"))


def caller():
    """The built caller (building it with the host compiler when absent)."""
    if not os.path.exists(CALLER):
        subprocess.run(["make", "-C", os.path.join(HERE, "capi")], check=True, capture_output=True)
    return CALLER


@pytest.mark.parametrize("k", range(5))
def test_compiled_caller_narrow_cones_match_reference(k, tmp_path):
    n, phi, theta, ty, tz = TARGETS[f"narrow{k}_params"]
    n = int(n)
    d = str(tmp_path) + "/"
    with open(d + "info_narrow_cone.txt", "w") as fh:
        fh.write(f"phi = {int(phi)}\ntheta = {theta}\nthetay = {ty}\nthetaz = {tz}\n")
    out = str(tmp_path / "cones.bin")
    subprocess.run([caller(), "narrow", d, str(n), out], check=True)
    v = np.fromfile(out, dtype=np.float64)
    assert v.size == 14 * n + 9
    np.testing.assert_array_equal(v[:7 * n].reshape(n, 7), TARGETS[f"narrow{k}_cones"])
    np.testing.assert_array_equal(v[7 * n:14 * n].reshape(n, 7), TARGETS[f"narrow{k}_rot"])
    np.testing.assert_array_equal(v[14 * n:].reshape(3, 3), TARGETS[f"narrow{k}_rotm1"])


def test_compiled_caller_usage_and_missing_file(tmp_path):
    r = subprocess.run([caller()], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([caller(), "narrow", str(tmp_path) + "/", "3", str(tmp_path / "x.bin")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and r.stderr.strip()
