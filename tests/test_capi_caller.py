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


def test_plane_segment_clamps_to_last_segment():
    """The plane crossing of a bundle ray (csrc/mpf_plane.h, host-compiled):
    a plane value equal to the last element's projection takes the last
    segment with weight 0 (the last element), where upper_bound ends past the
    trajectory; interior values interpolate within their segment."""
    exe = os.path.join(HERE, "capi", "_bin", "plane_segment_test")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(HERE, "capi")], check=True, capture_output=True)
    r = subprocess.run([exe, "3", "2", "1.5", "0", "--", "0", "1", "2", "3"], capture_output=True, text=True,
                       check=True)
    rows = [tuple(float(x) for x in line.split()) for line in r.stdout.split("\n") if line]
    assert rows == [(2, 0.0), (2, 1.0), (1, 0.5), (0, 1.0)]
