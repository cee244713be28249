"""GPU parity of the opt-in cross-lane middle pass (KWH_MIDDLE_XLANE=1: the inner radix of the
outermost-axis transforms exchanged with DPP moves instead of an LDS transpose). The switch is
read once per process, so the cases (xlane_cases.py) run in one child process with it set; each
is checked there against the float64 oracle at the reference's parity tolerances."""
import os
import subprocess
import sys

import pytest

from test_gpu_parity import _device  # noqa: F401  (module fixture: the library must load)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_xlane_middle_pass_parity():
    env = dict(os.environ, KWH_MIDDLE_XLANE="1")
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "xlane_cases.py")], env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("[ok]") == 6
