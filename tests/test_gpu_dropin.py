"""The C++ drop-in check (tests/native/test_rdme_dropin.cpp) on the GPU: reference-style
callers of population_box / generator / shifted_generator / sector_multigrid / num::gmres,
compiled against the header-only wrappers, pass the reference's own checks."""
import os
import subprocess

import pytest

from tests import drop_in

pytestmark = pytest.mark.gpu


def test_reference_style_checks_pass_on_the_gpu():
    if not os.path.exists(drop_in.BIN):
        drop_in.build()
    r = subprocess.run([drop_in.BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
