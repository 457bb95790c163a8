"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
`sanitize` target of oracle/Makefile builds oracle/san_check.c with the
oracle and runs it (edge count of the Kuhn box, symmetry of K, J = K and
R = K a for a linear material, no sanitizer report)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parent.parent / "oracle"


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None,
                    reason="no host C compiler")
def test_oracle_under_sanitizers():
    proc = subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=300)
    if proc.returncode != 0 and "cannot find -lasan" in proc.stdout:
        pytest.skip("the host compiler has no AddressSanitizer runtime")
    assert proc.returncode == 0, proc.stdout[-4000:]
    assert "all checks passed" in proc.stdout, proc.stdout[-4000:]
