import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def reference_dir():
    """The reference tree exists only in the build container; tests that read
    it are skipped elsewhere (the fixtures they derive are committed)."""
    p = Path(os.environ.get("MDFF_REFERENCE", "/root/reference"))
    if not (p / "src").is_dir():
        pytest.skip("reference tree not present")
    return p
