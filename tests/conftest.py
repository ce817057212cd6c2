import importlib
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def jm():
    return importlib.import_module("jax-mtp_amd")


@pytest.fixture(scope="session")
def O():
    import oracle

    oracle.build()
    return oracle


def oracle_for(O, params, with_tree=False, tmp=None):
    """Oracle bound to an MTPData (through the oracle's own .mtp reader)."""
    import tempfile

    jm = importlib.import_module("jax-mtp_amd")
    with tempfile.NamedTemporaryFile("w", suffix=".mtp", delete=False) as f:
        f.write(jm.format_mtp(params))
        path = f.name
    P = O.read_mtp(path)
    pathlib.Path(path).unlink()
    return O.Oracle(P, with_tree=with_tree)
