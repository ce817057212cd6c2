"""The contraction streams of the atom kernel, checked on the host.

``stream_plan_check.cpp`` builds the plan (jax-mtp_amd/csrc/mtp_plan.cpp)
for a workgroup size and executes every layer's entry streams on the CPU
the way the kernel walks them: lane accumulators, each lane storing its
sum at its own item ends (destination slots encoded in the multiplier
word), partial sums combined in their fixed order.  The moments and
their adjoints must equal a direct evaluation of alpha_index_times (the
reference's conversion.py:95-102 recurrence and its reverse) to 1e-12.
"""
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = ROOT / "jax-mtp_amd" / "csrc"
POT = ROOT / "jax-mtp_amd" / "potentials"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("spc") / "stream_plan_check"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{CSRC}", str(ROOT / "tests" / "stream_plan_check.cpp"),
                    str(CSRC / "mtp_plan.cpp"), str(CSRC / "mtp_format.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("level", [2, 6, 10, 12, 16, 20, 24])
@pytest.mark.parametrize("threads,layout", [(64, "one"), (128, "one"), (256, "one"), (512, "one"), (512, "pair"),
                                            (1024, "pair")])
def test_streams_reproduce_the_contraction(checker, level, threads, layout):
    """One row per workgroup (8-byte slots, byte offsets) and two interleaved rows (16-byte slots, slot indices)."""
    for max_nbr in (32, 80):
        r = subprocess.run([str(checker), str(POT / f"untrained_{level:02d}.mtp"), str(threads), str(max_nbr),
                            str(level), layout], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("level,threads", [(26, 512), (26, 256), (24, 512)])
def test_streams_with_slot_indices(checker, level, threads):
    """Operand fields as 8-byte slot indices from the LDS base (the layout of levels whose moments take
    more than 64 KB of LDS, level 26; checked at level 24 too, where byte offsets would also fit)."""
    r = subprocess.run([str(checker), str(POT / f"untrained_{level:02d}.mtp"), str(threads), "80", str(level), "abs"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
