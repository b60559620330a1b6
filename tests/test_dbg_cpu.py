"""Debug dumps (shyhip_dbg_*, host code) in the reference's format
(shympi_debug, utils/shympi/mpi_debug.f90:114-146,367-415), checked with
the reference's own tool check_shympi_debug, built from its sources into
oracle/_ref/check_dbg (skipped where that was not built): it lists every
time block and record of a file the library wrote, and its comparison of
two files finds no difference within -maxdiff and reports the ones beyond
it.  (The tool compares the first time block of the two files; the model's
first block holds ipv/ipev, which its difference reports look up.)"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "shyfemcm-ismar_amd"))
from shyhip import shyout  # noqa: E402

CHECK = os.path.join(ROOT, "oracle", "_ref", "check_dbg")
pytestmark = pytest.mark.skipif(not os.path.exists(CHECK), reason="reference debug checker not built")


def check(*args):
    """Run the tool beside the files (it keeps file names in 60 characters)."""
    files = [a for a in args if a.endswith(".dbg")]
    cwd = os.path.dirname(files[0])
    args = [os.path.basename(a) if a.endswith(".dbg") else a for a in args]
    p = subprocess.run([CHECK, *args], capture_output=True, text=True, cwd=cwd)
    return p.returncode, p.stdout + p.stderr


def _write(path, times, offset=0.0):
    w = shyout.DbgWriter(path)
    for t in times:
        w.time(t)
        w.record("ipv", np.arange(101, 111, dtype=np.int32))
        w.record("ipev", np.arange(201, 208, dtype=np.int32))
        w.record("znv", (np.arange(10, dtype=np.float32) * np.float32(0.01) + np.float32(offset)))
        w.record("utlnv", np.full((7, 3), 0.5, np.float32))
        w.record("dens", np.full((10, 2), 1025.0, np.float64))
        w.record("iwegv", np.zeros(7, np.int32))
        w.end_time()
    w.close()


def test_listing(tmp_path):
    f = str(tmp_path / "a.dbg")
    _write(f, (60.0, 120.0, 180.0))
    rc, out = check("-verbose", f)
    assert rc == 99, out                                  # the tool's own "finished" code
    assert "time records read:  3" in out
    rows = [line.split() for line in out.splitlines() if line.strip()[:1].isdigit()]
    assert [r[1:] for r in rows[:6]] == [["10", "1", "1", "ipv"], ["7", "1", "1", "ipev"], ["10", "1", "2", "znv"],
                                        ["7", "3", "2", "utlnv"], ["10", "2", "3", "dens"], ["7", "1", "1", "iwegv"]]


def test_compare(tmp_path):
    a, b, c = (str(tmp_path / n) for n in ("a.dbg", "b.dbg", "c.dbg"))
    _write(a, (60.0,))
    _write(b, (60.0,), offset=1e-5)
    _write(c, (60.0,), offset=0.0)
    rc, out = check("-quiet", "-maxdiff", "1e-4", a, b)
    assert "differences found:  0" in out, out
    rc, out = check("-quiet", a, c)
    assert "differences found:  0" in out, out
    rc, out = check("-quiet", "-maxdiff", "1e-7", a, b)
    assert "differences found:  0" not in out, out
