"""The fiber switch under esf_step_batch (eulersf_amd/host/esf_fiber.cpp),
on the CPU: fibers interleave in resume order, keep their own stacks, FP
control state and results, and run again after a reset."""
import math
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "..", "eulersf_amd", "host")


def test_fibers_interleave_and_keep_their_state(tmp_path):
    exe = tmp_path / "fiber_check"
    cmd = ["g++", "-O2", "-std=c++17", f"-I{HOST}", os.path.join(HERE, "native", "fiber_check.cpp"),
           os.path.join(HOST, "esf_fiber.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = dict((l.split(" ", 1)[0] + (l.split()[1] if l.startswith(("acc", "again")) else ""),
                  l.split(" ", 1)[1]) for l in out.stdout.splitlines())
    # round-robin: fiber i's step k comes at position 3 k + i
    order = [int(x) for x in lines["order"].split()]
    assert order == [10 * i + k for k in range(5) for i in range(3)]
    # five rounds end in a yield, one in the fibers finishing, one finds none live
    assert int(lines["rounds"]) == 7
    assert lines["round"] == "1"
    chain = sum(d * 0.5 + 7 for d in range(1, 201))
    for i in range(3):
        acc = float(lines[f"acc{i}"].split()[1])
        again = float(lines[f"again{i}"].split()[1])
        want = sum(math.sqrt(i + k + 1.0) + chain for k in range(5))
        assert abs(acc - want) <= 1e-9 * want
        if i != 1:
            assert acc == again
