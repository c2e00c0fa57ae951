"""bench.py host logic (no GPU): the script parses, refuses more than one GPU with a
message, builds every workload, and the atom-steps counting and byte model are as stated."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT


def test_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in json.loads(p.stdout.strip().splitlines()[-1])["error"]


def test_workloads_and_byte_model():
    sys.path.insert(0, ROOT)
    import bench
    for name, n_sys, atoms in (("c1", 1, 256000), ("c2", 1, 2048000)):
        systems, desc, cfg = bench.workload_systems(name)
        assert len(systems) == n_sys and systems[0].np == atoms and desc
    systems, _, cfg = bench.workload_systems("c3")
    assert len(systems) == 512 and cfg["skin_open"]
    # list (4 B per entry), own record, length, force out, kick
    assert bench.force_bytes_per_atom(100.0) == pytest.approx(4 * 100 + 36 + 68)
    with pytest.raises(SystemExit):
        bench.workload_systems("c9")
