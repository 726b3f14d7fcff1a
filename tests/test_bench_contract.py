"""The benchmark's counting conventions and command line, on the CPU: the
algorithmic bytes per cell of stream_collide (the figures of the survey's
measurement table), the interaction count behind the field kernel's
TFLOP/s (against a brute-force count on a small lattice) and the refusal of
more than one GPU."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from harness import make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("mhd,ft,graphics,expected", [
    (False, 2, True, 169),   # fi 2 x 19 x 4 + flags 1 + rho, u 16
    (False, 1, True, 93),    # 2-byte DDFs
    (False, 2, False, 153),  # no UPDATE_FIELDS
    (True, 2, True, 405),    # + ei 152, fqi 56, E_dyn/B_dyn 24, Q 4
    (True, 0, True, 225),
    (True, 1, True, 225),
])
def test_stream_collide_bytes_per_cell(mhd, ft, graphics, expected):
    cfg = make_config(mhd=mhd, float_type=ft, graphics=graphics, lod_depth=2)
    assert bench.bytes_per_cell(cfg) == expected


def test_field_interactions_match_a_brute_force_count():
    """Fluid cells x charged top-level LOD sources of the reference's range,
    less each cell's own LOD: counted pair by pair on 8 x 8 x 4 cells, D = 2
    (zero-charge sources, solid cells and sources outside the top level
    present)."""
    cfg = make_config(nx=8, ny=8, nz=4, mhd=True, lod_depth=2)
    rng = np.random.default_rng(3)
    D, N = 2, 8 * 8 * 4
    nd = 1 << D
    hi = ((1 << (3 * (D + 1))) - 1) // 7  # 73 LODs at depth 2: 1 + 8 + 64
    lo = hi - nd ** 3
    qu = rng.standard_normal((hi, 4)).astype(np.float32)
    qu[rng.uniform(size=hi) < 0.4, 0] = 0.0
    flags = np.where(rng.uniform(size=N) < 0.25, 1, 0).astype(np.uint8)
    count = 0
    for n in range(N):
        if flags[n] & 0x1F == 1:
            continue
        x, y, z = n % 8, n // 8 % 8, n // 64
        own = lo + x // (8 // nd) + (y // (8 // nd) + z // (4 // nd) * nd) * nd
        for d in range(lo, hi):
            if d != own and qu[d, 0] != 0:
                count += 1
    assert bench.field_interactions(cfg, flags, qu.reshape(-1)) == count


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
