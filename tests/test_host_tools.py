"""Host-side tools that need no GPU: the bench's one-GPU contract and the tile table."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_tile_table_roundtrip(tmp_path):
    from self_amd import autotune
    before = {(k, N): autotune.get_tile(N, k) for k in (1, 2, 3) for N in (3, 5, 7)}
    try:
        assert autotune.count(7, 2) == 18 and autotune.count(5, 3) == 0
        table = {"entries": {"sw:7": {"variant": 5}, "5": {"variant": 9}, "euler:7": {"variant": 2}}}
        p = tmp_path / "tiles.json"
        p.write_text(json.dumps(table))
        autotune.apply(str(p))
        assert autotune.get_tile(7, 2) == 5
        assert autotune.get_tile(5, 2) == 9
        assert autotune.get_tile(7, 3) == 2
        assert autotune.info(7, 5) == {"threads": 128, "alternate": False, "xcd_order": True, "min_waves": 1}
        assert autotune.info(7, 15) == {"threads": 128, "alternate": True, "xcd_order": False, "min_waves": 6}
    finally:
        for (k, N), v in before.items():
            if autotune.count(N, k):
                autotune.set_tile(N, v, k)


def test_committed_tile_table_is_valid():
    from self_amd import autotune
    with open(autotune.TABLE) as f:
        table = json.load(f)
    for key, ent in table["entries"].items():
        kname, _, n = key.rpartition(":")
        kind = autotune.KINDS[kname or "sw"]
        assert -1 <= ent["variant"] < autotune.count(int(n), kind)
