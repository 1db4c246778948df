"""bench.py's command line without a GPU: more than one GPU is refused with a
message before anything touches a device."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "runs on one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""
