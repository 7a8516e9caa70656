"""bench.py's contract on the host: more than one GPU is refused with a message
(exit status 2, before any GPU work), and the dominant kernel's HBM traffic is read
from the committed rocprof summary's TRAFFIC line."""
import importlib.util
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_traffic_line_of_the_force_pass():
    b = _bench()
    kernel = b.force_kernel_name({"kind": "block"})
    t = b.measured_traffic(kernel)
    assert t is not None and 1e9 < t < 1e10
    assert b.measured_traffic("no such kernel") is None
