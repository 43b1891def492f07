"""CPU checks of bench.py's command-line contract: more than one GPU is refused with a
message and exit status 2 before anything touches a device; the flags it documents
parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "one GPU" in r.stderr and r.stdout.strip() == ""


def test_help_lists_the_contract_flags():
    r = _bench("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--issue-probe", "--copy-probe",
                 "--skip-ensemble", "--tile-split"):
        assert flag in r.stdout, flag
