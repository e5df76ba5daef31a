"""bench.py's command-line contract on the host: a multi-GPU request (--gpus N
with N > 1, or a launcher's WORLD_SIZE > 1) ends with a message and a nonzero
status before any device work; this build measures one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "2"})])
def test_multi_gpu_request_exits(args, env):
    e = dict(os.environ, **env)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, env=e,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line
