"""The serving driver raises HIP's hardware-queue count before the runtime loads
(every request class runs on its own stream; classes sharing one of HIP's
default 4 queues run one after the other).  CPU only: importing the module
loads no HIP library."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _queues_after_import(env_value):
    env = dict(os.environ)
    if env_value is None:
        env.pop("GPU_MAX_HW_QUEUES", None)
    else:
        env["GPU_MAX_HW_QUEUES"] = env_value
    code = "import os, hyteg_amd.serving; print(os.environ['GPU_MAX_HW_QUEUES'])"
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, check=True)
    return out.stdout.strip()


def test_default_and_small_settings_are_raised_to_16():
    assert _queues_after_import(None) == "16"
    assert _queues_after_import("4") == "16"
    assert _queues_after_import("bogus") == "16"


def test_larger_settings_are_kept():
    assert _queues_after_import("24") == "24"
