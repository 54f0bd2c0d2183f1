"""The package's public modules import on a machine without a GPU (CPU only)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_star_import_exposes_every_listed_module():
    code = ("from feat3_amd import *\n"
            "import feat3_amd\n"
            "names = set(globals())\n"
            "missing = [m for m in feat3_amd.__all__ if m not in names]\n"
            "assert not missing, missing\n"
            "print('ok')\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("ok")
