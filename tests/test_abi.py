"""The C-ABI library loads and exports every symbol include/bcx.h declares;
the CLI and the benchmark reject what they do not support."""
import os
import re
import subprocess
import sys

import bcx


def declared_symbols():
    with open(os.path.join(bcx.HERE, "..", "include", "bcx.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bcx_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    L = bcx.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
        assert name in bcx.SIGNATURES, name


def test_oracle_not_linked_by_product():
    out = subprocess.run(["ldd", bcx.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    with open(os.path.join(bcx.HERE, "bcx.py")) as f:
        assert "import oracle" not in f.read()


def test_cli_usage_exit_code():
    exe = os.path.join(bcx.HERE, "betc")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = subprocess.run([exe, "/nonexistent.txt"], capture_output=True, text=True)
    assert r.returncode == 2 and "cannot be opened" in r.stdout


def test_bench_refuses_several_gpus():
    root = os.path.dirname(bcx.HERE)
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and not r.stdout
