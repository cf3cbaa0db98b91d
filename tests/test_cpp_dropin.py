"""A compiled C++ caller of the drop-in (tests/cpp/dropin_example.cpp, written
the way a NEXUS evaluator calls CKKSEvaluator): it compiles and links against
libnexus_ckks.so on the CPU, and on the GPU its decrypted products, rotations
(by a key and composed by NAF), add_many and reduced-error add are within 2^-16
of the plaintext arithmetic, with a level mismatch thrown as
std::invalid_argument."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_example")


def _build_module():
    import importlib.util

    spec = importlib.util.spec_from_file_location("nexus_amd_build", os.path.join(ROOT, "nexus_amd", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_example_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "nexus_amd", "libnexus_ckks.so")):
        pytest.skip("library not built")
    out = _build_module().build_example(str(tmp_path / "dropin_example"))
    assert os.path.exists(out)


@pytest.mark.gpu
def test_example_runs():
    nx = pytest.importorskip("nexus_amd")
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    assert os.path.exists(BIN), "tests/cpp/dropin_example is built by __graft_entry__.build()"
    print("[case] C++ drop-in example", flush=True)
    sys.stdout.flush()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout + r.stderr)
    assert r.returncode == 0
    assert r.stdout.strip().endswith("ok")
