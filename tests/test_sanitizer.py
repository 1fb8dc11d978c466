"""The host-code tests again under AddressSanitizer and
UndefinedBehaviorSanitizer: the model library (the product's host code for
density files, Earth-model checks, trajectory geometry and height bins, plus
the host build of the kernel arithmetic) and the oracle, built by
`make asan` in cudaprob3_amd/csrc and oracle, loaded in a child python with
the sanitizer runtimes preloaded.  Any report ends the child with an error."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

MODEL = os.path.join(ROOT, "cudaprob3_amd", "libcp3model_asan.so")
ORACLE = os.path.join(ROOT, "oracle", "liboracle_asan.so")


def runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True,
                         text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_host_code_under_asan_ubsan():
    if not (os.path.exists(MODEL) and os.path.exists(ORACLE)):
        pytest.skip("sanitizer builds not present (make asan)")
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc sanitizer runtimes not found")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan} {ubsan}",
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "CP3_MODEL_SO": MODEL,
        "CP3_ORACLE_SO": ORACLE,
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_model.py"),
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error:" not in r.stderr
