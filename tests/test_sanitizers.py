"""The CPU suite again with the host C code of the product (md-bench_amd/host/*.c: lattice set-up,
dump and funcfl readers and writers, EAM tables, drivers) and the oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (`make -C md-bench_amd asan`, `make -C oracle
asan`), the ASan runtime preloaded.  UBSan aborts on the first finding
(-fno-sanitize-recover=all), ASan on the first bad access; leak detection is off (the Python
interpreter itself leaks by design)."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROD = os.path.join(REPO, "md-bench_amd", "_build", "asan", "libmdgfx950.so")
ORC = os.path.join(REPO, "oracle", "_build", "asan", "liboracle_md.so")


def _asan_runtime():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True, timeout=30)
    except (OSError, subprocess.SubprocessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_env():
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("gcc's libasan is not available")
    for d in ("oracle", "md-bench_amd"):
        r = subprocess.run(["make", "-j8", "-C", os.path.join(REPO, d), "asan"],
                           capture_output=True, text=True, timeout=1200)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    env = dict(os.environ, LD_PRELOAD=rt, MDGFX950_LIB=PROD, ORACLE_MD_LIB=ORC,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    return env


def test_sanitized_libraries_are_the_ones_loaded(san_env):
    code = ("import sys; sys.path[:0] = ['tests', 'md-bench_amd']\n"
            "import oracle_bind as ob, importlib.util\n"
            "spec = importlib.util.spec_from_file_location('m', 'md-bench_amd/mdbench_gfx950.py')\n"
            "m = importlib.util.module_from_spec(spec); spec.loader.exec_module(m)\n"
            "ob.lib(); m.lib()\n"
            "maps = open('/proc/self/maps').read()\n"
            "print('asan' if 'libasan' in maps else 'no-asan')\n"
            "print(%r in maps, %r in maps)\n" % (PROD, ORC))
    r = subprocess.run([sys.executable, "-c", code], cwd=REPO, env=san_env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.split() == ["asan", "True", "True"], r.stdout


def test_cpu_suite_under_asan_and_ubsan(san_env):
    r = subprocess.run([sys.executable, "-m", "pytest", "tests", "-m", "not gpu", "-x", "-q",
                        "-p", "no:cacheprovider", "--deselect",
                        "tests/test_sanitizers.py::test_cpu_suite_under_asan_and_ubsan",
                        "--deselect",
                        "tests/test_sanitizers.py::test_sanitized_libraries_are_the_ones_loaded"],
                       cwd=REPO, env=san_env, capture_output=True, text=True, timeout=1800)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert " passed" in r.stdout and "failed" not in r.stdout, tail
