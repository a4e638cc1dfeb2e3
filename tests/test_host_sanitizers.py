"""The C ABI's host code under AddressSanitizer + UndefinedBehaviorSanitizer
(CPU only, no GPU): every HIP source is rebuilt with the sanitizers on its
host side only (``-Xarch_host`` before each ``-fsanitize=``; device code is
unchanged), linked into a copy of the library, and driven by
``tests/host/abi_host_checks.cpp`` (argument validation on every entry point,
size helpers, the SDV reader and writer on the reference's own SDV files).
A sanitizer report aborts the driver, failing the test.  LeakSanitizer is
off: the HIP runtime keeps process-lifetime allocations of its own."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "tt-toolbox_amd" / "csrc"
INCLUDE = ROOT / "include"
GOLDEN = ROOT / "tests" / "golden"
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def _hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(_hipcc()).exists(), reason="hipcc not available")
def test_abi_host_code_under_asan_ubsan(tmp_path):
    from ttgpu import _build
    flags = [f for f in _build.FLAGS if f != "-O3"] + ["-O1", "-g"]
    srcs = sorted(CSRC.glob("*.hip"))

    def comp(src):
        obj = tmp_path / (src.stem + ".o")
        cmd = [_hipcc(), *flags, *SAN, "-I", str(INCLUDE), "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
        p = subprocess.run(cmd, capture_output=True, text=True)
        return obj, p

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = []
        for obj, p in ex.map(comp, srcs):
            assert p.returncode == 0, p.stderr[-3000:]
            objs.append(str(obj))
    lib = tmp_path / "libttgpu_san.so"
    p = subprocess.run([_hipcc(), *flags, *SAN, "-shared", *objs, "-o", str(lib)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    # the driver is plain host C++, built by the same clang (one sanitizer runtime)
    clang = Path(_hipcc()).resolve().parent.parent / "llvm" / "bin" / "clang++"
    if not clang.exists():
        clang = Path("/opt/rocm/llvm/bin/clang++")
    exe = tmp_path / "abi_host_checks"
    p = subprocess.run([str(clang), "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I", str(INCLUDE),
                        str(ROOT / "tests" / "host" / "abi_host_checks.cpp"), str(lib), f"-Wl,-rpath,{tmp_path}",
                        "-o", str(exe)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", HIP_VISIBLE_DEVICES="")
    p = subprocess.run([str(exe), str(GOLDEN / "fmex_A_d80.sdv"), str(GOLDEN / "fmex_b_d80.sdv"), str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, (p.stdout + p.stderr)[-4000:]
    assert "host ABI checks passed" in p.stdout
