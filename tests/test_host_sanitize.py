"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The host C files of the library that run without a GPU (MT19937, the
kd-forest, the homogeneous kernel map, the distance kernels, the runtime
services) are compiled with gcc -fsanitize=address,undefined and linked, with
the library's unmodified HIP objects and the HIP runtime, into a driver
(tests/c_client/host_sanitize.c) that starts no GPU work.  The run must be
free of sanitizer reports, and its checksums must equal those of the same
driver linked against the product library."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HOST = ROOT / "vlfeat_amd" / "csrc" / "host"
BUILD = ROOT / "vlfeat_amd" / "_build"
LIB = ROOT / "vlfeat_amd" / "lib"
FILES = ["random", "kdtree", "homkermap", "mathop", "generic", "kmeans", "gmm", "fisher_vlad"]


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    if shutil.which("gcc") is None or not (LIB / "libvl.so").exists():
        pytest.skip("needs gcc and the built library")
    hip_objs = sorted(str(p) for p in BUILD.glob("*.hip.o"))
    if not hip_objs:
        pytest.skip("the library's HIP objects are not built here")
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
    objs = []
    for f in FILES:
        o = tmp_path / f"{f}.o"
        subprocess.run(["gcc", "-std=gnu99", "-O1", "-g", "-fPIC", "-fopenmp", "-ffp-contract=off", *san,
                        f"-I{ROOT / 'include'}", f"-I{HOST}", "-c", str(HOST / f"{f}.c"), "-o", str(o)], check=True)
        objs.append(str(o))
    driver = ROOT / "tests" / "c_client" / "host_sanitize.c"
    asan_exe = tmp_path / "host_sanitize_asan"
    subprocess.run(["g++", *san, "-fopenmp", f"-I{ROOT / 'include'}", "-x", "c", str(driver), "-x", "none", *objs,
                    *hip_objs, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(asan_exe)],
                   check=True)
    plain_exe = tmp_path / "host_sanitize_plain"
    subprocess.run(["gcc", "-std=gnu99", "-O2", f"-I{ROOT / 'include'}", str(driver), f"-L{LIB}", "-lvl",
                    f"-Wl,-rpath,{LIB}", "-o", str(plain_exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    a = subprocess.run([str(asan_exe)], capture_output=True, text=True, timeout=300, env=env)
    p = subprocess.run([str(plain_exe)], capture_output=True, text=True, timeout=300)
    print(a.stdout, a.stderr[-3000:])
    assert a.returncode == 0, a.stderr[-3000:]
    assert "AddressSanitizer" not in a.stderr and "runtime error" not in a.stderr
    assert p.returncode == 0
    assert a.stdout == p.stdout
