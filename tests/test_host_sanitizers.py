"""The library's host-only code (mesh readers, CellIni, generators, packing,
mesh / result writers, the "%f" formatter, the ini reader) built with g++
under AddressSanitizer + UndefinedBehaviorSanitizer and, for the threaded
reader and writer, under ThreadSanitizer, driven by tests/sanitize/
host_check.cpp on the shipped-style coarse mesh and on a mesh large enough
for the line-parallel reader.  Bar: the driver's checks pass and the
sanitizers report nothing (any report aborts the run)."""
import os
import shutil
import subprocess

import pytest

import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parallellbmcranfield_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SANITIZERS = {
    "address_undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
    "thread": ["-fsanitize=thread"],
}


@pytest.fixture(scope="module")
def meshes(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshes")
    small, large = d / "small", d / "large"
    meshgen.cavity(125).write(str(small))
    meshgen.cavity(400).write(str(large))  # D2node.dat of about 4.8 MB: the line-parallel reader
    assert os.path.getsize(large / "Mesh" / "D2node.dat") > 1 << 20
    return str(small / "Mesh"), str(large / "Mesh")


@pytest.mark.parametrize("kind", sorted(SANITIZERS))
def test_host_code_is_clean_under_sanitizers(kind, meshes, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / f"host_check_{kind}")
    cmd = (["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread"] +
           SANITIZERS[kind] +
           ["-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"), "-I", CSRC,
            os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
            os.path.join(CSRC, "case_io.cpp"), "-o", exe])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    out = tmp_path / "out"
    out.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    r = subprocess.run([exe, *meshes, str(out)], capture_output=True, text=True, timeout=600,
                       env=env)
    report = r.stdout + r.stderr
    assert r.returncode == 0 and "OK" in r.stdout, report[-6000:]
    for word in ("AddressSanitizer", "LeakSanitizer", "runtime error:", "ThreadSanitizer"):
        assert word not in report, report[-6000:]
