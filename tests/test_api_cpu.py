"""The drop-in C++ API's host side (settings, scenes, save delegate), driven
by a C++ program compiled against libclsph.so; no device is used."""
import os
import subprocess

from libclsph_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def compile_driver(src, out):
    subprocess.run(["g++", "-O1", "-std=c++17", src, build.API_OUT, build.OUT,
                    "-Wl,-rpath," + os.path.dirname(build.API_OUT), "-pthread", "-o", out],
                   check=True)


def test_api_host_side(tmp_path):
    exe = str(tmp_path / "api_cpu_test")
    compile_driver(os.path.join(ROOT, "tests", "cpp", "api_cpu_test.cpp"), exe)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    for name in ("load_settings", "settings_errors", "scene", "houdini", "util"):
        assert f"ok {name}" in r.stdout


def test_read_kernel_file(tmp_path):
    """readKernelFile (util/cl_boilerplate.cpp:90-112): the file's bytes, or a
    message on standard output and exit status 1 for a missing or empty file."""
    exe = str(tmp_path / "api_cpu_test")
    compile_driver(os.path.join(ROOT, "tests", "cpp", "api_cpu_test.cpp"), exe)
    text = "__kernel void k(void) {}\n// second line\n\ttabbed\n"
    src = tmp_path / "k.cl"
    src.write_bytes(text.encode())
    r = subprocess.run([exe, "read", str(src)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout == text
    missing = str(tmp_path / "nope.cl")
    r = subprocess.run([exe, "read", missing], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stdout == f"{missing} file not found.\n"
    empty = tmp_path / "empty.cl"
    empty.write_bytes(b"")
    r = subprocess.run([exe, "read", str(empty)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stdout == f"{empty} is empty.\n"
