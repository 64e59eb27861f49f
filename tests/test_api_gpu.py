"""The drop-in C++ API on the device: sph_simulation::simulate with callbacks
and checkpoints, and the command-line driver end to end."""
import os
import subprocess

import pytest

from libclsph_amd import build

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_simulate_callbacks_and_resume(tmp_path):
    exe = str(tmp_path / "api_gpu_test")
    subprocess.run(["g++", "-O1", "-std=c++17",
                    os.path.join(ROOT, "tests", "cpp", "api_gpu_test.cpp"), build.API_OUT,
                    build.OUT, "-Wl,-rpath," + os.path.dirname(build.API_OUT), "-pthread",
                    "-o", exe], check=True)
    print("case api_gpu_test", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    for name in ("callbacks", "abi_agreement", "resume", "overlapped_callbacks"):
        assert f"ok {name}" in r.stdout


def test_cli_end_to_end(tmp_path):
    (tmp_path / "fluid_properties").mkdir()
    (tmp_path / "simulation_properties").mkdir()
    (tmp_path / "scenes").mkdir()
    (tmp_path / "out_frames").mkdir()
    (tmp_path / "fluid_properties" / "water.json").write_text(
        '{ "fluid_density" : 1000, "dynamic_viscosity" : 0.0035, "restitution" : 0,\n'
        '  "k" : 3, "surface_tension_threshold" : 7.065, "surface_tension" : 0.0728,\n'
        '  "particles_inside_influence_radius" : 20 };')
    (tmp_path / "simulation_properties" / "small.json").write_text(
        '{ "particles_count" : 8192, "particle_mass" : 0.001, "simulation_time" : 0.004,\n'
        '  "target_fps" : 1000, "simulation_scale" : 1, "write_all_frames" : false,\n'
        '  "serialize" : true, "constant_acceleration" : { "x" : 0, "y" : -9.81, "z" : 0 } };')
    (tmp_path / "scenes" / "floor.obj").write_text(
        "v 1 -0.5 1\nv -1 -0.5 1\nv 1 -0.5 -1\nv -1 -0.5 -1\nf 2 1 3\nf 4 2 3\n")
    print("case cli", flush=True)
    r = subprocess.run([build.CLI_OUT, "water", "small", "floor.obj", "out_", "--yes"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    frames = sorted(os.listdir(tmp_path / "out_frames"))
    assert frames == [f"frame{k:07d}.geo" for k in range(1, 5)]
    text = (tmp_path / "out_frames" / "frame0000004.geo").read_text()
    assert text.startswith("PGEOMETRY V5\nNPoints 8192 NPrims 1\n")
    assert os.path.getsize(tmp_path / "last_frame.bin") == 8192 * 80
    # a second run resumes from the checkpoint; a wrong-sized one is refused
    (tmp_path / "last_frame.bin").write_bytes(b"\0" * 80)
    r = subprocess.run([build.CLI_OUT, "water", "small", "floor.obj", "out_", "--yes"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "incorrect size" in r.stdout
