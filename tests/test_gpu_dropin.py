"""The drop-in C++ Solver class (include/Solver.h) built against a reference-format config.h."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from incompflow import _abi, Solver

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "incompflow-cuda_amd", "examples")
LIBDIR = os.path.join(ROOT, "incompflow-cuda_amd", "incompflow")


def test_dropin_solver_example(tmp_path):
    cfg_text = open(os.path.join(EX, "cavity", "config.h")).read()
    cfg_text = re.sub(r"constexpr int nx = \d+;", "constexpr int nx = 24;", cfg_text)
    cfg_text = re.sub(r"constexpr int ny = \d+;", "constexpr int ny = 20;", cfg_text)
    cfg_text = re.sub(r"constexpr int numOuterIter = \d+;", "constexpr int numOuterIter = 1500;", cfg_text)
    (tmp_path / "config.h").write_text(cfg_text)
    shutil.copy(os.path.join(EX, "cavity", "types.h"), tmp_path / "types.h")
    exe = tmp_path / "example"
    subprocess.run(["g++", "-O2", "-std=c++20", f"-I{tmp_path}", f"-I{os.path.join(ROOT, 'include')}",
                    "-o", str(exe), os.path.join(EX, "main.cpp"), f"-L{LIBDIR}", "-lincompflow",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    out = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    assert text.startswith("iter: 1\npRelNorm : ")            # Solver.cu:270-285 log block
    assert "massNorm : " in text and "Elapsed time: " in text
    iters = [int(x) for x in re.findall(r"^iter: (\d+)$", text, re.M)]
    assert iters[0] == 1 and iters[-1] <= 1500
    vtk = (tmp_path / "output" / "test1.vtk").read_text()
    assert vtk.startswith("# vtk DataFile Version 3.0\n")
    # the same case through the Python binding reaches the same state
    c = _abi.cavity_config(n=24)
    c.ny = 20
    c.stop_rule, c.num_outer_iter, c.p_refine = _abi.STOP_FAITHFUL, 1500, 0
    c.tol[0] = c.tol[1] = c.tol[2] = 1e-4
    c.tol[3] = 1e-2
    c.tol_mass = 1e-9
    s = Solver(c)
    r = s.solve(1500, history=False)
    assert r["iterations"] == iters[-1]
    lines = vtk.split("\n")
    k = lines.index("pressure 1 480 float")
    p_vtk = np.array(lines[k + 1].split(), dtype=np.float64)
    np.testing.assert_allclose(p_vtk, s.get(_abi.FIELD_P).ravel(), rtol=1e-5, atol=1e-6)   # 6 digits


def test_dropin_solver_with_energy(tmp_path):
    """IFL_SOLVE_ENERGY turns the temperature equation on behind the unchanged Solver surface; the
    config.h's TempBCs (north 293 K, south 373 K, east/west adiabatic in the example case) drive it."""
    cfg_text = open(os.path.join(EX, "cavity", "config.h")).read()
    cfg_text = re.sub(r"constexpr int nx = \d+;", "constexpr int nx = 24;", cfg_text)
    cfg_text = re.sub(r"constexpr int ny = \d+;", "constexpr int ny = 20;", cfg_text)
    cfg_text = re.sub(r"constexpr int numOuterIter = \d+;", "constexpr int numOuterIter = 1500;", cfg_text)
    (tmp_path / "config.h").write_text(cfg_text)
    shutil.copy(os.path.join(EX, "cavity", "types.h"), tmp_path / "types.h")
    exe = tmp_path / "example_energy"
    subprocess.run(["g++", "-O2", "-std=c++20", "-DIFL_SOLVE_ENERGY", f"-I{tmp_path}",
                    f"-I{os.path.join(ROOT, 'include')}", "-o", str(exe), os.path.join(EX, "main.cpp"),
                    f"-L{LIBDIR}", "-lincompflow", f"-Wl,-rpath,{LIBDIR}"], check=True)
    out = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = (tmp_path / "output" / "test1.vtk").read_text().split("\n")
    k = lines.index("temperature 1 480 float")
    T = np.array(lines[k + 1].split(), dtype=np.float64).reshape(20, 24)
    assert T.min() > 290.0 and T.max() < 376.0
    rows = T.mean(axis=1)
    assert rows[0] > rows[-1] + 40.0              # hot south wall, cold north wall
    assert 300.0 < rows[10] < 366.0
