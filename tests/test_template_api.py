"""The header-only template API (include/morphstore_hip.hpp): a C++ query
program written in the reference's operator shape compiles and links
against the C-ABI library here; on the GPU it runs and checks itself
against brute force."""
import os
import subprocess

import pytest

import engine_amd as ea

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "template_api_test.cpp")


def build(tmp_path):
    exe = str(tmp_path / "template_api_test")
    libdir = os.path.dirname(ea.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L", libdir, "-lmorphstore_hip", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def test_template_api_compiles_and_links(tmp_path):
    build(tmp_path)


@pytest.mark.gpu
def test_template_api_runs(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "template api ok" in r.stdout
