"""include/rcg/RayleighChebyshevLM.hpp: the C++ mirror of the reference's
template API compiles against the C ABI (CPU) and solves through librcg.so
(GPU), with the reference's calling sequence (RayleighChebyshevLM.h:712-738)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "solve_laplacian.cpp")
LIB = os.path.join(ROOT, "rayleighchebyshevgpgpu_amd", "librcg.so")


def test_header_compiles_standalone(tmp_path):
    # header alone, strict warnings, no HIP headers needed by the caller
    src = tmp_path / "h.cpp"
    src.write_text('#include "rcg/RayleighChebyshevLM.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def test_cpp_caller_links(tmp_path):
    if not os.path.exists(LIB):
        pytest.fail("librcg.so is not built")
    out = tmp_path / "solve_laplacian"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", str(out), LIB], check=True)
    assert out.exists()


@pytest.mark.gpu
def test_cpp_caller_solves():
    import __graft_entry__ as ge
    exe = ge.CPP_EXAMPLE
    deps = [SRC, LIB, os.path.join(ROOT, "include", "rcg_c_api.h"),
            os.path.join(ROOT, "include", "rcg", "RayleighChebyshevLM.hpp")]
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(d) for d in deps):
        ge.build_cpp_example()  # stale against the header or the library
    print("cpp caller: 24x24 Dirichlet Laplacian, k=10", flush=True)
    p = subprocess.run([exe, "24", "10", "1e-10"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    lines = p.stdout.strip().splitlines()
    rows = np.array([[float(t) for t in ln.split()] for ln in lines[:-2]])
    assert rows.shape == (10, 5)
    assert np.max(np.abs(rows[:, 0] - rows[:, 1])) < 1e-10
    # the RayleighChebyshev (host-vector) interval solve with the caller's bounds
    words = lines[-2].split()
    assert words[0] == "interval" and words[4] == words[6]
    assert lines[-1].startswith("found 10 ")


@pytest.mark.gpu
def test_cpp_matrix_free_caller_solves():
    """A C++ caller whose operator is its own HIP stencil kernel (no matrix),
    through rcg::CallbackOperator: the Dirichlet spectrum in closed form."""
    import __graft_entry__ as ge
    exe = ge.CPP_MATRIX_FREE
    deps = [exe + ".hip", LIB, os.path.join(ROOT, "include", "rcg_c_api.h"),
            os.path.join(ROOT, "include", "rcg", "RayleighChebyshevLM.hpp")]
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(d) for d in deps):
        ge.build_cpp_matrix_free()
    print("cpp matrix-free caller: 20x20 Dirichlet stencil kernel, k=8", flush=True)
    p = subprocess.run([exe, "20", "8", "1e-10"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    lines = p.stdout.strip().splitlines()
    rows = np.array([[float(t) for t in ln.split()] for ln in lines[:-1]])
    assert rows.shape == (8, 3)
    assert np.max(np.abs(rows[:, 0] - rows[:, 1])) < 1e-10
    words = lines[-1].split()
    assert words[0] == "found" and words[1] == "8" and int(words[3]) > 8
