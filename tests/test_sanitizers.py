"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the Matrix Market parser (rcg_io.cpp) on well-formed and malformed files --
truncated, out-of-range indices, rectangular, bad headers, huge sizes, empty
-- and the CPU oracle solving a Dirichlet Laplacian (real and complex).  A
sanitizer report fails the run; parser results and eigenvalues are checked."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

FILES = {
    "ok_real.mtx": "%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n1 1 2.0\n2 1 -1\n2 2 2\n3 3 2\n",
    "ok_cplx.mtx": ("%%MatrixMarket matrix coordinate complex hermitian\n% comment\n2 2 3\n"
                    "1 1 1.0 0\n2 1 0.5 0.25\n2 2 1 0\n"),
    "ok_pattern.mtx": "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n2 1\n",
    "dup.mtx": "%%MatrixMarket matrix coordinate integer general\n2 2 3\n1 1 1\n1 1 2\n2 2 5\n",
    "truncated.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 2.0\n2 1\n",
    "oob.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 2.0\n",
    "zero_index.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 1\n0 1 2.0\n",
    "rect.mtx": "%%MatrixMarket matrix coordinate real general\n3 4 1\n1 1 2.0\n",
    "bad_header.mtx": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "no_header.mtx": "3 3 1\n1 1 1\n",
    "huge.mtx": "%%MatrixMarket matrix coordinate real general\n9999999999 9999999999 1\n1 1 1\n",
    "neg_nnz.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 -5\n",
    "empty.mtx": "",
    "garbage.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 2\nx y z\n\x00\x01\x02\n",
}
EXPECT_OK = {"ok_real.mtx": (3, 5), "ok_cplx.mtx": (2, 4), "ok_pattern.mtx": (2, 2),
             "dup.mtx": (2, 2)}


@pytest.fixture(scope="module")
def asan_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("asan")
    exe = d / "ingest_oracle_asan"
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
    cmd = [HIPCC, "-std=c++17", "-O1", "-g", "-fopenmp", *san,
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"),
           "-x", "hip", "--cuda-host-only",
           os.path.join(ROOT, "rayleighchebyshevgpgpu_amd", "csrc", "rcg_io.cpp"),
           os.path.join(ROOT, "tests", "cpp", "ingest_oracle_asan.cpp"),
           "-o", str(exe), "-L" + os.path.join(ROOT, "rayleighchebyshevgpgpu_amd"), "-lrcg",
           "-Wl,-rpath," + os.path.join(ROOT, "rayleighchebyshevgpgpu_amd")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + p.stderr[-3000:])
    for name, text in FILES.items():
        (d / name).write_bytes(text.encode("latin-1"))
    return d, exe


def test_parser_and_oracle_under_asan_ubsan(asan_exe):
    d, exe = asan_exe
    # the OpenMP runtime keeps a thread-pool allocation until exit
    supp = d / "lsan.supp"
    supp.write_text("leak:___kmp_allocate\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               LSAN_OPTIONS=f"suppressions={supp}",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    p = subprocess.run([str(exe), str(d), *FILES], capture_output=True, text=True, env=env,
                       timeout=300)
    print(p.stdout)
    print(p.stderr[-4000:])
    assert p.returncode == 0
    assert "Sanitizer" not in p.stderr and "runtime error" not in p.stderr
    rows = {ln.split()[0]: ln.split()[1:] for ln in p.stdout.splitlines()
            if ln.endswith(".mtx") or ".mtx " in ln}
    for name in FILES:
        st = int(rows[name][0])
        if name in EXPECT_OK:
            n, nnz = EXPECT_OK[name]
            assert st == 0 and int(rows[name][1]) == n and int(rows[name][2]) == nnz, name
            assert int(rows[name][3]) == 0, name
        else:
            assert st != 0, name
    assert p.stdout.count("oracle ") == 2
