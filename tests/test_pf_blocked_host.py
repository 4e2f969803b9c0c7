"""The blocked Pfaffian pf_blocked2 (neuralhubbard_amd/csrc/nh_pf_blocked.h) and the blocked LU
lu_blocked (nh_lu_blocked.h) run on the CPU under a host emulation of their workgroup
(tests/host/pf_blocked_host.cpp: one OS thread per GPU thread, the wave builtins as exchanges
between barriers of 64 lanes), built with AddressSanitizer and the alignment check: every access
stays inside the matrix and the scratch the kernels allot, no value read from scratch that was
never written reaches the result (the scratch starts as NaN), readfirstlane only sees
wave-uniform values, and (sign, log|Pf|) matches the FP64 oracle, (sign, log|det|) LAPACK.
Pfaffian: the panel widths in use (8 steps, 4 in the LDS form above n = 240) and wider ones (12,
16), at the sizes of the image kernels (n = 224 for 16x16 at 112 + 112 electrons; n = 896 for
32x32 at 448 + 448 with 1024 threads and 10-bit row keys, the form of the 32x32 image kernel and
of nh_slogpf_ws) and the standalone nh_slogpf form (1024 threads, 8-bit row keys); the lazy-pivot
pf_lazy (nh_pf_lazy.h) of nh_slogpf_ws's column form with 256 and 512 threads, the latter also
under ThreadSanitizer.  LU: the BCS kernels' 16-column panels at n = 100 to 448 (the 32x32 BCS
state)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CASES = [(256, 8, 224), (256, 4, 130), (256, 12, 224), (256, 16, 256), (256, 12, 66), (1024, 8, 224), (1024, 8, 896),
         (1024, 4, 256)]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("pfh") / "pf_blocked_host")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,alignment", "-fno-sanitize-recover=all",
           "-fno-strict-aliasing", "-pthread", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "neuralhubbard_amd", "csrc"),
           os.path.join(HERE, "host", "pf_blocked_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(exe, tmp_path, NT, P, A):
    n = A.shape[0]
    i, j = np.tril_indices(n, -1)  # row-major strict lower triangle: row i at i (i - 1) / 2
    path = str(tmp_path / f"tri_{NT}_{P}_{n}.bin")
    A[i, j].astype(np.float32).tofile(path)
    r = subprocess.run([exe, str(NT), str(P), str(n), path], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    sg, la = r.stdout.split()
    return float(sg), float(la)


@pytest.mark.parametrize("NT,P,n", CASES)
def test_pf_blocked_host_matches_oracle(harness, tmp_path, NT, P, n):
    rng = np.random.default_rng(1000 * P + n)
    A = rng.standard_normal((n, n)).astype(np.float32)
    A = A - A.T
    sg, la = _run(harness, tmp_path, NT, P, A)
    rs, rl = O.slogpf(A.astype(np.float64))
    assert np.isfinite(la), "a NaN scratch value reached the result"
    assert sg == rs
    assert abs(la - rl) < 1e-4 * n  # FP32 elimination, FP64 log-sum


def test_pf_blocked_host_pins_a_known_pfaffian(harness, tmp_path):
    """Block-diagonal [[0, d_k], [-d_k, 0]] blocks, rows and columns permuted at random: Pf is
    sgn(perm) prod d_k, a value fixed without the oracle, reached only through pivoting."""
    n = 130
    rng = np.random.default_rng(7)
    d = rng.uniform(0.5, 2.0, n // 2) * rng.choice([-1.0, 1.0], n // 2)
    B = np.zeros((n, n))
    B[np.arange(0, n, 2), np.arange(1, n, 2)] = d
    B -= B.T
    perm = rng.permutation(n)
    A = np.zeros_like(B)
    A[np.ix_(perm, perm)] = B  # A[perm[i]][perm[j]] = B[i][j], Pf A = sgn(perm) Pf B
    sgn_perm = np.linalg.det(np.eye(n)[perm])
    sg, la = _run(harness, tmp_path, 256, 8, A.astype(np.float32))
    assert sg == np.sign(sgn_perm) * np.sign(np.prod(d))
    assert abs(la - np.sum(np.log(np.abs(d.astype(np.float32).astype(np.float64))))) < 1e-4


def _run_lu(exe, tmp_path, NT, A, keys="lu"):
    n = A.shape[0]
    path = str(tmp_path / f"mat_{keys}_{NT}_{n}.bin")
    np.ascontiguousarray(A, dtype=np.float32).tofile(path)
    r = subprocess.run([exe, keys, str(NT), str(n), path], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    sg, la = r.stdout.split()
    return float(sg), float(la)


@pytest.mark.parametrize("NT,n,keys", [(256, 100, "lu"), (256, 256, "lu"), (512, 130, "lu"), (512, 448, "lu"),
                                        (512, 112, "lu8"), (512, 192, "lu8"), (256, 256, "lu8")])
def test_lu_blocked_host_matches_lapack(harness, tmp_path, NT, n, keys):
    """lu_blocked (the BCS determinant: 16-column panels, rows kept in place) with 9-bit row keys
    (the workspace form, up to n = 448 of the 32x32 BCS state) and with the 8-bit keys of the
    LDS form and nh_slogdet (n = 112 of the 16x16 BCS workload, 192, 256), against LAPACK's
    slogdet in FP64."""
    rng = np.random.default_rng(n + NT + (7 if keys == "lu8" else 0))
    A = rng.standard_normal((n, n)).astype(np.float32)
    sg, la = _run_lu(harness, tmp_path, NT, A, keys)
    rs, rl = np.linalg.slogdet(A.astype(np.float64))
    assert np.isfinite(la), "a NaN scratch value reached the result"
    assert sg == rs
    assert abs(la - rl) < 1e-4 * n


def test_lu_blocked_host_pins_a_permuted_diagonal(harness, tmp_path):
    """det of a row-permuted diagonal matrix is sgn(perm) prod d: reached only through pivoting,
    and its sign only through the parity count of the pivot rows."""
    n = 200
    rng = np.random.default_rng(11)
    d = rng.uniform(0.5, 2.0, n) * rng.choice([-1.0, 1.0], n)
    perm = rng.permutation(n)
    A = np.zeros((n, n))
    A[perm, np.arange(n)] = d  # column j holds d_j in row perm[j]
    sg, la = _run_lu(harness, tmp_path, 256, A.astype(np.float32))
    assert sg == np.linalg.det(np.eye(n)[perm]) * np.sign(np.prod(d))
    assert abs(la - np.sum(np.log(np.abs(d.astype(np.float32).astype(np.float64))))) < 1e-4


@pytest.fixture(scope="module")
def harness_tsan(tmp_path_factory):
    """The harness under ThreadSanitizer (no dummy slots in pf_lazy: any report is a race)."""
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("pft") / "pf_blocked_tsan")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=thread", "-fno-strict-aliasing", "-pthread",
           "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "neuralhubbard_amd", "csrc"),
           os.path.join(HERE, "host", "pf_blocked_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run_lazy(exe, tmp_path, NT, A, mode="lazy"):
    n = A.shape[0]
    i, j = np.tril_indices(n, -1)
    path = str(tmp_path / f"{mode}_{NT}_{n}.bin")
    A[i, j].astype(np.float32).tofile(path)
    return subprocess.run([exe, mode, str(NT), str(n), path], capture_output=True, text=True, timeout=600)


@pytest.mark.parametrize("mode,NT,n", [("lazy", 256, 100), ("lazy", 256, 224), ("lazy12", 512, 258),
                                       ("lazy12", 512, 352), ("lazy12", 512, 512), ("lazy", 512, 460)])
def test_pf_lazy_host_matches_oracle(harness, tmp_path, mode, NT, n):
    """pf_lazy on the column-major square of nh_slogpf_ws's column form, in the two forms the entry
    runs: 256 threads, 8-bit row keys, 16-step panels (100 <= n <= 256); 512 threads, 10-bit keys,
    12-step panels (256 < n <= 352); and at n = 460 / 512 with 512 threads (measured, not used:
    the packed triangle is faster there in the low-rank hop ratios' batches), against the FP64
    oracle."""
    rng = np.random.default_rng(3000 + n)
    A = (rng.standard_normal((n, n)) * 0.3).astype(np.float32)
    A = A - A.T
    r = _run_lazy(harness, tmp_path, NT, A, mode)
    assert r.returncode == 0, r.stderr[-3000:]
    sg, la = (float(x) for x in r.stdout.split())
    rs, rl = O.slogpf(A.astype(np.float64))
    assert np.isfinite(la), "a NaN value reached the result"
    assert sg == rs
    assert abs(la - rl) < 1e-4 * n


def test_pf_lazy_host_no_data_race(harness_tsan, tmp_path):
    """The 512-thread pf_lazy (eight waves: the step's double-buffered keys, masks and smallest
    rows, the column buffer written by candidate rows only) shows no data race at n = 460."""
    rng = np.random.default_rng(460)
    A = (rng.standard_normal((460, 460)) * 0.3).astype(np.float32)
    A = A - A.T
    r = _run_lazy(harness_tsan, tmp_path, 512, A)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stderr[-3000:]
    sg, la = (float(x) for x in r.stdout.split())
    rs, rl = O.slogpf(A.astype(np.float64))
    assert sg == rs and abs(la - rl) < 1e-4 * 460
