"""The Pfaffian image kernel's column-major form (image_pf_cm in
neuralhubbard_amd/csrc/nh_image_pf.h: occupied-mode gather under a group element and a hop,
feature staging, the full skew pairing square built in the workspace, and the lazy-pivot
blocked Parlett-Reid pf_lazy of nh_pf_lazy.h) runs whole on the CPU under a host emulation of
its 256-thread workgroup (tests/host/image_pf_host.cpp: one OS thread per GPU thread, the wave
builtins as exchanges between barriers of 64 lanes), built with AddressSanitizer and the
alignment check.  Every buffer is exactly as large as the library makes it (the image's
workspace slice, the launch's dynamic LDS), so each access is shown in bounds; workspace and LDS
start as NaN, so no value that was never written reaches the result; readfirstlane only sees
wave-uniform values.  (sign, log|Pf|) matches the FP64 oracle's image amplitude
(oracle.pairing_occ of the mapped modes, oracle.slogpf, the reversal sign)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("imgh") / "image_pf_host")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,alignment", "-fno-sanitize-recover=all",
           "-fno-strict-aliasing", "-pthread", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "neuralhubbard_amd", "csrc"),
           os.path.join(HERE, "host", "image_pf_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.fixture(scope="module")
def harness_tsan(tmp_path_factory):
    """The same harness under ThreadSanitizer: the emulated workgroup's threads synchronise only
    through __syncthreads (a std::barrier) and the wave exchanges, so an access to LDS or to the
    workspace that a barrier does not order against another thread's write is reported."""
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("imgt") / "image_pf_tsan")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=thread", "-fno-strict-aliasing", "-pthread",
           "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "neuralhubbard_amd", "csrc"),
           os.path.join(HERE, "host", "image_pf_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _case(LX, LY, n_up, n_dn, D, seed, f0=False):
    rng = np.random.default_rng(seed)
    N = LX * LY
    s = -np.ones(2 * N, np.int8)
    s[rng.choice(N, n_up, replace=False)] = 1
    s[N + rng.choice(N, n_dn, replace=False)] = 1
    Y = rng.standard_normal((2 * N, D)).astype(np.float32)
    W = rng.standard_normal((3, N, D)) / np.sqrt(D)
    Wt = O.build_wtilde(W, 0.55, LX, LY).astype(np.float32)
    F0 = None
    if f0:
        M = rng.standard_normal((2 * N, 2 * N)).astype(np.float32) / np.sqrt(N)
        F0 = ((M - M.T) / 2).astype(np.float32)
    return s, Y, Wt, F0


def _expected(LX, LY, s, Y, Wt, F0, g, hop):
    s = s.copy()
    if hop is not None:
        s[hop[0]], s[hop[1]] = -1, 1
    perm = O.group_perm(g, LX, LY)
    modes = perm[np.flatnonzero(s > 0)]
    F = O.pairing_occ(Y.astype(np.float64), modes, Wt.astype(np.float64), LX, LY)
    if F0 is not None:
        F = F + F0.astype(np.float64)[np.ix_(modes, modes)]
    sg, la = O.slogpf(F)
    ne = len(modes)
    return sg * (-1) ** (ne * (ne - 1) // 2), la


def _run(exe, tmp_path, LX, LY, n_up, n_dn, D, P, g, s, Y, Wt, F0, hop, raw=False):
    N = LX * LY
    hp = -1 if hop is None else (hop[0] | (hop[1] << 16))
    head = np.array([N, LX, LY, n_up + n_dn, n_up, D, g, hp, int(F0 is not None), P], np.int32)
    path = str(tmp_path / f"case_{LX}x{LY}_{D}_{P}_{g}.bin")
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(s.astype(np.int8).tobytes())
        f.write(Y.astype(np.float32).tobytes())
        f.write(Wt.astype(np.float32).tobytes())
        if F0 is not None:
            f.write(F0.astype(np.float32).tobytes())
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=900)
    if raw:
        return r
    assert r.returncode == 0, r.stderr[-3000:]
    sg, la = r.stdout.split()
    return float(sg), float(la)


# (LX, LY, n_up, n_dn, D, P, g, hop?, F0?): the 16x16 workload (n = 224, d = 16) under three group
# elements and a hop, the released states' d = 32, 8-step panels, a base matrix F0, an odd number
# of 4-row column groups (n % 4 == 2: the mirrored 16-byte pieces reach the padding rows), and a
# small size
CASES = [
    (16, 16, 112, 112, 16, 16, 0, False, False),
    (16, 16, 112, 112, 16, 16, 5, True, False),
    (16, 16, 112, 112, 16, 16, 7, False, False),
    (16, 16, 112, 112, 32, 16, 3, True, False),
    (16, 16, 112, 112, 16, 8, 6, False, False),
    (10, 10, 44, 44, 8, 16, 1, True, True),
    (9, 10, 40, 38, 16, 16, 4, False, False),
    (12, 12, 36, 36, 16, 16, 2, True, False),
]


@pytest.mark.parametrize("LX,LY,n_up,n_dn,D,P,g,hop,f0", CASES)
def test_image_pf_cm_host_matches_oracle(harness, tmp_path, LX, LY, n_up, n_dn, D, P, g, hop, f0):
    seed = 1000 * LX + 10 * D + g
    s, Y, Wt, F0 = _case(LX, LY, n_up, n_dn, D, seed, f0)
    N = LX * LY
    h = None
    if hop:  # an occupied up mode to an empty neighbour site
        occ = np.flatnonzero(s[:N] > 0)
        for m in occ:
            x, y = divmod(int(m), LY)
            to = x * LY + (y + 1) % LY
            if s[to] < 0:
                h = (int(m), to)
                break
        assert h is not None
    sg, la = _run(harness, tmp_path, LX, LY, n_up, n_dn, D, P, g, s, Y, Wt, F0, h)
    rs, rl = _expected(LX, LY, s, Y, Wt, F0, g, h)
    assert np.isfinite(la), "a value never written reached the result"
    assert sg == rs
    assert abs(la - rl) < 1e-4  # FP32 elimination: <= 3.4e-5 measured


def _first_hop(s, N, LY):
    occ = np.flatnonzero(s[:N] > 0)
    for m in occ:
        x, y = divmod(int(m), LY)
        to = x * LY + (y + 1) % LY
        if s[to] < 0:
            return int(m), to
    return None


@pytest.mark.parametrize("LX,LY,n_up,n_dn,D,P,g,hop,f0", [
    (16, 16, 112, 112, 16, 16, 5, True, False),
    (16, 16, 112, 112, 32, 16, 3, False, False),
    (10, 10, 44, 44, 8, 16, 1, False, True),
    (9, 10, 40, 38, 16, 8, 4, True, False),
])
def test_image_pf_cm_host_no_data_race(harness_tsan, tmp_path, LX, LY, n_up, n_dn, D, P, g, hop, f0):
    """No data race in the image kernel's body: the pairing build, the workspace square and every
    pivot step and trailing update of pf_lazy (each step's LDS exchanges, the two buffers of the
    compacted complements), with the result still the oracle's; d = 16 (the synthetic 16x16
    workload) and d = 32 (the released states)."""
    s, Y, Wt, F0 = _case(LX, LY, n_up, n_dn, D, 1000 * LX + 10 * D + g, f0)
    h = _first_hop(s, LX * LY, LY) if hop else None
    r = _run(harness_tsan, tmp_path, LX, LY, n_up, n_dn, D, P, g, s, Y, Wt, F0, h, raw=True)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stderr[-3000:]
    sg, la = (float(x) for x in r.stdout.split())
    rs, rl = _expected(LX, LY, s, Y, Wt, F0, g, h)
    assert sg == rs
    assert abs(la - rl) < 1e-4
