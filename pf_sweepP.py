import os, ctypes
import numpy as np, torch
here = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng(0)
n, B = 224, 2048
A = rng.standard_normal((B, n, n)).astype(np.float32) * 0.3
A = np.triu(A, 1); A = A - np.transpose(A, (0, 2, 1))
t = torch.tensor(A, device="cuda")
ref = None
for P in (6, 8, 10, 12):
    lib = ctypes.CDLL(os.path.join(here, f"libpfprobeP{P}.so"))
    lib.probe_slogpf.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.probe_set.argtypes = [ctypes.c_void_p]
    for nt in (512, 1024):
        prof = torch.zeros((B, 8), dtype=torch.int64, device="cuda")
        sg = torch.empty(B, device="cuda"); la = torch.empty(B, dtype=torch.float64, device="cuda")
        assert lib.probe_set(prof.data_ptr()) == 0
        for rep in range(2):
            prof.zero_(); torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True); ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            assert lib.probe_slogpf(nt, B, n, t.data_ptr(), sg.data_ptr(), la.data_ptr()) == 0
            ev1.record(); torch.cuda.synchronize()
        p = prof.cpu().numpy().astype(np.float64).mean(axis=0)
        l = la.cpu().numpy()
        if ref is None: ref = l
        print(f"P={P} nt={nt}: {ev0.elapsed_time(ev1)*1e3/B:.3f} us/matrix; per step p1 {p[0]/112:.0f} p2 {p[1]/112:.0f}; trailing {p[2]:.0f}; max|dla| vs first {np.abs(l-ref).max():.2e}", flush=True)
