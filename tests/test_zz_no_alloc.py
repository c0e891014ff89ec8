"""The allocation-free qr_compact! entry point (mak_qr_compact_noalloc_d)
against the workspace entry point (mak_qr_compact_d): two calls in a row on
the shared static scratch, the second with a smaller input, each within 1e-14
(relative Frobenius) of the workspace entry point's Q and R on the same input
(the same kernels run on the same data) and within 1e-12 of the oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

import mak_mi355x as mak
from mak_mi355x import _lib
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _call_workspace(lib, A, dev):
    m, n = A.shape
    At = mak.colmajor(torch.from_numpy(A).to(dev))
    Q, R = mak.initialize_output_qr(At, False)
    nb = C.c_size_t(0)
    assert lib.mak_qr_workspace_size(0, 0, m, n, C.byref(nb)) == 0
    ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    code = lib.mak_qr_compact_d(s, m, n, C.c_void_p(At.data_ptr()), m, C.c_void_p(Q.data_ptr()), m,
                                C.c_void_p(R.data_ptr()), min(m, n), 1, C.c_void_p(ws.data_ptr()), nb.value)
    torch.cuda.synchronize()
    assert code == 0
    return Q.cpu().numpy(), R.cpu().numpy()


def _call_noalloc(lib, A, dev):
    m, n = A.shape
    At = mak.colmajor(torch.from_numpy(A).to(dev))
    Q, R = mak.initialize_output_qr(At, False)
    s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    code = lib.mak_qr_compact_noalloc_d(s, m, n, C.c_void_p(At.data_ptr()), m, C.c_void_p(Q.data_ptr()), m,
                                        C.c_void_p(R.data_ptr()), min(m, n), 1)
    torch.cuda.synchronize()
    assert code == 0
    return Q.cpu().numpy(), R.cpu().numpy()


def test_noalloc_qr_matches_workspace_entry_point_twice(dev):
    lib = _lib.load()
    for (m, n), seed in (((1024, 512), 5), ((300, 70), 6)):   # the second call is smaller
        A = O.gaussian(m, n, np.float64, seed=seed)
        print(f"noalloc qr_compact {m}x{n}", flush=True)
        Qn, Rn = _call_noalloc(lib, A, dev)
        Qw, Rw = _call_workspace(lib, A, dev)
        assert rel(Qn, Qw) <= 1e-14 and rel(Rn, Rw) <= 1e-14, (rel(Qn, Qw), rel(Rn, Rw))
        Qo, Ro = O.qr_compact(A)
        assert rel(Qn, Qo) < 1e-12 and rel(Rn, Ro) < 1e-12


def test_noalloc_qr_rejects_problems_beyond_the_scratch(dev):
    lib = _lib.load()
    s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    null = C.c_void_p(0)
    # 8192 x 1024 needs ~435 MiB of workspace: rejected before any launch
    assert lib.mak_qr_compact_noalloc_d(s, 8192, 1024, null, 8192, null, 8192, null, 1024, 1) == 7
    assert lib.mak_qr_compact_noalloc_d(s, -1, 4, null, 1, null, 1, null, 1, 1) == -2
