"""lai_dev_eri_quartets_device_noalloc (the entry point that allocates no
device memory: its quartet descriptors live in the library's static staging
array) against lai_dev_eri_quartets_device, bit for bit, on two calls in a
row, the second with fewer quartets; and the size limit as a host-side
error."""
import ctypes

import numpy as np
import pytest

from libaccint import _cabi, systems
import libaccint as lai

pytestmark = pytest.mark.gpu

P64 = ctypes.POINTER(ctypes.c_int64)
PI = ctypes.POINTER(ctypes.c_int)


def blocks(L, fn, hb, q, stream, torch):
    n = len(q)
    offs = np.zeros(n + 1, dtype=np.int64)
    _cabi.check(fn(hb, n, q.ctypes.data_as(PI), None, offs.ctypes.data_as(P64), None))
    out = torch.full((max(int(offs[-1]), 1),), float("nan"), dtype=torch.float64, device="cuda")
    offs2 = np.zeros(n + 1, dtype=np.int64)
    _cabi.check(fn(hb, n, q.ctypes.data_as(PI), ctypes.c_void_p(out.data_ptr()), offs2.ctypes.data_as(P64),
                   ctypes.c_void_p(stream.cuda_stream)))
    assert np.array_equal(offs, offs2)
    torch.cuda.synchronize()
    return out.cpu().numpy()[:int(offs[-1])], offs


def test_noalloc_blocks_equal_allocating_entry_point():
    torch = pytest.importorskip("torch")
    atoms = systems.water_cluster(2, 8.0, seed=3)
    b = lai.basis_set("cc-pvtz", atoms)
    e = lai.Engine(b)
    L = _cabi.lib()
    hb = ctypes.c_void_p(e.device_basis_handle())
    s = torch.cuda.Stream()
    rng = np.random.default_rng(21)
    for n in (5000, 700):  # the second call is smaller: stale descriptors must not leak in
        print("no-alloc ERI blocks: %d quartets" % n, flush=True)
        q = rng.integers(0, b.n_shells(), size=(n, 4)).astype(np.int32)
        ref, offs = blocks(L, L.lai_dev_eri_quartets_device, hb, q, s, torch)
        got, offs2 = blocks(L, L.lai_dev_eri_quartets_device_noalloc, hb, q, s, torch)
        assert np.array_equal(offs, offs2)
        assert not np.any(np.isnan(got))
        assert np.array_equal(got, ref)


def test_noalloc_rejects_too_many_quartets():
    pytest.importorskip("torch")
    atoms = systems.h2o_sto3g_geometry()
    b = lai.basis_set("sto-3g", atoms)
    e = lai.Engine(b)
    L = _cabi.lib()
    hb = ctypes.c_void_p(e.device_basis_handle())
    n = _cabi.LAI_NOALLOC_MAX_QUARTETS + 1
    q = np.zeros((n, 4), dtype=np.int32)
    offs = np.zeros(n + 1, dtype=np.int64)
    rc = L.lai_dev_eri_quartets_device_noalloc(hb, n, q.ctypes.data_as(PI), None, offs.ctypes.data_as(P64), None)
    assert rc == -1
    assert b"staging limit" in L.lai_dev_last_error()
