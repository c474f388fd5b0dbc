"""fftwpp_amd_batch_rows_noalloc (partial results in the library's fixed
device array, for callers that cannot allocate) against fftwpp_amd_batch_rows
(a caller buffer): the same
kernels on the same inputs must give bit-identical outputs, on two calls in a
row, the second with fewer items; and a launch whose partial results exceed
the 16 MiB array is refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH = 16 << 20


def _items(lengths):
    offs = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    items = np.zeros((len(lengths), 3), dtype=np.int64)
    items[:, 0] = offs[:-1]
    items[:, 1] = lengths
    items[:, 2] = offs[:-1]
    return offs, items


def _run(lib, fn, d, dcode, dtype, gpu, lengths, fs, gs, D, q, m, part=None):
    import torch
    from fftwpp_amd import _lib
    offs, items = _items(lengths)
    f = torch.tensor(np.concatenate(fs), dtype=dtype, device=gpu)
    g = torch.tensor(np.concatenate(gs), dtype=dtype, device=gpu)
    out = torch.zeros_like(f)
    it = torch.tensor(items.ravel(), device=gpu)
    tps = lib.fftwpp_amd_batch_rows_tps(d, dcode)
    threads = tps
    G = q if D == 0 else -(-q // D)
    units = len(lengths) * G
    nblocks = -(-units // 8) * 8
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [d, dcode, 1, it.data_ptr(), len(lengths), f.data_ptr(), g.data_ptr(),
            out.data_ptr(), 1, threads, nblocks, 0, D]
    if part is not None:
        args.append(part.data_ptr())
    args += [int(offs[-1]), 1.0 / (m * q), stream]
    _lib.check(fn(*args))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("dtype_name,D", [("complex128", 0), ("complex128", 4), ("complex64", 0),
                                          ("complex64", 2)])
def test_noalloc_matches_batch_rows(gpu, dtype_name, D):
    import torch
    from fftwpp_amd import _lib
    lib = _lib.lib()
    dtype = getattr(torch, dtype_name)
    dcode = 0 if dtype == torch.complex128 else 1
    m, q = 4096, 8
    rng = np.random.default_rng(5 + D)
    calls = [[12289, 13000, 16000, 16384, 14001, 15000], [16384, 12500]]
    d = ctypes.c_void_p()
    _lib.check(lib.fftwpp_amd_dim_create(ctypes.byref(d), 16384, m, q, 0))
    try:
        for lengths in calls:
            fs = [rng.uniform(-1, 1, L) + 1j * rng.uniform(-1, 1, L) for L in lengths]
            gs = [rng.uniform(-1, 1, L) + 1j * rng.uniform(-1, 1, L) for L in lengths]
            total = sum(lengths)
            words = len(lengths) * q * m if D == 0 else -(-q // D) * total
            part = torch.empty(words, dtype=dtype, device=gpu)
            ref = _run(lib, lib.fftwpp_amd_batch_rows, d, dcode, dtype, gpu, lengths, fs, gs,
                       D, q, m, part)
            got = _run(lib, lib.fftwpp_amd_batch_rows_noalloc, d, dcode, dtype, gpu, lengths,
                       fs, gs, D, q, m)
            print(f"noalloc {dtype_name} D={D}: {len(lengths)} items, "
                  f"{words * (16 if dcode == 0 else 8) / 2**20:.2f} MiB of scratch", flush=True)
            assert np.array_equal(got, ref)
    finally:
        lib.fftwpp_amd_dim_destroy(d)


def test_noalloc_refuses_oversized(gpu):
    import torch
    from fftwpp_amd import _lib
    lib = _lib.lib()
    d = ctypes.c_void_p()
    _lib.check(lib.fftwpp_amd_dim_create(ctypes.byref(d), 16384, 4096, 8, 0))
    try:
        count = SCRATCH // (16 * 8 * 4096) + 1       # sequences of count items overflow
        rc = lib.fftwpp_amd_batch_rows_noalloc(d, 0, 1, 1, count, 1, 1, 1, 1, 512, 8, 0, 0,
                                               count * 16384, 1.0, None)
        assert rc == -1
        assert b"16 MiB" in lib.fftwpp_amd_last_error()
        assert torch.cuda.is_available()
    finally:
        lib.fftwpp_amd_dim_destroy(d)
