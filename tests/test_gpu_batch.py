"""Batches of independent 1D convolutions of uneven lengths (one plan, one
graph launch) against the oracle: each item must equal the single dealiased
convolution of its pair, which the reference computes with one Convolution
object per length (convolve.h:982-1166, tests/hybridconv.cc)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# lengths that cross every sub-FFT size boundary, plus the extremes
EDGE = [1, 2, 3, 15, 16, 17, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513,
        1000, 1023, 1024, 1025, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 8192, 8193,
        16383, 16384]


def _inputs(lengths, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    fs, gs = [], []
    for L in lengths:
        fs.append(rng.uniform(-1, 1, L) + 1j * rng.uniform(-1, 1, L))
        gs.append(rng.uniform(-1, 1, L) + 1j * rng.uniform(-1, 1, L))
    return fs, gs


def _expected(oracle, f, g, mult="binary"):
    L = len(f)
    if mult == "correlation":
        # h_j = sum_b f_{b+j} conj(g_b), j < L (F0 *= conj F1, convolve.cc:26-110);
        # numpy's correlate gives c_k = sum_n f_{n+k} conj(g_n) at index L-1+k
        return np.correlate(f, g, "full")[L - 1:]
    if L <= 256:
        return oracle.direct_conv(f, g)
    return oracle.explicit_conv(f, g, 1 << (2 * L - 1).bit_length())


def _check_items(oracle, plan, out, fs, gs, tol, mult="binary"):
    got = plan.unpack(out)
    worst = 0.0
    for i, (f, g) in enumerate(zip(fs, gs)):
        e = _expected(oracle, f, g, mult)
        err = oracle.rel_l2(got[i], e)
        worst = max(worst, err)
        assert err < tol, f"item {i} (L={len(f)}, geometry {plan.geometry(i)}): rel L2 {err:.3e}"
    return worst


@pytest.mark.parametrize("graph", [False, True])
def test_batch_edges_fp64(gpu, oracle, graph):
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    lengths = EDGE + [int(v) for v in oracle.batched_lengths(40, 1, 3000, seed=5)]
    fs, gs = _inputs(lengths, 11)
    plan = BatchedConvolution(lengths, dtype=torch.complex128, graph=graph)
    for i, L in enumerate(lengths):
        m, q, N = plan.geometry(i)
        assert N == m * q and N >= 2 * L - 1
    f = plan.pack(fs)
    g = plan.pack(gs)
    out = plan(f, g)
    torch.cuda.synchronize()
    _check_items(oracle, plan, out, fs, gs, 1e-12)


def test_batch_config_M2L_fp64_fp32(gpu, oracle):
    """The batched workload's padding M = 2L on a sample of its lengths, FP64
    to 1e-12 and FP32 to 1e-5 against FP64."""
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    lengths = [int(v) for v in oracle.batched_lengths(256)]
    fs, gs = _inputs(lengths, 12)
    p64 = BatchedConvolution(lengths, M=[2 * L for L in lengths], dtype=torch.complex128)
    out64 = p64(p64.pack(fs), p64.pack(gs))
    p32 = BatchedConvolution(lengths, M=[2 * L for L in lengths], dtype=torch.complex64)
    out32 = p32(p32.pack(fs), p32.pack(gs))
    torch.cuda.synchronize()
    idx = list(range(0, len(lengths), 8))
    sub64 = p64.unpack(out64)
    for i in idx:
        e = _expected(oracle, fs[i], gs[i])
        assert oracle.rel_l2(sub64[i], e) < 1e-12, (i, lengths[i])
    a = out64.cpu().numpy()
    b = out32.cpu().numpy().astype(np.complex128)
    err = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert err < 1e-5, err


def test_batch_graph_replays(gpu, oracle):
    """Replays of the recorded graph read the current contents of the arrays;
    a second set of arrays records its own graph."""
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    lengths = [5, 300, 1024, 77, 2500, 9000, 64]
    plan = BatchedConvolution(lengths, dtype=torch.complex128, graph=True)
    eager = BatchedConvolution(lengths, dtype=torch.complex128, graph=False)
    out = torch.empty(plan.size, dtype=torch.complex128, device=gpu)
    out2 = torch.empty_like(out)
    for seed in range(4):
        fs, gs = _inputs(lengths, 100 + seed)
        f = plan.pack(fs)
        g = plan.pack(gs)
        plan(f, g, out)
        ref = eager(f, g)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        _check_items(oracle, plan, out, fs, gs, 1e-12)
    # new tensors for f, g and out: another graph
    fs, gs = _inputs(lengths, 200)
    f2 = plan.pack(fs)
    g2 = plan.pack(gs)
    for _ in range(3):
        plan(f2, g2, out2)
        torch.cuda.synchronize()
        _check_items(oracle, plan, out2, fs, gs, 1e-12)


def test_batch_tuned(gpu, oracle):
    """The autotuned plan picks valid geometries and still matches."""
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    lengths = [int(v) for v in oracle.batched_lengths(120, 8, 8000, seed=3)]
    fs, gs = _inputs(lengths, 13)
    plan = BatchedConvolution(lengths, M=[2 * L for L in lengths], dtype=torch.complex128,
                              tune=True)
    text = plan.describe()
    assert "m=" in text
    for i, L in enumerate(lengths):
        m, q, N = plan.geometry(i)
        assert N >= 2 * L
    out = plan(plan.pack(fs), plan.pack(gs))
    torch.cuda.synchronize()
    _check_items(oracle, plan, out, fs, gs, 1e-12)


def test_batch_correlation(gpu, oracle):
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    lengths = [1, 7, 200, 257, 1500]
    fs, gs = _inputs(lengths, 14)
    plan = BatchedConvolution(lengths, dtype=torch.complex128, mult="correlation")
    out = plan(plan.pack(fs), plan.pack(gs))
    torch.cuda.synchronize()
    _check_items(oracle, plan, out, fs, gs, 1e-12, mult="correlation")


def test_batch_errors(gpu):
    import torch
    from fftwpp_amd._lib import FFTWPPError
    from fftwpp_amd.batch import BatchedConvolution
    with pytest.raises(FFTWPPError):
        BatchedConvolution([4, 0, 3])
    with pytest.raises(FFTWPPError):
        BatchedConvolution([10], M=[5])
    plan = BatchedConvolution([3, 4])
    f = torch.zeros(7, dtype=torch.complex128, device=gpu)
    g = torch.zeros_like(f)
    with pytest.raises(FFTWPPError):
        plan(f, g, f)
    with pytest.raises(TypeError):
        plan(f.cpu(), g)
    with pytest.raises(ValueError):
        plan(torch.zeros(8, dtype=torch.complex128, device=gpu), g)


@pytest.mark.parametrize("dtype_name,tol", [("complex128", 1e-12), ("complex64", 1e-5)])
def test_batch_workload_every_item(gpu, oracle, dtype_name, tol):
    """The batched workload at full size: 4096 lengths log-uniform in
    [64, 16384] from seed 20261202, M = 2L, uniform[-1, 1] data, an autotuned
    plan replayed from its hipGraph (the measured configuration).  Every item of
    two replays is checked against the oracle (explicit zero padding, FP64)."""
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    dtype = getattr(torch, dtype_name)
    lengths = [int(v) for v in oracle.batched_lengths(4096)]
    fs, gs = _inputs(lengths, 20261206)
    plan = BatchedConvolution(lengths, M=[2 * L for L in lengths], dtype=dtype, tune=True,
                              graph=True)
    f = plan.pack(fs)
    g = plan.pack(gs)
    out = torch.empty_like(f)
    for _ in range(3):  # direct run + recording, then two replays
        out.zero_()
        plan(f, g, out)
        torch.cuda.synchronize()
    worst = _check_items(oracle, plan, out, fs, gs, tol)
    print(f"batched workload {dtype_name}: worst item rel L2 {worst:.3e}", flush=True)


@pytest.mark.parametrize("dtype_name,m,lengths,tol", [
    ("complex128", 4096, [12289, 13000, 16000, 16384, 14001], 1e-12),
    ("complex64", 4096, [12289, 13000, 16000, 16384, 14001], 1e-5),
    ("complex128", 256, [129, 200, 256, 131, 255, 190, 170, 140, 250], 1e-12),
    ("complex128", 2048, [4097, 5000, 6144], 1e-12),
])
def test_batch_rows_sequence_mode(gpu, oracle, dtype_name, m, lengths, tol):
    """The sequence mode of the batched kernel (D = 0: one residue per work
    unit writing its inverse-transformed sequence, then the backward fold over
    the blocks in a second launch) called through the kernel-layer ABI for
    items that share (m, q), checked item by item against explicit padding."""
    import ctypes
    import torch
    from fftwpp_amd import _lib
    lib = _lib.lib()
    dtype = getattr(torch, dtype_name)
    dcode = 0 if dtype == torch.complex128 else 1

    def res(L, M):
        p = -(-L // m)
        P = 1 if p == 2 else p
        n = -(-M // (P * m))
        return P * n
    qs = {res(L, 2 * L) for L in lengths}
    assert len(qs) == 1
    q = qs.pop()
    fs, gs = _inputs(lengths, 31 + m)
    offs = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    items = np.zeros((len(lengths), 3), dtype=np.int64)
    items[:, 0] = offs[:-1]
    items[:, 1] = lengths
    items[:, 2] = offs[:-1]
    d = ctypes.c_void_p()
    _lib.check(lib.fftwpp_amd_dim_create(ctypes.byref(d), max(lengths), m, q, 0))
    try:
        tps = lib.fftwpp_amd_batch_rows_tps(d, dcode)
        assert tps > 0
        f = torch.tensor(np.concatenate(fs), dtype=dtype, device=gpu)
        g = torch.tensor(np.concatenate(gs), dtype=dtype, device=gpu)
        out = torch.zeros_like(f)
        it = torch.tensor(items.ravel(), device=gpu)
        part = torch.empty(len(lengths) * q * m, dtype=dtype, device=gpu)
        threads = max(tps, 256) if tps <= 64 else tps
        units = len(lengths) * q
        nblocks = -(-units // 8) * 8 if tps > 64 else -(-units // (threads // tps))
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.fftwpp_amd_batch_rows(d, dcode, 1, it.data_ptr(), len(lengths),
                                             f.data_ptr(), g.data_ptr(), out.data_ptr(),
                                             1, threads, nblocks, 0, 0, part.data_ptr(),
                                             int(offs[-1]), 1.0 / (m * q), stream))
        torch.cuda.synchronize()
        h = out.cpu().numpy()
        for i, L in enumerate(lengths):
            e = _expected(oracle, fs[i], gs[i])
            err = oracle.rel_l2(h[offs[i]:offs[i + 1]].astype(np.complex128), e)
            assert err < tol, f"item {i} (L={L}, m={m}, q={q}): rel L2 {err:.3e}"
    finally:
        lib.fftwpp_amd_dim_destroy(d)


def test_batch_tune_cache(gpu, oracle, tmp_path, monkeypatch):
    """The autotuner's choices persist (the analogue of FFTW wisdom): the
    second tuned plan of the same batch is built from the stored entry
    without timing (well under 50 ms), has the same classes and tiles, and
    still matches the oracle; another batch appends its own entry."""
    import time
    import torch
    from fftwpp_amd.batch import BatchedConvolution
    path = tmp_path / "tune.json"
    monkeypatch.setenv("FFTWPP_AMD_TUNE_CACHE", str(path))
    lengths = [int(v) for v in oracle.batched_lengths(256, 64, 16384, seed=9)]
    M = [2 * L for L in lengths]
    t0 = time.perf_counter()
    p1 = BatchedConvolution(lengths, M=M, dtype=torch.complex128, tune=True)
    t1 = time.perf_counter()
    assert path.exists() and len(path.read_text().splitlines()) == 1
    p2 = BatchedConvolution(lengths, M=M, dtype=torch.complex128, tune=True)
    t2 = time.perf_counter()
    print(f"tuned creation {t1 - t0:.3f} s, from the cache {t2 - t1:.3f} s", flush=True)
    assert t2 - t1 < 0.05

    def tiles(plan):
        return [line.split(" ms=")[0] for line in plan.describe().splitlines()]
    assert tiles(p2) == tiles(p1)
    assert len(path.read_text().splitlines()) == 1
    fs, gs = _inputs(lengths, 19)
    out = p2(p2.pack(fs), p2.pack(gs))
    torch.cuda.synchronize()
    got = p2.unpack(out)
    for i in range(0, len(lengths), 16):
        assert oracle.rel_l2(got[i], _expected(oracle, fs[i], gs[i])) < 1e-12
    other = lengths[:100]
    BatchedConvolution(other, M=M[:100], dtype=torch.complex128, tune=True)
    assert len(path.read_text().splitlines()) == 2
