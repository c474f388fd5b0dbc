"""Timed sub-FFT sizes for single convolutions (fftwpp_amd_conv_create_tuned,
DeviceConvolution(tune=True)): the GPU counterpart of the reference's
optimizer, which times candidate parameters when a plan is made
(fftBase::OptBase, convolve.h:164-236, convolve.cc:239-470).

A tuned plan may run with other sub-FFT sizes than the library's choice; its
results must equal that plan's (itself checked against explicit padding and
direct sums by the parity tests) to 1e-12 (FP64) / 1e-5 (FP32) relative L2.
The choice is cached: a second creation of the same plan reads it back (same
sizes, no timing)."""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # shape, kind, lengths, dtype name
    ((1000,), "complex", None, "c128"),
    ((3000,), "complex", None, "c64"),
    ((40, 33), "complex", None, "c128"),
    ((63, 32), "hermitian", [63, 63], "c128"),
    ((12, 10, 24), "complex", None, "c128"),
    ((600,), "real", None, "f64"),
    ((5,), "complex", None, "c128"),
    ((24, 40), "real", None, "f64"),
]


def _data(torch, gpu, shape, dt, rng):
    if dt.is_complex:
        v = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    else:
        v = rng.uniform(-1, 1, shape)
    return torch.tensor(v, dtype=dt, device=gpu)


@pytest.fixture
def tune_cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("FFTWPP_AMD_TUNE_CACHE", str(path))
    return path


@pytest.mark.parametrize("shape,kind,lengths,dname", CASES)
def test_tuned_plan_equals_library_choice(gpu, tune_cache, shape, kind, lengths, dname):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    dt = {"c128": torch.complex128, "c64": torch.complex64, "f64": torch.float64}[dname]
    rng = np.random.default_rng(sum(shape))
    a, b = _data(torch, gpu, shape, dt, rng), _data(torch, gpu, shape, dt, rng)
    base = DeviceConvolution(shape, kind=kind, dtype=dt, lengths=lengths)
    tuned = DeviceConvolution(shape, kind=kind, dtype=dt, lengths=lengths, tune=True)
    want, got = torch.empty_like(a), torch.empty_like(a)
    base.convolve_out(want, a, b)
    tuned.convolve_out(got, a, b)
    torch.cuda.synchronize()
    w, g = want.cpu().numpy(), got.cpu().numpy()
    err = float(np.linalg.norm((g - w).ravel()) / np.linalg.norm(w.ravel()))
    tol = 1e-5 if dt == torch.complex64 else 1e-12
    print(f"{shape} {kind} {dname}: library {base.geometry()[1][:len(shape)]}, "
          f"tuned {tuned.geometry()[1][:len(shape)]}, rel diff {err:.2e}", flush=True)
    assert err < tol, err


def test_tuned_choice_is_cached(gpu, tune_cache):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    shape, lengths = (127, 64), [127, 127]
    t0 = time.perf_counter()
    first = DeviceConvolution(shape, kind="hermitian", lengths=lengths, tune=True)
    t1 = time.perf_counter()
    second = DeviceConvolution(shape, kind="hermitian", lengths=lengths, tune=True)
    t2 = time.perf_counter()
    assert first.geometry() == second.geometry()
    lines = tune_cache.read_text().splitlines()
    assert len(lines) == 1 and '"key":"c' in lines[0] and '"m":[' in lines[0], lines
    print(f"tuning {1e3 * (t1 - t0):.1f} ms, cached creation {1e3 * (t2 - t1):.1f} ms",
          flush=True)
    assert t2 - t1 < 0.5 * (t1 - t0) + 0.05
    # retune appends a fresh entry
    DeviceConvolution(shape, kind="hermitian", lengths=lengths, tune=True, retune=True)
    assert len(tune_cache.read_text().splitlines()) == 2


def test_tune_excludes_forced_sizes(gpu, tune_cache):
    from fftwpp_amd.torch_conv import DeviceConvolution
    with pytest.raises(ValueError):
        DeviceConvolution((100,), m=[64], tune=True)
    assert not os.path.exists(tune_cache)
