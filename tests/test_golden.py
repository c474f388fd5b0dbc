"""Golden vectors (tests/golden/golden.json, written by make_golden.py with
exact integer arithmetic from the reference's own ramp inputs and direct
sums, tests/direct.{h,cc}).  The oracle's restatement must reproduce every
input and output (CPU); the library must reproduce the outputs through its
public plans within 1e-12 relative L2 in FP64 (GPU)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "golden.json")))["cases"]


def arr(case, key):
    v = np.array(case[key], dtype=np.float64)
    return (v[:, 0] + 1j * v[:, 1]).reshape(case["shape"])


def ids(c):
    return c["kind"] + "x".join(str(s) for s in c["shape"])


def test_golden_file_is_reproducible(tmp_path):
    """The fixtures are what make_golden.py writes (exact integers)."""
    import importlib.util
    import shutil
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    mg.HERE = str(tmp_path)
    mg.main()
    assert (tmp_path / "golden.json").read_text() == open(os.path.join(HERE, "golden", "golden.json")).read()
    shutil.rmtree(tmp_path)


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_oracle_matches_golden(oracle, case):
    f, g, h = arr(case, "f"), arr(case, "g"), arr(case, "h")
    kind, shape = case["kind"], case["shape"]
    if kind in ("conv1", "convC1"):
        assert np.array_equal(f, oracle.ramp1(shape[0], 0)) and np.array_equal(g, oracle.ramp1(shape[0], 1))
        got = (oracle.direct_conv if kind == "conv1" else oracle.direct_convC)(f, g)
    elif kind == "convh1":
        assert np.array_equal(f, oracle.rampH1(shape[0], 0))
        got = oracle.direct_convh(f, g)
    elif kind == "conv2":
        assert np.array_equal(f, oracle.ramp2(*shape, 0))
        got = oracle.direct_conv2(f, g)
    elif kind == "convh2":
        assert np.array_equal(f, oracle.ramph2(*shape, 0)) and np.array_equal(g, oracle.ramph2(*shape, 1))
        got = oracle.direct_convh2(f, g)
    else:
        assert np.array_equal(f, oracle.ramp3(*shape, 0))
        got = oracle.direct_conv3(f, g)
    assert np.array_equal(got, h), kind


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=ids)
def test_library_matches_golden(gpu, case):
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    f, g, h = arr(case, "f"), arr(case, "g"), arr(case, "h")
    kind = {"conv1": "complex", "convC1": "centered", "convh1": "hermitian", "conv2": "complex",
            "convh2": "hermitian", "conv3": "complex"}[case["kind"]]
    lengths = case.get("L")
    if isinstance(lengths, int):
        lengths = [lengths]
    plan = DeviceConvolution(f.shape, kind=kind, lengths=lengths)
    a = torch.tensor(f, device=gpu)
    b = torch.tensor(g, device=gpu)
    plan.convolve_(a, b)
    torch.cuda.synchronize()
    got = a.cpu().numpy()
    err = float(np.linalg.norm((got - h).ravel()) / max(np.linalg.norm(h.ravel()), 1e-300))
    assert err < 1e-12, f"{case['kind']} {case['shape']}: rel L2 {err:.3e}"
