#!/usr/bin/env python3
"""bench.py -- 3D complex dealiased convolution 256^3 -> 512^3 on one MI355X.

One step = one implicitly dealiased convolution h = f (*) g (Convolution3 of
two 256^3 complex arrays padded to 512^3, multBinary), with f and g resident
in HBM and the result written out of place, FP64 (the drop-in precision).
The same workload is then timed in FP32 (FP64-generated twiddles) and its
relative L2 error against the FP64 result is reported.

Prints ONE JSON line.  Per-kernel durations come from HIP events recorded by
the library after each of its launches on the benchmark stream.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "3D conv/s at 256³→512³ FP32 vs FP64 (rel. L2 err) + HBM GB/s % of peak"
PEAK_GBS = 8000.0  # MI355X HBM3E
SEED = 20261202


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run a 512 MiB device-to-device copy first (the known-size "
                         "reference for HBM traffic counters)")
    return ap.parse_args()


def inputs(L, seed=SEED):
    """Uniform[-1,1] real then imaginary parts, f then g, PCG64(seed)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(2):
        re = rng.uniform(-1.0, 1.0, size=(L, L, L))
        im = rng.uniform(-1.0, 1.0, size=(L, L, L))
        out.append(re + 1j * im)
    return out


# Algorithmic HBM bytes of each launch of the Convolution3 pipeline, in units
# of E*L^3 (E = bytes per complex word), for q = 2 residues per dimension:
# reads of its inputs plus writes of its outputs.
STAGE_UNITS = {
    "x forward f": 1 + 2, "x forward g": 1 + 2,
    "y forward f": 2 + 4, "y forward g": 2 + 4,
    "z rows": 8 + 4, "y backward": 4 + 2, "x backward": 2 + 1,
}
CONV_UNITS = 15  # (A+B)(1 + 2 rho_x), A=2, B=1, rho_x=2


def time_plan(torch, plan, out, a, b, steps, warmup, stream):
    from fftwpp_amd import _lib
    L = _lib.lib()
    ms = (ctypes.c_float * 32)()
    n = ctypes.c_int()
    for _ in range(warmup):
        _lib.check(L.fftwpp_amd_conv_execute_timed(plan._h, out.data_ptr(), a.data_ptr(),
                                                   b.data_ptr(), stream, ms, 32, ctypes.byref(n)))
    torch.cuda.synchronize()
    sums = None
    names = None
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        _lib.check(L.fftwpp_amd_conv_execute_timed(plan._h, out.data_ptr(), a.data_ptr(),
                                                   b.data_ptr(), stream, ms, 32, ctypes.byref(n)))
        k = n.value
        if sums is None:
            sums = [0.0] * k
            names = [L.fftwpp_amd_conv_stage_name(plan._h, i).decode() for i in range(k)]
        for i in range(k):
            sums[i] += ms[i]
    stop.record()
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    stages = {names[i]: sums[i] / steps for i in range(len(names))}
    return total_ms / steps, stages


def roofline_of(stages, E, L3):
    """Dominant kernel (longest average duration) and its achieved bandwidth
    against its algorithmic bytes per launch."""
    name = max(stages, key=stages.get)
    bytes_alg = STAGE_UNITS.get(name, 0) * E * L3
    achieved = bytes_alg / (stages[name] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_GBS, 4), "traffic": None, "kernel": name,
            "bytes_alg_per_launch": bytes_alg, "avg_launch_ms": round(stages[name], 4)}


def sampled_direct_error(f, g, H, npts=6, seed=7):
    """Relative L2 error of H against direct O(N) sums at sampled outputs
    h[i,j,k] = sum_{p<=i,q<=j,r<=k} f[p,q,r] g[i-p,j-q,k-r]."""
    import numpy as np
    L = f.shape[0]
    rng = np.random.default_rng(seed)
    pts = [(L - 1, L - 1, L - 1)] + [tuple(int(v) for v in rng.integers(0, L, 3)) for _ in range(npts)]
    num = den = 0.0
    for i, j, k in pts:
        d = np.sum(f[:i + 1, :j + 1, :k + 1] * g[i::-1, j::-1, k::-1])
        num += abs(H[i, j, k] - d) ** 2
        den += abs(d) ** 2
    return math.sqrt(num / den)


def cpu_baseline(f, g, L):
    """The oracle's C port of the hybrid algorithm (OpenMP) on the host cores,
    one full convolution of the same workload."""
    from oracle import oracle as O
    O.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 64))
    O.lib().oracle_set_threads(cores)
    t0 = time.perf_counter()
    O.hybrid_conv3(f, g, L, L, 2, L, 2, L, 2)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 4), "unit": "conv/s", "cores": cores, "kind": "port",
            "sample": f"one full {L}^3 complex128 convolution (m={L}, q=2 per dimension) "
                      f"with the oracle's C port of the residue algorithm, OpenMP, "
                      f"{dt:.2f} s wall"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(see DESIGN.md for how the path would shard)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution

    L = args.size
    L3 = L ** 3
    dev = torch.device("cuda:0")
    if args.calibrate:
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    f, g = inputs(L)

    # FP64 drop-in path
    a = torch.tensor(f, device=dev)
    b = torch.tensor(g, device=dev)
    out = torch.empty_like(a)
    plan = DeviceConvolution((L, L, L), M=[2 * L] * 3)
    N, mm, qq = plan.geometry()
    ms64, st64 = time_plan(torch, plan, out, a, b, args.steps, args.warmup, stream)
    H64 = out.cpu().numpy()
    err64 = sampled_direct_error(f, g, H64)

    result = {
        "metric": METRIC,
        "value": round(1000.0 / ms64, 3),
        "unit": "conv/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"conv3d complex {L}^3 -> {2 * L}^3 (Convolution3, multBinary, "
                               f"A=2, B=1), inputs uniform[-1,1] seed {SEED}",
                   "L": [L] * 3, "M": [2 * L] * 3, "N": list(N), "m": list(mm), "q": list(qq)},
        "rel_l2": {"fp64_vs_direct_sampled": err64},
        "roofline": roofline_of(st64, 16, L3),
        "conv_roofline": {"bytes_alg": CONV_UNITS * 16 * L3,
                          "achieved": round(CONV_UNITS * 16 * L3 / (ms64 * 1e-3) / 1e9, 1),
                          "peak": PEAK_GBS, "unit": "GB/s",
                          "frac": round(CONV_UNITS * 16 * L3 / (ms64 * 1e-3) / 1e9 / PEAK_GBS, 4)},
        "stages_ms": {k: round(v, 4) for k, v in st64.items()},
    }
    del a, b, out, plan
    torch.cuda.empty_cache()

    if not args.no_fp32:
        a = torch.tensor(f, device=dev, dtype=torch.complex64)
        b = torch.tensor(g, device=dev, dtype=torch.complex64)
        out = torch.empty_like(a)
        plan = DeviceConvolution((L, L, L), dtype=torch.complex64, M=[2 * L] * 3)
        ms32, st32 = time_plan(torch, plan, out, a, b, args.steps, args.warmup, stream)
        H32 = out.cpu().numpy().astype(np.complex128)
        err32 = float(np.linalg.norm((H32 - H64).ravel()) / np.linalg.norm(H64.ravel()))
        result["fp32"] = {
            "value": round(1000.0 / ms32, 3), "unit": "conv/s", "ms_per_step": round(ms32, 4),
            "dtype": "f32", "rel_l2_vs_fp64": err32,
            "roofline": roofline_of(st32, 8, L3),
            "conv_roofline": {"bytes_alg": CONV_UNITS * 8 * L3,
                              "achieved": round(CONV_UNITS * 8 * L3 / (ms32 * 1e-3) / 1e9, 1),
                              "frac": round(CONV_UNITS * 8 * L3 / (ms32 * 1e-3) / 1e9 / PEAK_GBS, 4)},
            "stages_ms": {k: round(v, 4) for k, v in st32.items()},
        }
        result["rel_l2"]["fp32_vs_fp64"] = err32
        del a, b, out, plan
        torch.cuda.empty_cache()

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(f, g, L)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
