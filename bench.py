#!/usr/bin/env python3
"""bench.py -- 3D complex dealiased convolution 256^3 -> 512^3 on one MI355X.

One step = one implicitly dealiased convolution h = f (*) g (Convolution3 of
two 256^3 complex arrays padded to 512^3, multBinary), with f and g resident
in HBM and the result written out of place, FP64 (the drop-in precision).
The same workload is then timed in FP32 (FP64-generated twiddles) and its
relative L2 error against the FP64 result is reported.

Prints ONE JSON line.  Per-kernel durations come from HIP events recorded by
the library after each of its launches on the benchmark stream.  The batched
workload runs in a child process of its own before this process touches the
GPU: after the 3D legs' multi-GB buffers have been allocated and freed, the
batch measured in the same process runs 10-25% slower than in a fresh one
(not a thermal effect: a 20 s pause does not remove it), and a user of the
batch API does not carry that history.
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


def progress(msg):
    """One line per finished leg on stderr (the JSON result stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched, 1D and 2D Hermitian workloads")
    ap.add_argument("--traffic", action="store_true", default=True,
                    help="measure the dominant kernel's HBM bytes per launch with two "
                         "rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over a short "
                         "child run, scaled by a 512 MiB copy (the default; roofline.traffic "
                         "stays null when rocprofv3 is missing or a pass fails)")
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the counter passes (roofline.traffic null)")
    ap.add_argument("--batched-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--workload", default=None,
                    choices=["hconv2d", "real3d", "conv1d", "user_multiplier", "batched"],
                    help="run one secondary workload alone and print its JSON (for "
                         "profiling that workload's kernels)")
    ap.add_argument("--traffic-child", default=None,
                    choices=["f64", "f32", "f64-plane", "f32-plane"],
                    help=argparse.SUPPRESS)
    ap.add_argument("--calibrate", action="store_true",
                    help="run a 512 MiB device-to-device copy first (the known-size "
                         "reference for HBM traffic counters)")
    return ap.parse_args()


def calibration_copy(torch, dev):
    """Two 512 MiB device-to-device copies: the known-size reference that scales
    the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


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
    # one launch for y forward, z rows and y backward of every x-plane: reads
    # both x-transformed arrays, writes the plane results
    "y-z planes": 4 + 2,
}
CONV_UNITS = 15  # (A+B)(1 + 2 rho_x), A=2, B=1, rho_x=2


def time_plan(torch, plan, out, a, b, steps, warmup, stream):
    """Milliseconds per step of `steps` back-to-back executions (after
    `warmup` untimed ones), bracketed by synchronisations; then, untimed,
    the same number of executions with a HIP event after every kernel (on
    the plan's stream) for the per-stage durations.  The event pass waits
    for each execution, so it is kept out of the timed steps."""
    from fftwpp_amd import _lib
    L = _lib.lib()
    for _ in range(warmup):
        _lib.check(L.fftwpp_amd_conv_execute_out(plan._h, out.data_ptr(), a.data_ptr(),
                                                 b.data_ptr(), stream))
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        _lib.check(L.fftwpp_amd_conv_execute_out(plan._h, out.data_ptr(), a.data_ptr(),
                                                 b.data_ptr(), stream))
    stop.record()
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms = (ctypes.c_float * 32)()
    n = ctypes.c_int()
    sums = None
    names = None
    for _ in range(steps):
        _lib.check(L.fftwpp_amd_conv_execute_timed(plan._h, out.data_ptr(), a.data_ptr(),
                                                   b.data_ptr(), stream, ms, 32, ctypes.byref(n)))
        k = n.value
        if sums is None:
            sums = [0.0] * k
            names = [L.fftwpp_amd_conv_stage_name(plan._h, i).decode() for i in range(k)]
        for i in range(k):
            sums[i] += ms[i]
    torch.cuda.synchronize()
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


def bench_plane_engine(torch, f, g, L, H64, steps, warmup, stream):
    """The 3D workload under a memory limit that selects the plane kernel for
    the y-z stage (2 x-transformed arrays plus its L2 exchange areas instead
    of 2 + 2 q_y arrays): conv/s, work bytes of both engines, and the
    difference from the default engine's FP64 result."""
    import numpy as np
    from fftwpp_amd.torch_conv import DeviceConvolution
    dev = torch.device("cuda:0")
    out = {}
    for key, dt, E in (("f64", torch.complex128, 16), ("f32", torch.complex64, 8)):
        a = torch.tensor(f, device=dev, dtype=dt)
        b = torch.tensor(g, device=dev, dtype=dt)
        o = torch.empty_like(a)
        plan = DeviceConvolution((L, L, L), dtype=dt, M=[2 * L] * 3)
        full, _ = plan.work_bytes()
        plan.set_memory_limit(3 * 2 * L * L * L * E)
        need, plane = plan.work_bytes()
        if not plane:
            return {"note": "plane kernel not available on this device"}
        ms, st = time_plan(torch, plan, o, a, b, steps, warmup, stream)
        H = o.cpu().numpy().astype(np.complex128)
        err = float(np.linalg.norm((H - H64).ravel()) / np.linalg.norm(H64.ravel()))
        pk = st.get("y-z planes", 0.0)
        out[key] = {"conv_per_s": round(1000.0 / ms, 3), "ms_per_conv": round(ms, 4),
                    "work_bytes": need, "work_bytes_default_engine": full,
                    "plane_kernel_ms": round(pk, 4),
                    "plane_kernel_alg_GBs": round(STAGE_UNITS["y-z planes"] * E * L ** 3 / (pk * 1e-3) / 1e9, 1)
                    if pk else None,
                    "rel_l2_vs_default_fp64": err,
                    "stages_ms": {k: round(v, 4) for k, v in st.items()}}
        del a, b, o, plan
        torch.cuda.empty_cache()
    out["note"] = ("memory-capped engine: HBM traffic of the 15-unit model, about a third of "
                   "the work memory; slower than the default engine on one MI355X")
    return out


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


def cpu_library_baseline(f, g, L):
    """The same convolution by explicit zero padding to (2L)^3 with a library
    FFT on the host cores (scipy.fft's pocketfft, multithreaded): three
    transforms of the padded arrays and a product.  A library-backed CPU
    yardstick beside the port (FFTW3, which the reference uses, is not
    installed)."""
    import numpy as np
    import scipy.fft as sf
    cores = _workers()
    N = 2 * L
    t0 = time.perf_counter()
    F = np.zeros((N, N, N), dtype=np.complex128)
    F[:L, :L, :L] = f
    F = sf.fftn(F, overwrite_x=True, workers=cores)
    G = np.zeros((N, N, N), dtype=np.complex128)
    G[:L, :L, :L] = g
    G = sf.fftn(G, overwrite_x=True, workers=cores)
    F *= G
    del G
    F = sf.ifftn(F, overwrite_x=True, workers=cores)
    h = F[:L, :L, :L].copy()
    dt = time.perf_counter() - t0
    del F
    return {"value": round(1.0 / dt, 4), "unit": "conv/s", "cores": cores,
            "kind": "library (scipy.fft explicit padding)",
            "sample": f"one full {L}^3 complex128 convolution by explicit zero padding to "
                      f"{N}^3, scipy.fft with workers={cores}, {dt:.2f} s wall"}, h


# Kernel of each stage of the 3D plan, as named in rocprofv3 records.
STAGE_KERNEL = {"z rows": "rowconv256p_kernel", "x forward": "colfwd256_kernel",
                "y forward": "colfwd256_kernel", "y backward": "colbwd256_kernel",
                "x backward": "colbwd256_kernel", "y-z planes": "plane256_kernel"}
COPY_BYTES = 512 << 20


def traffic_child(key, L):
    """Body profiled by --traffic: a 512 MiB device copy (the counters'
    calibration), then two steps of the 3D plan."""
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution
    dev = torch.device("cuda:0")
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    dt = torch.complex128 if key.startswith("f64") else torch.complex64
    f, g = inputs(L)
    a = torch.tensor(f, device=dev, dtype=dt)
    b = torch.tensor(g, device=dev, dtype=dt)
    out = torch.empty_like(a)
    plan = DeviceConvolution((L, L, L), dtype=dt, M=[2 * L] * 3)
    if key.endswith("-plane"):
        # the memory-capped engine: the y-z stage as the plane kernel
        E = 16 if key.startswith("f64") else 8
        plan.set_memory_limit(3 * 2 * L ** 3 * E)
    for _ in range(2):
        plan.convolve_out(out, a, b)
    torch.cuda.synchronize()


def measure_traffic(L, key="f64"):
    """Run two rocprofv3 counter passes over traffic_child in child processes
    (before this process touches the GPU) and return {kernel: bytes per
    launch} scaled by the copy kernel's reading, plus the scale factors."""
    import csv
    import glob
    import subprocess
    import tempfile
    out = {}
    scale = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = ["timeout", "-s", "KILL", "180", "rocprofv3", "--pmc", counter,
                   "--output-format", "csv", "-d", d, "-o", "pmc", "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-child", key,
                   "--size", str(L)]
            # the exit status is not trusted: rocprofv3 may end abnormally after
            # the profiled program finished and wrote complete records
            subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tmp)
            rows = []
            for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
                with open(path) as fh:
                    rows += list(csv.DictReader(fh))
            per = {}
            for r in rows:
                if r["Counter_Name"] != counter:
                    continue
                per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024)
            copy = [v for k, vs in per.items() if "copyBuffer" in k for v in vs]
            if not copy:
                raise RuntimeError("no copy kernel in the counter records")
            scale[counter] = COPY_BYTES / max(copy)
            for k, vs in per.items():
                name = next((n for n in set(STAGE_KERNEL.values()) if n in k), None)
                if name:
                    out.setdefault(name, {})[counter] = sum(vs) / len(vs) * scale[counter]
    return out, scale


def batched_lengths(count=4096, lo=64, hi=16384, seed=SEED):
    """L = floor(exp(U(ln lo, ln hi))), PCG64(seed)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(np.int64)


def _time_calls(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_batched(torch, dev, steps, warmup, per_call=True):
    """4096 independent 1D convolutions, L log-uniform in [64,16384], M=2L:
    one autotuned plan replayed as a hipGraph, the same plan without tuning
    run eagerly, and one DeviceConvolution call per item."""
    import numpy as np
    from fftwpp_amd.batch import BatchedConvolution
    from fftwpp_amd.torch_conv import DeviceConvolution
    Ls = batched_lengths()
    total = int(Ls.sum())
    rng = np.random.Generator(np.random.PCG64(SEED + 4))
    fh = rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)
    gh = rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)
    res = {"workload": "4096 1D complex convolutions, L log-uniform in [64,16384] "
                       f"(sum L = {total}), M = 2L, multBinary, uniform[-1,1] seed {SEED + 4}"}
    outs = {}
    import tempfile
    cache = tempfile.NamedTemporaryFile(prefix="fftwpp_amd_tune_", suffix=".json", delete=False)
    cache.close()
    os.unlink(cache.name)
    saved = os.environ.get("FFTWPP_AMD_TUNE_CACHE")
    os.environ["FFTWPP_AMD_TUNE_CACHE"] = cache.name   # a fresh cache: the first plan tunes
    for dt, key, E in [(torch.complex128, "f64", 16), (torch.complex64, "f32", 8)]:
        f = torch.tensor(fh, dtype=dt, device=dev)
        g = torch.tensor(gh, dtype=dt, device=dev)
        out = torch.empty_like(f)
        t0 = time.perf_counter()
        plan = BatchedConvolution(Ls, M=2 * Ls, dtype=dt, tune=True, graph=True)
        tune_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        again = BatchedConvolution(Ls, M=2 * Ls, dtype=dt, tune=True, graph=True)
        cached_s = time.perf_counter() - t0
        del again
        ms = _time_calls(torch, lambda: plan(f, g, out), steps, warmup)
        outs[key] = out.clone()
        eager = BatchedConvolution(Ls, M=2 * Ls, dtype=dt, tune=False, graph=False)
        ms_e = _time_calls(torch, lambda: eager(f, g, out), steps, warmup)
        bytes_alg = 3 * E * total
        r = {"batches_per_s": round(1000.0 / ms, 2), "ms_per_batch": round(ms, 4),
             "achieved_GBs": round(bytes_alg / (ms * 1e-3) / 1e9, 1),
             "frac_of_hbm_peak": round(bytes_alg / (ms * 1e-3) / 1e9 / PEAK_GBS, 4),
             "bytes_alg": bytes_alg, "tune_s": round(tune_s, 2),
             "create_from_tune_cache_s": round(cached_s, 4),
             "untuned_eager_ms_per_batch": round(ms_e, 4),
             "classes": plan.describe().strip().splitlines()}
        if per_call and key == "f64":
            plans = [DeviceConvolution((int(L),), dtype=dt, M=[2 * int(L)]) for L in Ls]
            off = plan.offsets
            fv = [f[off[i]:off[i + 1]] for i in range(len(Ls))]
            gv = [g[off[i]:off[i + 1]].clone() for i in range(len(Ls))]
            ov = [torch.empty_like(v) for v in fv]

            def calls():
                for i in range(len(Ls)):
                    plans[i].convolve_out(ov[i], fv[i], gv[i])
            ms_c = _time_calls(torch, calls, 2, 1)
            ref = torch.cat(ov)
            r["per_call_ms_per_batch"] = round(ms_c, 3)
            r["speedup_vs_per_call"] = round(ms_c / ms, 2)
            r["rel_l2_vs_per_call"] = float(torch.linalg.norm(outs[key] - ref) / torch.linalg.norm(ref))
            del plans
        res[key] = r
        progress(f"batched {key}: {ms:.4f} ms per batch")
        del f, g, out
        torch.cuda.empty_cache()
    if saved is None:
        del os.environ["FFTWPP_AMD_TUNE_CACHE"]
    else:
        os.environ["FFTWPP_AMD_TUNE_CACHE"] = saved
    if os.path.exists(cache.name):
        os.unlink(cache.name)
    a, b = outs["f64"], outs["f32"].to(torch.complex128)
    res["f32"]["rel_l2_vs_f64"] = float(torch.linalg.norm(a - b) / torch.linalg.norm(a))
    return res


def bench_batched_replay(torch, dev, steps, warmup):
    """The batched workload's tuned graph plans alone (FP64 and FP32), built
    through the tuning cache of the environment (FFTWPP_AMD_TUNE_CACHE or the
    default file): a second run finds its entries and times nothing at plan
    creation, so a kernel trace of it holds only the replayed batches."""
    import numpy as np
    from fftwpp_amd.batch import BatchedConvolution
    Ls = batched_lengths()
    total = int(Ls.sum())
    rng = np.random.Generator(np.random.PCG64(SEED + 4))
    fh = rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)
    gh = rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)
    res = {"workload": "4096 1D complex convolutions, L log-uniform in [64,16384], M = 2L "
                       "(tuned graph replay only)"}
    for dt, key in [(torch.complex128, "f64"), (torch.complex64, "f32")]:
        f = torch.tensor(fh, dtype=dt, device=dev)
        g = torch.tensor(gh, dtype=dt, device=dev)
        out = torch.empty_like(f)
        t0 = time.perf_counter()
        plan = BatchedConvolution(Ls, M=2 * Ls, dtype=dt, tune=True, graph=True)
        create_s = time.perf_counter() - t0
        ms = _time_calls(torch, lambda: plan(f, g, out), steps, warmup)
        res[key] = {"ms_per_batch": round(ms, 4), "create_s": round(create_s, 3),
                    "classes": plan.describe().strip().splitlines()}
        del plan, f, g, out
        torch.cuda.empty_cache()
    return res


def bench_conv1d(torch, dev, steps):
    """One 1D convolution L=1024, M=2048, FP64: latency per call."""
    import numpy as np
    from fftwpp_amd.torch_conv import DeviceConvolution
    rng = np.random.Generator(np.random.PCG64(SEED))
    f = torch.tensor(rng.uniform(-1, 1, 1024) + 1j * rng.uniform(-1, 1, 1024), device=dev)
    g = torch.tensor(rng.uniform(-1, 1, 1024) + 1j * rng.uniform(-1, 1, 1024), device=dev)
    out = torch.empty_like(f)
    plan = DeviceConvolution((1024,), M=[2048])
    ms = _time_calls(torch, lambda: plan.convolve_out(out, f, g), 50 * steps, 10)
    return {"workload": "1D complex L=1024, M=2048, FP64", "us_per_call": round(ms * 1e3, 2),
            "conv_per_s": round(1000.0 / ms, 1), "geometry": [list(v) for v in plan.geometry()]}


def _workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def hermitian_explicit(f, g, N):
    """Centred Hermitian 2D convolution by explicit zero padding to N x N
    (scipy.fft, multithreaded): f, g store kx = i - Lx//2 (odd Lx) and
    ky = j >= 0; the modes ky < 0 are conj f(-kx, -ky)."""
    import numpy as np
    import scipy.fft as sf
    Lx, Hy = f.shape
    rows = np.mod(np.arange(Lx) - Lx // 2, N)
    cols = np.mod(np.arange(2 * Hy - 1) - (Hy - 1), N)
    ix = np.ix_(rows, cols)
    spec = []
    for a in (f, g):
        full = np.empty((Lx, 2 * Hy - 1), dtype=np.complex128)
        full[:, Hy - 1:] = a
        full[:, :Hy - 1] = np.conj(a[::-1, 1:][:, ::-1])
        P = np.zeros((N, N), dtype=np.complex128)
        P[ix] = full
        spec.append(sf.fft2(P, overwrite_x=True, workers=_workers()))
    H = sf.ifft2(spec[0] * spec[1], overwrite_x=True, workers=_workers())
    return H[np.ix_(rows, np.arange(Hy))]


def bench_hconv2d(torch, dev, steps, warmup):
    """2D centred Hermitian convolution 2047 x 1024 stored, M = 3070, FP64;
    after the timed region, the result against explicit zero padding."""
    import numpy as np
    from fftwpp_amd import _lib
    from fftwpp_amd.torch_conv import DeviceConvolution
    Lx, Hy = 2047, 1024
    rng = np.random.Generator(np.random.PCG64(SEED))
    kx = (np.arange(Lx) - Lx // 2)[:, None]
    ky = np.arange(Hy)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.where(k > 0, 1.0 / np.where(k > 0, k, 1.0), 0.0)
    ts = []
    for _ in range(2):
        t = torch.tensor(amp * np.exp(1j * rng.uniform(0, 2 * math.pi, (Lx, Hy))), device=dev)
        _lib.lib().fftwpp_HermitianSymmetrizeX((Lx + 1) // 2, Hy, Lx // 2, t.data_ptr())
        ts.append(t)
    f, g = ts
    out = torch.empty_like(f)
    plan = DeviceConvolution((Lx, Hy), kind="hermitian", M=[3070, 3070])
    # 0.17 ms calls: at least 200 untimed ones, so that the first timed
    # plan does not pay for the clocks ramping up
    warm = max(warmup, 200)
    ms = _time_calls(torch, lambda: plan.convolve_out(out, f, g), max(steps, 100), warm)
    torch.cuda.synchronize()
    ref = hermitian_explicit(f.cpu().numpy(), g.cpu().numpy(), 3070)
    err = float(np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref))
    # the same plan with timed sub-FFT sizes (timed afresh, then cached)
    t0 = time.perf_counter()
    tuned = DeviceConvolution((Lx, Hy), kind="hermitian", M=[3070, 3070], tune=True,
                              retune=True)
    tune_s = time.perf_counter() - t0
    out_t = torch.empty_like(f)
    ms_t = _time_calls(torch, lambda: tuned.convolve_out(out_t, f, g), max(steps, 100), warm)
    torch.cuda.synchronize()
    err_t = float(np.linalg.norm(out_t.cpu().numpy() - ref) / np.linalg.norm(ref))
    bytes_alg = 12 * 16 * Lx * Hy
    return {"workload": "2D centred Hermitian 2047x1024 stored, M=3070, |k|^-1 random phase, FP64",
            "conv_per_s": round(1000.0 / ms, 1), "ms_per_conv": round(ms, 4),
            "achieved_GBs": round(bytes_alg / (ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(bytes_alg / (ms * 1e-3) / 1e9 / PEAK_GBS, 4),
            "rel_l2_vs_explicit_padding": err,
            "geometry": [list(v) for v in plan.geometry()],
            "tuned": {"conv_per_s": round(1000.0 / ms_t, 1), "ms_per_conv": round(ms_t, 4),
                      "geometry": [list(v) for v in tuned.geometry()],
                      "tune_s": round(tune_s, 2), "rel_l2_vs_explicit_padding": err_t}}


def bench_real3d(torch, dev, steps, warmup):
    """3D real convolution 256^3 -> 512^3 (fftPadReal outer dimension), FP64
    and FP32, checked against the complex path on the same real data.  The x
    dimension runs on the real residue transforms: 257 kept x planes (N/2 + 1
    of 512) go through the complex y and z stages."""
    import numpy as np
    from fftwpp_amd.torch_conv import DeviceConvolution
    L = 256
    rng = np.random.Generator(np.random.PCG64(SEED + 5))
    f = torch.tensor(rng.uniform(-1, 1, (L, L, L)), device=dev)
    g = torch.tensor(rng.uniform(-1, 1, (L, L, L)), device=dev)
    out = torch.empty_like(f)
    plan = DeviceConvolution((L, L, L), kind="real", dtype=torch.float64, M=[2 * L] * 3)
    stream = torch.cuda.current_stream().cuda_stream
    ms, stages = time_plan(torch, plan, out, f, g, steps, warmup, stream)
    ref = DeviceConvolution((L, L, L), M=[2 * L] * 3)
    fc, gc = f.to(torch.complex128), g.to(torch.complex128)
    hc = torch.empty_like(fc)
    ref.convolve_out(hc, fc, gc)
    err = float(torch.linalg.norm(out - hc.real) / torch.linalg.norm(hc.real))
    del ref, fc, gc
    f32, g32 = f.float(), g.float()
    out32 = torch.empty_like(f32)
    plan32 = DeviceConvolution((L, L, L), kind="real", dtype=torch.float32, M=[2 * L] * 3)
    ms32, stages32 = time_plan(torch, plan32, out32, f32, g32, steps, warmup, stream)
    err32 = float(torch.linalg.norm(out32.double() - hc.real) / torch.linalg.norm(hc.real))
    bytes_alg = CONV_UNITS * 8 * L ** 3
    return {"workload": "3D real 256^3 -> 512^3 (M = 2L), FP64, uniform[-1,1] seed %d" % (SEED + 5),
            "conv_per_s": round(1000.0 / ms, 2), "ms_per_conv": round(ms, 4),
            "achieved_GBs": round(bytes_alg / (ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(bytes_alg / (ms * 1e-3) / 1e9 / PEAK_GBS, 4),
            "bytes_alg": bytes_alg, "rel_l2_vs_complex_path": err,
            "stages_ms": {k: round(v, 4) for k, v in stages.items()},
            "fp32": {"conv_per_s": round(1000.0 / ms32, 2), "ms_per_conv": round(ms32, 4),
                     "rel_l2_vs_complex_fp64": err32,
                     "stages_ms": {k: round(v, 4) for k, v in stages32.items()}},
            "note": "x on the real residue transforms (257 of 512 x planes kept)"}


def bench_user_multiplier(torch, dev, steps, warmup):
    """The 3D 256^3 -> 512^3 workload through the user-multiplier path (a
    device binary product of tests/support/multipliers.hip: every dimension
    expanded, the multiplier called once), against the fused path."""
    import ctypes
    import numpy as np
    from fftwpp_amd.torch_conv import DeviceConvolution, UserConvolution
    lib = os.path.join(ROOT, "tests", "support", "libtestmult.so")
    if not os.path.exists(lib):
        return {"skipped": "tests/support/libtestmult.so not built"}
    mult = ctypes.cast(ctypes.CDLL(lib).testmult_binary, ctypes.c_void_p)
    L = 256
    rng = np.random.Generator(np.random.PCG64(SEED + 6))
    f = torch.tensor(rng.uniform(-1, 1, (L, L, L)) + 1j * rng.uniform(-1, 1, (L, L, L)), device=dev)
    g = torch.tensor(rng.uniform(-1, 1, (L, L, L)) + 1j * rng.uniform(-1, 1, (L, L, L)), device=dev)
    plan = UserConvolution((L, L, L), 2, 1, mult, M=[2 * L] * 3, m=[256] * 3)
    a = f.clone()
    ms = _time_calls(torch, lambda: plan(a, g), steps, warmup)
    a.copy_(f)
    plan(a, g)
    fused = DeviceConvolution((L, L, L), M=[2 * L] * 3)
    h = torch.empty_like(f)
    fused.convolve_out(h, f, g)
    err = float(torch.linalg.norm(a - h) / torch.linalg.norm(h))
    del plan, a
    # the fused user multiplier (include/fftwpp_amd_fused.h): a device functor
    # F G / (1 + c |k|^2) compiled into the z-row kernel (fusedmult.hip)
    fn = ctypes.CDLL(lib).fusedmult_weighted
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p] * 4 + [ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                          ctypes.c_double, ctypes.c_void_p]
    Ns = (ctypes.c_int64 * 3)(2 * L, 2 * L, 2 * L)
    out = torch.empty_like(f)
    stream = torch.cuda.current_stream().cuda_stream

    fb = ctypes.CDLL(lib).fusedmult_binary
    fb.restype = ctypes.c_int
    fb.argtypes = [ctypes.c_void_p] * 5
    from fftwpp_amd import _lib

    def fused_call(c):
        _lib.check(fn(fused._h, out.data_ptr(), f.data_ptr(), g.data_ptr(), Ns, 3, c, stream))

    def binary_call():
        _lib.check(fb(fused._h, out.data_ptr(), f.data_ptr(), g.data_ptr(), stream))
    # the fused calls take 2.3 ms: more steps than the expanded path's
    fsteps, fwarm = max(steps, 20), max(warmup, 5)
    ms_b = _time_calls(torch, binary_call, fsteps, fwarm)
    torch.cuda.synchronize()
    same_b = bool(torch.equal(out, h))
    ms_f = _time_calls(torch, lambda: fused_call(1e-6), fsteps, fwarm)
    fused_call(0.0)
    torch.cuda.synchronize()
    same = bool(torch.equal(out, h))
    del f, g, h, out
    torch.cuda.empty_cache()
    bytes_alg = CONV_UNITS * 16 * L ** 3
    return {"workload": "3D complex 256^3 -> 512^3, FP64, user device multiplier F0 *= F1",
            "conv_per_s": round(1000.0 / ms, 2), "ms_per_conv": round(ms, 4),
            "achieved_GBs": round(bytes_alg / (ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(bytes_alg / (ms * 1e-3) / 1e9 / PEAK_GBS, 4),
            "bytes_alg": bytes_alg, "rel_l2_vs_fused": err,
            "fused_functor": {
                "workload": "the same multiplier (F G) as a caller's device functor compiled "
                            "into the z-row kernel (include/fftwpp_amd_fused.h)",
                "conv_per_s": round(1000.0 / ms_b, 2), "ms_per_conv": round(ms_b, 4),
                "equals_library_multBinary": same_b},
            "fused_functor_weighted": {
                "workload": "a device functor F G / (1 + 1e-6 |k|^2), k the padded grid "
                            "frequency in x, y, z",
                "conv_per_s": round(1000.0 / ms_f, 2), "ms_per_conv": round(ms_f, 4),
                "unit_weight_equals_library_multBinary": same}}


def bench_fft(torch, dev, steps, warmup):
    """Standalone FFTs: 1D complex n = 2^20 and 3D complex 256^3, FP64, sign
    -1; algorithmic bytes one read and one write of the array."""
    import numpy as np
    from fftwpp_amd.torch_fft import DeviceFFT
    res = {}
    rng = np.random.Generator(np.random.PCG64(SEED + 7))
    for name, shape in [("fft1d_2^20", (1 << 20,)), ("fft3d_256", (256, 256, 256))]:
        x = torch.tensor(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape), device=dev)
        y = torch.empty_like(x)
        plan = DeviceFFT(shape if len(shape) > 1 else shape[0], -1, dtype=torch.float64)
        ms = _time_calls(torch, lambda: plan(x, y), 5 * steps, warmup)
        err = float(torch.linalg.norm(y - torch.fft.fftn(x)) / torch.linalg.norm(y))
        # torch.fft (the vendor FFT library) on the same array, for comparison
        ms_t = _time_calls(torch, lambda: torch.fft.fftn(x), 5 * steps, warmup)
        nbytes = 2 * 16 * x.numel()
        res[name] = {"ms": round(ms, 4), "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_of_hbm_peak": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_GBS, 4),
                     "rel_l2_vs_torch_fft": err, "torch_fft_ms": round(ms_t, 4)}
        del x, y, plan
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.traffic_child:
        traffic_child(args.traffic_child, args.size)
        return
    if args.batched_child:
        import torch
        print(json.dumps(bench_batched(torch, torch.device("cuda:0"), args.steps, args.warmup)),
              flush=True)
        return
    if args.workload:
        import torch
        dev = torch.device("cuda:0")
        if args.calibrate:
            calibration_copy(torch, dev)
        fn = {"hconv2d": lambda: bench_hconv2d(torch, dev, args.steps, args.warmup),
              "real3d": lambda: bench_real3d(torch, dev, args.steps, args.warmup),
              "conv1d": lambda: bench_conv1d(torch, dev, args.steps),
              "batched": lambda: bench_batched_replay(torch, dev, args.steps, args.warmup),
              "user_multiplier": lambda: bench_user_multiplier(torch, dev, min(args.steps, 5),
                                                               1)}[args.workload]
        print(json.dumps(fn()), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus N>1 is not supported (the batched "
                                   "items would shard across GPUs with no exchange, the 3D and "
                                   "2D Hermitian outer residues by an all-to-all per residue "
                                   "pair)"}))
        sys.exit(2)
    traffic = None
    # never nest counter passes inside a profiled process
    profiled = any(k.startswith("ROCPROF") for k in os.environ) or \
        "rocprof" in os.environ.get("LD_PRELOAD", "")
    if args.traffic and profiled:
        print("traffic measurement skipped: running under a profiler", file=sys.stderr, flush=True)
    elif args.traffic:
        # child processes first: this process has not touched the GPU yet
        try:
            traffic = measure_traffic(args.size)
            progress("traffic counters done")
        except Exception as e:  # no rocprofv3, no records: the field stays null
            print(f"traffic measurement skipped: {e}", file=sys.stderr, flush=True)
    batched = None
    if not args.no_extra:
        # the batched workload in a fresh process (see the module docstring)
        import subprocess
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--batched-child",
                            "--steps", str(args.steps), "--warmup", str(args.warmup)],
                           stdout=subprocess.PIPE, text=True)
        lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode != 0 or not lines:
            print(f"batched child failed with status {r.returncode}", file=sys.stderr, flush=True)
            sys.exit(1)
        batched = json.loads(lines[-1])
        progress("batched workload done (child process)")
    import numpy as np
    import torch
    from fftwpp_amd.torch_conv import DeviceConvolution

    L = args.size
    L3 = L ** 3
    dev = torch.device("cuda:0")
    if args.calibrate:
        calibration_copy(torch, dev)
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
    progress(f"3D FP64 done: {ms64:.4f} ms")

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
    if traffic is not None:
        per, scale = traffic
        kname = STAGE_KERNEL.get(result["roofline"]["kernel"])
        if kname in per and len(per[kname]) == 2:
            t = per[kname]["FETCH_SIZE"] + per[kname]["WRITE_SIZE"]
            result["roofline"]["traffic"] = round(t)
            result["roofline"]["traffic_kernel"] = kname
        result["traffic_scale"] = {k: round(v, 4) for k, v in scale.items()}
        result["traffic_per_kernel"] = {k: {c: round(v) for c, v in d.items()}
                                        for k, d in per.items()}
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
        progress(f"3D FP32 done: {ms32:.4f} ms")
        del a, b, out, plan
        torch.cuda.empty_cache()

    if not args.no_extra:
        result["plane_engine"] = bench_plane_engine(torch, f, g, L, H64, min(args.steps, 10),
                                                    2, stream)
        progress("plane engine done")
        result["batched"] = batched
        result["conv1d"] = bench_conv1d(torch, dev, args.steps)
        progress("1D done")
        result["hconv2d"] = bench_hconv2d(torch, dev, args.steps, args.warmup)
        progress("2D Hermitian done")
        result["real3d"] = bench_real3d(torch, dev, args.steps, args.warmup)
        progress("3D real done")
        result["user_multiplier"] = bench_user_multiplier(torch, dev, min(args.steps, 5), 1)
        progress("user multiplier done")
        result["fft"] = bench_fft(torch, dev, args.steps, args.warmup)

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(f, g, L)
        progress("CPU baseline (port) done")
        lib_base, href = cpu_library_baseline(f, g, L)
        result["cpu_baseline_library"] = lib_base
        result["rel_l2"]["fp64_vs_explicit_padding"] = float(
            np.linalg.norm((H64 - href).ravel()) / np.linalg.norm(href.ravel()))
        del href
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
