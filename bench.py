"""Benchmark: TV-ADMM solves/s at 256x256x160 on one MI355X.

Workload (default): 256x256x160 at 0.9x0.9x1.0 mm, B0 = z, a seeded
40-ellipsoid phantom (local field = Re IFFT(D FFT(chi)) + N(0, 0.005 ppm),
zero outside the mask), lambda = 1e-3, rho = 0.1, tol = 0 so exactly 200
iterations run.  One "step" = one full solve of every volume in the batch:
spectral setup (FFT of the field, dipole kernel and inverse operator in
FP64), 200 ADMM iterations, masking and the f64 output, with inputs already
resident in HBM.  FFTs and ADMM state are FP32; operator construction and
convergence norms are FP64.

The JSON line also carries
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its mean launch time (HIP events on the solver's stream),
                against 8 TB/s;
  cpu_baseline  the C++ f64 oracle (restatement of the reference's `parallel`
                CPU path: OpenMP over FFT axis loops only) timed on this host
                for one full solve of the first volume, on every CPU the job
                may use (host CPU model and counts recorded);
  parity        the GPU result of the timed path for that volume against the
                oracle's solve, at the same iteration count (NRMSE, max-abs).
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

Other workloads (--workload), each printing one JSON line of its own:
  uneven32  32 volumes, each axis from {144,...,256}, voxel sizes U(0.6,1.2) mm,
            B0 tilted U(0,20) deg, 12-ellipsoid phantoms, lambda 1e-3,
            rho 0.1, tol 0, 200 iterations, one batch
  chain8    8 volumes 384x384x256 at 0.6 mm: total field (phantom + 3
            external dipoles + noise) -> batched V-SHARP (radii 12..1 mm)
            -> batched TV-ADMM (lambda 1e-3, rho 0.1, tol 0, 200 iterations)
            on the V-SHARP outputs, device-resident end to end
  vjp       192x192x128 at 1 mm: K = 20 unrolled TV-ADMM iterations and the
            vector-Jacobian product of ||chi_K - chi_true||^2 w.r.t. the
            local field and lambda (one step = forward + backward)
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "qsm.rs_amd")
METRIC = "TV-ADMM solves/s at 256²×160: FP32+batched vs plain, NRMSE≤1e-3 vs f64"
HBM_PEAK_GBS = 8000.0

KERNELS = ["r2c_x+div (previous relchange, z/u update, bdiv, R2C x)", "c2c_y fwd",
           "c2c_z + x-update + inverse z", "c2c_y inv", "c2r_x + relchange sums"]


def _load(name, path, search=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=search)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def alg_bytes(dims):
    """Algorithmic HBM bytes per launch of each iteration kernel (FP32):
    every input read once, every output written once.  A half spectrum holds
    (nx + 1) // 2 complex values per row (even nx: DC and Nyquist share one),
    i.e. Nh = ceil(nx / 2) ny nz elements of 8 bytes."""
    nx, ny, nz = dims
    N = nx * ny * nz
    Nh = ((nx + 1) // 2) * ny * nz
    return [
        28 * N + 8 * Nh,   # read x (4N) + u (12N); write u (12N) + half spectrum (8Nh)
        16 * Nh,           # read + write half spectrum
        24 * Nh,           # spectrum r/w (16Nh) + f^ (8Nh); rho*inv_a/N is formed in-kernel
        16 * Nh,
        8 * Nh + 8 * N,    # read spectrum + x_prev, write x
    ]


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_tv_256x256x160.txt")


def traffic_for(kernel_index, nvol, dims):
    """Measured HBM bytes per launch of one iteration kernel, from the
    committed rocprofv3 PMC summary (256x256x160, one volume; scaled by the
    batch)."""
    names = ["k_r2c_x", "k_y_pass", "k_z_tv", "k_y_pass", "k_c2r_tv"]
    if tuple(dims) != (256, 256, 160) or kernel_index >= len(names) or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if f and f[0] == names[kernel_index]:
            return int(float(f[4]) * 1e6 * nvol)   # columns: name, dispatches, fetch, write, HBM MB
    return None


def vsharp_kernel_bytes(dims):
    """Algorithmic HBM bytes per volume of each V-SHARP launch as built
    (every input read once, every output written once; Nh = ceil(nx/2) ny nz
    complex values of 8 bytes).  Keys are the launch kinds; "radius" kinds
    run once per radius."""
    nx, ny, nz = dims
    N = nx * ny * nz
    Nh = ((nx + 1) // 2) * ny * nz
    return {
        # setup: F^ = FFT(field) and M^ = FFT(mask), each R2C + y + z
        "r2c_f64": 8 * N + 8 * Nh, "r2c_u8": N + 8 * Nh,
        "y_fwd_x2": 2 * 16 * Nh, "z_fwd_x2": 2 * 16 * Nh,   # one y and one z launch per spectrum
        "vs_init": 5 * N,                      # clear the selected field (f32) and mask (u8)
        # per radius: sphere generated in the R2C pass (nothing read)
        "radius_r2c_sphere": 8 * Nh,
        "radius_y_fwd": 16 * Nh,
        # z: S^ in, F^ and M^ read, P = M^ S^ and Q = F^ (1 - S^) written
        "radius_z": 40 * Nh,
        "radius_y_inv2": 32 * Nh,              # inverse y of P and Q in one launch
        "radius_c2r_select": 16 * Nh + N,      # P + iQ rows, the final-mask flags
        "select_writes": 5 * N,                # each voxel's value and flag written once over all radii
        "sfirst": 4 * Nh,                      # S^ of the largest radius kept (first radius)
        # deconvolution with the largest kernel
        "r2c_f32": 4 * N + 8 * Nh, "deconv_y_fwd": 16 * Nh, "z_deconv": 20 * Nh,
        "deconv_y_inv": 16 * Nh, "c2r_out": 8 * Nh + 10 * N,
    }


def vsharp_alg_bytes(dims, R):
    """Algorithmic bytes of one V-SHARP run with R radii, as built:
    FFT(field) and FFT(mask) once, per radius the sphere's R2C / y / z
    (P and Q) / inverse y of both / one C2R of P + iQ with the selection,
    then the deconvolution round trip and the output pass."""
    k = vsharp_kernel_bytes(dims)
    per_radius = sum(v for n, v in k.items() if n.startswith("radius_"))
    once = sum(v for n, v in k.items() if not n.startswith("radius_"))
    return once + R * per_radius


def host_cpu_info():
    """What this host offers the CPU baseline: logical CPUs (nproc), the ones
    this process may run on (affinity), the cgroup CPU quota if one is set,
    physical cores among the allowed CPUs, and the CPU model."""
    info = {"nproc": os.cpu_count() or 1}
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = list(range(info["nproc"]))
    info["affinity"] = len(allowed)
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    info["cgroup_quota_cpus"] = quota
    model, cores, cur = None, set(), {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur.get("processor") is not None and int(cur["processor"]) in allowed:
                    cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                cur = {}
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            cur[k] = v
            if k == "model name" and model is None:
                model = v
        if cur.get("processor") is not None and int(cur["processor"]) in allowed:
            cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
    except OSError:
        pass
    info["physical_cores_allowed"] = len(cores) or None
    info["model"] = model
    usable = info["affinity"]
    if quota is not None:
        usable = max(1, min(usable, int(quota)))
    info["usable_threads"] = usable
    return info


def cpu_baseline(ph, params, threads, x_gpu=None, cpu_info=None):
    """One full f64 oracle solve (setup + params.max_iter iterations, tol 0)
    of the workload's first volume, wall-clock timed.  With x_gpu (the GPU
    result of the timed path for the same volume) the two are compared:
    NRMSE of tests/common/mod.rs:34-51, raw and with the mean difference
    inside the mask removed, and max-abs error.  The oracle is the checker
    here; the GPU result is what is checked."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    O.set_num_threads(threads)
    print(f"# cpu baseline: oracle solve of {'x'.join(map(str, ph.dims))} x{params.max_iter} "
          f"on {threads} threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    x_ref, it_ref, _ = O.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir,
                                 params.lambda_, params.rho, 0.0, params.max_iter, return_info=True)
    t_solve = time.perf_counter() - t0
    parity = None
    if x_gpu is not None:
        m = np.asarray(ph.mask).ravel() > 0
        d = (np.asarray(x_gpu).ravel() - np.asarray(x_ref).ravel())[m]
        xr = np.asarray(x_ref).ravel()[m]
        rng = float(xr.max() - xr.min())
        parity = {
            "iterations": int(it_ref),
            "nrmse": float(O.nrmse(x_gpu, x_ref, ph.mask)),
            "nrmse_mean_removed": float(np.sqrt(np.mean((d - d.mean()) ** 2)) / rng),
            "max_abs": float(np.max(np.abs(d))),
            "max_abs_mean_removed": float(np.max(np.abs(d - d.mean()))),
            "oracle_range": rng,
            "bar": 1e-3,
            "pass": bool(float(O.nrmse(x_gpu, x_ref, ph.mask)) <= 1e-3),
            "note": ("GPU FP32 result of the timed batched path (volume 0) vs one full f64 oracle "
                     "solve of the same phantom, same iterations; the mean difference is the "
                     "reference's DC term for anisotropic voxels"),
        }
    ci = cpu_info or {}
    return {
        "parity": parity,
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": int(O.num_threads()),
        "kind": "port",
        "sample": (f"1 full solve: C++ f64 restatement of the reference's `parallel` path (own FFT, "
                   f"OpenMP over FFT axis loops only, other passes serial), setup + "
                   f"{params.max_iter} iterations on {'x'.join(map(str, ph.dims))}, "
                   f"{t_solve:.1f} s wall on {int(O.num_threads())} threads"),
        "host": {k: ci.get(k) for k in ("model", "nproc", "affinity", "cgroup_quota_cpus",
                                         "physical_cores_allowed", "usable_threads")},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4,
                    help="volumes per launch on the batched path (value); 1 = plain only")
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads for the CPU baseline (default: every CPU this job may use)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true",
                    help="time only the batched path (so a profiler sees batched launches only)")
    ap.add_argument("--dims", type=str, default="256,256,160")
    ap.add_argument("--workload", choices=["tv256", "uneven32", "chain8", "vjp", "hipfft"], default="tv256")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="before timing, copy BYTES device-to-device once (a plain copy "
                         "kernel of known size for scaling HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); volumes are "
              "independent, so a multi-GPU run would shard whole volumes across devices.",
              file=sys.stderr)
        sys.exit(2)

    qsm = _load("qsm_rs_amd", os.path.join(PKG_DIR, "__init__.py"), [PKG_DIR])
    phantom = _load("qsm_phantom", os.path.join(PKG_DIR, "phantom.py"))
    if a.workload == "uneven32":
        return bench_uneven(a, qsm, phantom)
    if a.workload == "chain8":
        return bench_chain(a, qsm, phantom)
    if a.workload == "vjp":
        return bench_vjp(a, qsm, phantom)
    if a.workload == "hipfft":
        return bench_vendor_fft(a, qsm, phantom)

    dims = tuple(int(v) for v in a.dims.split(","))
    vs = (0.9, 0.9, 1.0)
    params = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=a.iters)
    grid = qsm.Grid(dims, vs)
    phantoms = [phantom.make_phantom(dims, voxel_size=vs, bdir=(0.0, 0.0, 1.0), n_ellipsoids=40,
                                     seed=1001 + 1000 * v) for v in range(max(1, a.batch))]
    ph = phantoms[0]
    keep, vols = [], []
    for p_ in phantoms:
        tf = qsm.DeviceArray.from_numpy(p_.local_field)
        tm = qsm.DeviceArray.from_numpy(p_.mask)
        tc = qsm.DeviceArray(p_.local_field.size, np.float64)
        keep.append((tf, tm, tc))
        vols.append((grid, p_.bdir, tf.ptr, tm.ptr, tc.ptr))
    qsm.synchronize()
    if a.copy_calibration:
        n = a.copy_calibration // 4
        src = qsm.DeviceArray.from_numpy(np.ones(n, np.float32))
        dst = qsm.DeviceArray(n, np.float32)
        qsm.copy_device(dst, src)
        qsm.synchronize()
        src.free()
        dst.free()

    def timed(batch, nvol):
        for _ in range(a.warmup):
            batch.run()
        qsm.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            batch.run()
        qsm.synchronize()
        dt = time.perf_counter() - t0
        return a.steps * nvol / dt, dt / a.steps * 1e3

    nk = len(KERNELS)
    plain_value = plain_ms = plain_prof = x_plain = None
    if not (a.skip_plain and a.batch > 1):
        # plain: one volume per launch
        plain = qsm.TvBatch(vols[:1], params)
        plain_value, plain_ms = timed(plain, 1)
        iters_done = plain.status(0)[0]
        plain_prof = plain.profile(a.profile_iters)[:nk]
        plain.close()
        x_plain = keep[0][2].numpy()
    # batched: every volume in one launch per kernel
    if a.batch > 1:
        batch = qsm.TvBatch(vols, params)
        value, ms_per_step = timed(batch, a.batch)
        iters_done = batch.status(0)[0]
        ms = batch.profile(a.profile_iters)[:nk]
        batch.close()
    else:
        value, ms_per_step, ms = plain_value, plain_ms, plain_prof

    # per-kernel mean launch time (HIP events on the solver's stream)
    nbytes = [b * max(1, a.batch) for b in alg_bytes(dims)]
    dom = int(np.argmax(ms))
    achieved = nbytes[dom] / (ms[dom] * 1e-3) / 1e9
    iter_bytes = sum(nbytes)
    iter_ms = float(sum(ms))
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_for(dom, max(1, a.batch), dims),
        "kernel": KERNELS[dom],
        "kernel_ms": round(float(ms[dom]), 5),
        "alg_bytes_per_launch": int(nbytes[dom]),
        "iteration_GBs": round(iter_bytes / (iter_ms * 1e-3) / 1e9, 1),
        "per_kernel_ms": {k: round(float(v), 5) for k, v in zip(KERNELS, ms)},
    }
    # what a plain device copy reaches on this GPU, for reading the roofline
    copy_gbs = qsm.hbm_copy_probe(1 << 30, 5)
    roofline["copy_probe_GBs"] = round(copy_gbs, 1)
    roofline["frac_of_copy"] = round(achieved / copy_gbs, 4)
    if dom == 0:
        # the same access shape as the fused R2C pass: x and the three dual
        # components read, the three dual components and the spectrum written,
        # element by element over the batch
        nvox = int(np.prod(dims)) * max(1, a.batch)
        shape_gbs = qsm.hbm_stream_probe(nvox, 4, 4, 5)
        roofline["shape_probe_GBs"] = round(shape_gbs, 1)
        roofline["frac_of_shape"] = round(achieved / shape_gbs, 4)
    cpu = None
    if not a.no_cpu:
        ci = host_cpu_info()
        threads = a.cpu_threads if a.cpu_threads > 0 else ci["usable_threads"]
        # the timed path's own output for volume 0 (every run writes the same)
        x_gpu = keep[0][2].numpy()
        cpu = cpu_baseline(ph, params, threads, x_gpu, ci)
        cpu["ratio_gpu_over_cpu"] = round(value / cpu["value"], 1)
    parity = cpu.pop("parity") if cpu else None

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"TV-ADMM {dims[0]}x{dims[1]}x{dims[2]} @0.9x0.9x1.0 mm, 40-ellipsoid "
                        f"phantom, lambda 1e-3, rho 0.1, tol 0, {a.iters} iterations",
            "dims": list(dims),
            "batch": a.batch,
            "iterations": iters_done,
            "path": "batched" if a.batch > 1 else "plain",
        },
        "plain": None if plain_value is None else {
            "value": round(plain_value, 3), "ms_per_step": round(plain_ms, 3),
            "per_kernel_ms": {k: round(float(v), 5) for k, v in zip(KERNELS, plain_prof)}},
        "batched_over_plain": None if plain_value is None else round(value / plain_value, 3),
        # volume 0 solved alone and inside the batch: the same bits
        "plain_equals_batched": (None if x_plain is None or a.batch <= 1
                                 else bool(np.array_equal(x_plain, keep[0][2].numpy()))),
        "voxel_iterations_per_s": round(value * int(np.prod(dims)) * iters_done, 1),
        "fft_schedule": {f"y{dims[1]}": qsm.fft_schedule("y", dims[1]),
                         f"z{dims[2]}": qsm.fft_schedule("z", dims[2])},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out))


def _timed(a, run, nvol, sync):
    for _ in range(a.warmup):
        run()
    sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run()
    sync()
    dt = time.perf_counter() - t0
    return a.steps * nvol / dt, dt / a.steps * 1e3


def _roofline(nbytes, ms, names):
    dom = int(np.argmax(ms))
    achieved = nbytes[dom] / (ms[dom] * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": names[dom],
        "kernel_ms": round(float(ms[dom]), 5), "alg_bytes_per_launch": int(nbytes[dom]),
        "iteration_GBs": round(sum(nbytes) / (float(sum(ms)) * 1e-3) / 1e9, 1),
        "per_kernel_ms": {k: round(float(v), 5) for k, v in zip(names, ms)},
    }


def bench_uneven(a, qsm, phantom):
    """32 volumes of uneven dims / voxel sizes / B0 tilts in one TV batch."""
    count = 32
    shapes = phantom.uneven_batch_shapes(3002, count)
    params = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=a.iters)
    keep, vols, nbytes = [], [], np.zeros(len(KERNELS))
    t0 = time.perf_counter()
    for v, (dims, vs) in enumerate(shapes):
        b = phantom.tilted_bdir(3002 * 1000 + v)
        ph = phantom.make_phantom(dims, voxel_size=vs, bdir=b, n_ellipsoids=12, seed=3002 * 1000 + v)
        tf = qsm.DeviceArray.from_numpy(ph.local_field)
        tm = qsm.DeviceArray.from_numpy(ph.mask)
        tc = qsm.DeviceArray(ph.local_field.size, np.float64)
        keep.append((tf, tm, tc))
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
        nbytes += np.array(alg_bytes(dims), dtype=np.float64)
        print(f"# volume {v + 1}/{count} {dims} ready ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
    qsm.synchronize()
    batch = qsm.TvBatch(vols, params)
    value, ms_per_step = _timed(a, batch.run, count, qsm.synchronize)
    iters_done = batch.status(0)[0]
    ms = batch.profile(a.profile_iters)[: len(KERNELS)]
    batch.close()
    voxels = sum(int(np.prod(d)) for d, _ in shapes)
    print(json.dumps({
        "metric": "TV-ADMM solves/s, 32 uneven volumes in one batch",
        "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "32 volumes, axes from {144..256 step 16}, voxel U(0.6,1.2) mm, "
                               "B0 tilt U(0,20) deg, 12-ellipsoid phantoms, lambda 1e-3, rho 0.1, "
                               f"tol 0, {a.iters} iterations", "volumes": count,
                   "total_voxels": voxels, "iterations": iters_done,
                   "distinct_lengths": sorted({n for d, _ in shapes for n in d})},
        "voxel_iterations_per_s": round(voxels * iters_done / (ms_per_step * 1e-3), 1),
        "fft_schedule": {f"{ax}{n}": qsm.fft_schedule(ax, n)
                         for ax, k in (("y", 1), ("z", 2))
                         for n in sorted({d[k] for d, _ in shapes})},
        "roofline": _roofline(nbytes, ms, KERNELS),
    }))


def bench_chain(a, qsm, phantom):
    """V-SHARP then TV-ADMM on 8 volumes of 384x384x256, device-resident."""
    count = 8
    dims, vs = (384, 384, 256), (0.6, 0.6, 0.6)
    grid = qsm.Grid(dims, vs)
    vp = qsm.VsharpParams()          # 0.05, 12 mm, 1 mm: radii 12, 11, ..., 1
    radii = qsm.vsharp_radii(vp)
    params = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=a.iters)
    keep, vvols, tvols = [], [], []
    n = int(np.prod(dims))
    t0 = time.perf_counter()
    for v in range(count):
        seed = 4004 * 1000 + v
        ph = phantom.add_background(phantom.make_phantom(dims, voxel_size=vs, n_ellipsoids=12,
                                                         seed=seed), seed=seed)
        tf = qsm.DeviceArray.from_numpy(ph.total_field)
        tm = qsm.DeviceArray.from_numpy(ph.mask)
        tl = qsm.DeviceArray(n, np.float64)
        te = qsm.DeviceArray(n, np.uint8)
        tc = qsm.DeviceArray(n, np.float64)
        keep.append((tf, tm, tl, te, tc))
        vvols.append((grid, tf.ptr, tm.ptr, tl.ptr, te.ptr))
        tvols.append((grid, ph.bdir, tl.ptr, te.ptr, tc.ptr))
        print(f"# volume {v + 1}/{count} ready ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
        del ph
    qsm.synchronize()
    vsb = qsm.VsharpBatch(vvols, radii, vp.threshold)
    tvb = qsm.TvBatch(tvols, params)
    # stage times
    vs_vps, vs_ms = _timed(a, vsb.run, count, qsm.synchronize)
    tv_vps, tv_ms = _timed(a, tvb.run, count, qsm.synchronize)

    def chain():
        vsb.run()
        tvb.run()
    value, ms_per_step = _timed(a, chain, count, qsm.synchronize)
    ms = tvb.profile(a.profile_iters)[: len(KERNELS)]
    vsb.close()
    tvb.close()
    vs_bytes = vsharp_alg_bytes(dims, len(radii)) * count
    print(json.dumps({
        "metric": "V-SHARP -> TV-ADMM volumes/s, 8 x 384x384x256 chained on device",
        "value": round(value, 4), "unit": "volumes/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "8 x 384x384x256 @0.6 mm, total field = 12-ellipsoid phantom + 3 "
                               "external dipoles + N(0,0.005); V-SHARP radii 12..1 mm, thr 0.05; "
                               f"TV lambda 1e-3, rho 0.1, tol 0, {a.iters} iterations",
                   "volumes": count, "radii": radii},
        "vsharp": {"ms_per_batch": round(vs_ms, 3), "alg_bytes": int(vs_bytes),
                   "alg_GBs": round(vs_bytes / (vs_ms * 1e-3) / 1e9, 1)},
        "tv": {"ms_per_batch": round(tv_ms, 3), "solves_per_s": round(tv_vps, 3)},
        "roofline": _roofline(np.array(alg_bytes(dims), dtype=np.float64) * count, ms, KERNELS),
    }))


def bench_vjp(a, qsm, phantom):
    """K = 20 unrolled iterations + VJP at 192x192x128 (one volume)."""
    dims, K = (192, 192, 128), 20
    ph = phantom.make_phantom(dims, n_ellipsoids=12, seed=3003)
    grid = qsm.Grid(dims, ph.voxel_size)
    params = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=K)
    n = ph.mask.size
    tf, tm = qsm.DeviceArray.from_numpy(ph.local_field), qsm.DeviceArray.from_numpy(ph.mask)
    tc, tfb = qsm.DeviceArray(n, np.float64), qsm.DeviceArray(n, np.float64)
    u = qsm.TvUnrolled([(grid, ph.bdir, tf.ptr, tm.ptr, tc.ptr)], params)
    u.forward()
    tcb = qsm.DeviceArray.from_numpy(2.0 * (tc.numpy() - ph.chi * ph.mask))

    def step():
        u.forward()
        u.backward([tcb.ptr], [tfb.ptr])
    fwd_ps, fwd_ms = _timed(a, u.forward, 1, qsm.synchronize)
    value, ms_per_step = _timed(a, step, 1, qsm.synchronize)
    u.close()
    N = n
    Nh = ((dims[0] + 1) // 2) * dims[1] * dims[2]
    # backward iteration: adjoint source (c-bar 4N, s-bar 12N in, 12N out, q 3N,
    # acc 8N) + spectrum 8Nh, y/z/y passes 48Nh, C2R 8Nh + 4N
    bwd_bytes = K * (43 * N + 64 * Nh)
    fwd_bytes = sum(alg_bytes(dims)) * K + 3 * N * (K - 1)
    print(json.dumps({
        "metric": "unrolled TV-ADMM (K=20) forward+VJP per second at 192x192x128",
        "value": round(value, 3), "unit": "gradients/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "192x192x128 @1 mm, 12-ellipsoid phantom, lambda 1e-3, rho 0.1, "
                               "K=20, loss ||chi_K - chi_true||^2", "iterations": K},
        "forward_ms": round(fwd_ms, 3), "backward_ms": round(ms_per_step - fwd_ms, 3),
        "alg_GBs": round((fwd_bytes + bwd_bytes) / (ms_per_step * 1e-3) / 1e9, 1),
    }))


def bench_vendor_fft(a, qsm, phantom):
    """Timing comparator: the installed hipFFT's 3-D R2C and C2R (FP32, out of
    place, nothing fused) on the batch this bench times, against one whole
    TV-ADMM iteration of this library on the same volumes (both 3-D
    transforms plus every element-wise step).  hipFFT is only run here, never
    linked into the library."""
    import ctypes as C
    dims = tuple(int(v) for v in a.dims.split(","))
    nb = max(1, a.batch)
    vs = (0.9, 0.9, 1.0)
    params = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=a.iters)
    grid = qsm.Grid(dims, vs)
    keep, vols = [], []
    for v in range(nb):
        p_ = phantom.make_phantom(dims, voxel_size=vs, bdir=(0.0, 0.0, 1.0), n_ellipsoids=40,
                                  seed=1001 + 1000 * v)
        tf, tm = qsm.DeviceArray.from_numpy(p_.local_field), qsm.DeviceArray.from_numpy(p_.mask)
        tc = qsm.DeviceArray(p_.local_field.size, np.float64)
        keep.append((tf, tm, tc))
        vols.append((grid, p_.bdir, tf.ptr, tm.ptr, tc.ptr))
    batch = qsm.TvBatch(vols, params)
    batch.run()
    ms = batch.profile(a.profile_iters)[:len(KERNELS)]
    batch.close()
    iter_ms = float(sum(ms))

    lib = None
    for name in ("libhipfft.so", "/opt/rocm/lib/libhipfft.so"):
        try:
            lib = C.CDLL(name)
            break
        except OSError:
            continue
    if lib is None:
        print(json.dumps({"metric": "hipFFT comparator", "error": "libhipfft.so not found"}))
        return
    HIPFFT_R2C, HIPFFT_C2R = 0x2A, 0x2C
    nx, ny, nz = dims
    N, Nh = nx * ny * nz, (nx // 2 + 1) * ny * nz
    x = qsm.DeviceArray(nb * N, np.float32)
    X = qsm.DeviceArray(2 * nb * Nh, np.float32)
    n = (C.c_int * 3)(nz, ny, nx)  # C order: x fastest
    pf, pi = C.c_void_p(), C.c_void_p()
    for plan, typ, idist, odist in ((pf, HIPFFT_R2C, N, Nh), (pi, HIPFFT_C2R, Nh, N)):
        rc = lib.hipfftPlanMany(C.byref(plan), 3, n, None, 1, idist, None, 1, odist, typ, nb)
        if rc != 0:
            print(json.dumps({"metric": "hipFFT comparator", "error": f"hipfftPlanMany returned {rc}"}))
            return

    def r2c():
        assert lib.hipfftExecR2C(pf, C.c_void_p(x.ptr), C.c_void_p(X.ptr)) == 0

    def c2r():
        assert lib.hipfftExecC2R(pi, C.c_void_p(X.ptr), C.c_void_p(x.ptr)) == 0

    def pair():
        r2c()
        c2r()
    reps = max(a.steps, 20)
    out = {}
    for name, f in (("r2c", r2c), ("c2r", c2r), ("pair", pair)):
        for _ in range(3):
            f()
        qsm.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        qsm.synchronize()
        out[name] = (time.perf_counter() - t0) / reps * 1e3
    lib.hipfftDestroy(pf)
    lib.hipfftDestroy(pi)
    x.free()
    X.free()
    print(json.dumps({
        "metric": f"one TV-ADMM iteration vs hipFFT's bare 3-D R2C + C2R, {nb} x "
                  f"{dims[0]}x{dims[1]}x{dims[2]} FP32",
        "value": round(iter_ms, 4), "unit": "ms per iteration of the batch", "n_gpus": 1,
        "steps": reps, "warmup": 3, "ms_per_step": round(iter_ms, 4), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nb} x {dims[0]}x{dims[1]}x{dims[2]} @0.9x0.9x1.0 mm"},
        "tv_iteration_ms": round(iter_ms, 4),
        "tv_per_kernel_ms": {k: round(float(v), 5) for k, v in zip(KERNELS, ms)},
        "hipfft_r2c_ms": round(out["r2c"], 4), "hipfft_c2r_ms": round(out["c2r"], 4),
        "hipfft_pair_ms": round(out["pair"], 4),
        "pair_over_iteration": round(out["pair"] / iter_ms, 3),
    }))


if __name__ == "__main__":
    main()
