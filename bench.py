"""Benchmark of the k-space first-order PSTD time loop on one MI355X.

Workload (one "step" = one PSTD time step of the whole grid): the 3-D
heterogeneous absorbing nonlinear configuration -- 256^3, dx = 0.1 mm,
PML 20, 8 seeded ellipsoids in c0/rho0, alpha0 = 0.75 dB/(MHz^y cm),
y = 1.5, B/A = 5, plane pressure source, full x-y sensor plane recording
p, p_max and p_rms, Float32 kernels.  Inputs are resident in HBM before the
timed region; the timed region is K steps between two device syncs.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints ONE JSON line (metric, value = voxel-updates/s, the step roofline on
64 B per voxel-update, the roofline of the dominant kernel -- the largest
share of the step -- measured with HIP events on the simulation stream, each
kernel's figures, the FP32-vs-FP64 rel-L2 over the workload's full 1200
steps, and the oracle timed on the host cores as the CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "k-wave-julia_amd"))

METRIC = "voxel-updates/s + % HBM roofline, FP32 vs FP64 form at ≤1e-4 rel-L2"
HBM_PEAK_GBS = 8000.0


def kernel_bytes_per_voxel(name: str, nx: int, s: int = 4) -> float:
    """Algorithmic (compulsory) bytes per grid voxel of one launch of each kernel.

    Directional passes read p, u_d, rho_d, rho0 and write u_d, rho_d; the x
    pass also reads rho_y, rho_z, c0^2 and writes the half spectrum
    (Nx/2+1 complex per Nx reals); spectral passes read and write the half
    spectrum; the c2r pass reads it and writes p.
    """
    half = 2 * s * (nx // 2 + 1) / nx
    return {
        "k_dir_pass_y": 6 * s,
        "k_dir_pass_z": 6 * s,
        "k_dir_pass_x_eos": 9 * s + half,
        "k_spec_pass_y_fwd": 2 * half,
        "k_spec_pass_z_mul": 2 * half,
        "k_spec_pass_y_inv": 2 * half,
        "k_c2r_x": half + s,
        "k_plane_c2r": half + s,
    }.get(name, 0.0)


# Kernel functions of one 3-D step (as rocprofv3 names them) and the bench.py pass
# classes each one runs as: the strided directional pass is one kernel launched twice
# per step (y and z lines).
KERNELS = {
    "k_dir_pass<float,256,false> (strided y/z pass)": ("k_dir_pass_y", "k_dir_pass_z"),
    "k_dir_pass<float,256,true> (x pass + EOS r2c)": ("k_dir_pass_x_eos",),
    "k_spec_pass<float,256,0> (spectral y forward)": ("k_spec_pass_y_fwd",),
    "k_spec_pass<float,256,2> (spectral z fwd*M*inv)": ("k_spec_pass_z_mul",),
    "k_spec_pass<float,256,1> (spectral y inverse)": ("k_spec_pass_y_inv",),
    "k_c2r_x<float,256> (c2r x pass)": ("k_c2r_x",),
    "k_plane_c2r<float,256,256> (y inverse + c2r per z plane)": ("k_plane_c2r",),
}


def measured_traffic(kernel_class: str, n: int):
    """HBM bytes per launch of a pass class on the n^3 workload from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel
    calibration), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"pmc_pstd_step_{n}cube_fp32.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{kernel_class}]:"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    return None


def host_cpu():
    """CPU model and the CPUs this process may use (the CPU baseline's host)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, visible, os.cpu_count()


def cpu_baseline(n_steps_max: int = 20, budget_s: float = 20.0):
    """Oracle (FP32 restatement, scipy.fft threads) on the same workload, truncated."""
    import numpy as np
    from oracle import kwave_oracle as O
    from kwave_amd.configs import ellipsoids_3d

    # the GPU host allots 16 CPUs to each GPU: the FFT thread pool is sized to that share
    model, visible, total = host_cpu()
    workers = min(visible, 16)
    O.set_fft_workers(workers)
    kg, med, src, sen, kw = ellipsoids_3d(Nt=1200)
    stamps = []
    t_start = time.perf_counter()

    class Stop(Exception):
        pass

    def cb(t, Nt, p):
        stamps.append(time.perf_counter())
        # step 1 is the warm-up; stop once the budget or the step cap is used
        if len(stamps) >= 2 and (stamps[-1] - stamps[0] > budget_s or len(stamps) > n_steps_max):
            raise Stop()

    try:
        O.kspace_first_order(kg, med, src, sen, pml_size=kw["pml_size"], pml_alpha=kw["pml_alpha"],
                             smooth_p0=False, data_cast=np.float32, progress_callback=cb)
    except Stop:
        pass
    n = len(stamps) - 1
    dt = stamps[-1] - stamps[0]
    vox = 256 ** 3
    return {"value": vox * n / dt, "unit": "voxel-updates/s", "cores": workers, "kind": "port",
            "sample": f"oracle FP32 restatement of the 256^3 workload: 1 warm-up step + {n} timed steps "
                      f"({dt:.1f} s; scipy.fft with {workers} threads, NumPy elementwise on 1 thread), "
                      f"setup {stamps[0] - t_start:.1f} s excluded",
            "cpu_model": model, "cpus_visible": visible, "cpus_in_host": total}


def bench_batch(a, tune: bool = True, seq_steps_max: int = 200):
    """64 uneven 2-D firings on 256^2..1024^2 grids; batched (one launch per pass per line
    length, line lengths on concurrent streams, hipGraph step) against the same simulations
    run one after another.  With tune=False the batch runs once, with the default shapes."""
    import numpy as np
    import kwave_amd as K
    from kwave_amd.configs import uneven_2d_batch

    sims, kw = uneven_2d_batch(Nt=a.warmup + a.steps)
    descs, keep = [], []
    for s in sims:
        d, k, _ = K.build_desc(*s, data_cast=np.float32, **kw)
        descs.append(d)
        keep.append(k)
    vox = sum(int(np.prod(s[0].grid_size())) for s in sims)

    def run_batch(persistent=False):
        b = K.Batch(descs, keep, use_graph=True, persistent=persistent)
        b.run(a.warmup)
        b.synchronize()
        u1 = time.perf_counter()
        b.run(a.steps)
        b.synchronize()
        u2 = time.perf_counter()
        p = b.sensor(0, "p")
        assert np.all(np.isfinite(p))
        b.close()
        return u2 - u1

    K.tuning.clear_table()
    print(f"[bench] batch of {len(sims)}: {a.steps} steps", file=sys.stderr, flush=True)
    default_secs = run_batch()
    tune_s, table, persistent_secs = None, {}, None
    t1, t2 = 0.0, default_secs
    if tune:
        # autotuner: trials of the whole batch (untimed)
        tt = time.perf_counter()
        K.tuning.autotune_batch(sims, steps=10, data_cast=np.float32, **kw)
        tune_s = time.perf_counter() - tt
        table = K.tuning.table()
        t2 = run_batch()
        persistent_secs = run_batch(persistent=True)
        K.tuning.clear_table()
    batched = vox * a.steps / (t2 - t1)
    # the same simulations, one plain launch sequence each (truncated run; a rate)
    seq_steps = max(1, min(a.steps, seq_steps_max))
    print(f"[bench] the same {len(sims)} simulations one after another: {seq_steps} steps each",
          file=sys.stderr, flush=True)
    tot = 0.0
    for d, k in zip(descs, keep):
        s = K.Simulation(d, k)
        s.run(a.warmup)
        s.synchronize()
        u1 = time.perf_counter()
        s.run(seq_steps)
        s.synchronize()
        tot += time.perf_counter() - u1
        s.close()
    sequential = vox * seq_steps / tot
    sizes = sorted({s[0].Nx for s in sims})
    out = {
        "metric": METRIC, "value": batched, "unit": "voxel-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": (t2 - t1) * 1e3 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch of {len(sims)} 2D firings, uneven grids {sizes}, heterogeneous c0, "
                               f"alpha0=0.5 y=1.1, linear array source/sensor, hipGraph step",
                   "voxels_per_step": vox, "parallelism": "single GPU"},
        "sequential_value": sequential, "batched_over_sequential": batched / sequential,
        "untuned_batched_value": vox * a.steps / default_secs, "autotune_s": tune_s,
        "persistent_batched_value": vox * a.steps / persistent_secs if persistent_secs else None,
        "tuning_table": {f"{k[1]}:{k[2]}": v for k, v in sorted(table.items())},
        "sequential_sample": f"{seq_steps} steps per simulation after {a.warmup} warm-up steps",
    }
    return out


def bench_adjoint(a, nseq_max: int = 8):
    """Adjoint form: 2-D 512^2 ring-array misfit gradient, FP32, checkpoints every 50 steps.

    Observed data come from a forward run on the true model (untimed); the timed
    region is gradient_c0 on the homogeneous starting model for `a.firings`
    firings (each: forward with checkpoints, recompute, adjoint sweep).
    """
    import numpy as np
    import kwave_amd as K
    from kwave_amd.configs import ring_2d_adjoint

    kg, c_true, c_start, firings, kw = ring_2d_adjoint()
    firings = firings[: a.firings]
    med_true = K.KWaveMedium(sound_speed=c_true, density=1000.0)
    med0 = K.KWaveMedium(sound_speed=c_start, density=1000.0)
    print(f"[bench] observed data: {len(firings)} stacked forward firings", file=sys.stderr, flush=True)
    outs = K.kspace_first_order_firings(kg, med_true, [f[0] for f in firings], [f[1] for f in firings],
                                        data_cast=np.float32, **kw)
    obs = [o["p"] for o in outs]
    del outs
    srcs, sens = [f[0] for f in firings], [f[1] for f in firings]
    K.gradient_c0(kg, med0, srcs[:2], sens[:2], obs[:2], checkpoint_every=50, data_cast=np.float32, **kw)
    print(f"[bench] stacked gradient over {len(firings)} firings", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    loss_total, g = K.gradient_c0(kg, med0, srcs, sens, obs, checkpoint_every=50, data_cast=np.float32, **kw)
    t2 = time.perf_counter()
    assert np.all(np.isfinite(g))
    vox = int(np.prod(kg.grid_size()))
    secs = t2 - t1
    # the same gradient one firing after another (stack=1) on a subset, for the batched/sequential ratio
    nseq = min(nseq_max, len(firings))
    print(f"[bench] sequential gradient over {nseq} firings", file=sys.stderr, flush=True)
    t3 = time.perf_counter()
    _, g_seq = K.gradient_c0(kg, med0, srcs[:nseq], sens[:nseq], obs[:nseq], checkpoint_every=50,
                             data_cast=np.float32, stack=1, **kw)
    t4 = time.perf_counter()
    seq_per_firing = (t4 - t3) / nseq
    # accuracy (untimed): FP64 adjoint of firing 0 against central differences of the FP64
    # forward run on `a.fd_pixels` seeded pixels (h = 1e-3 m/s, dt held), and the FP32
    # adjoint of the same firing against the FP64 one
    fd = None
    if a.fd_pixels > 0:
        s0, r0 = firings[0]
        _, g64 = K.gradient_c0(kg, med0, s0, r0, obs[0], checkpoint_every=50, data_cast=np.float64, **kw)
        _, g32 = K.gradient_c0(kg, med0, s0, r0, obs[0], checkpoint_every=50, data_cast=np.float32, **kw)

        def loss(c):
            m = K.KWaveMedium(sound_speed=c, density=1000.0)
            p = K.kspace_first_order(kg, m, s0, r0, data_cast=np.float64, **kw)["p"]
            return 0.5 * float(np.sum((p - obs[0]) ** 2))

        rng = np.random.default_rng(42)
        pix = rng.choice(vox, size=a.fd_pixels, replace=False)
        gf, h, errs = g64.ravel(order="F"), 1e-3, []
        for i in pix:
            e = np.zeros(vox)
            e[i] = h
            e = e.reshape(c_start.shape, order="F")
            d = (loss(c_start + e) - loss(c_start - e)) / (2 * h)
            errs.append(abs(d - gf[i]))
        gmax = float(np.abs(gf).max())
        fd = {"pixels": int(a.fd_pixels), "h_m_per_s": h, "firing": 0,
              "max_abs_err_over_max_grad": max(errs) / gmax,
              "fp32_vs_fp64_grad_rel_l2": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))}
    out = {
        "metric": METRIC, "value": vox * kg.Nt * len(firings) / secs, "unit": "voxel-updates/s", "n_gpus": 1,
        "steps": kg.Nt * len(firings), "warmup": kg.Nt, "ms_per_step": secs * 1e3 / (kg.Nt * len(firings)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"adjoint dL/dc0, 2D {kg.Nx}^2, ring of 128 point elements, {len(firings)} firings, "
                               f"Nt={kg.Nt}, checkpoints every 50 steps; value counts forward time steps "
                               f"(each one forward + one recompute + one adjoint step)",
                   "parallelism": "single GPU"},
        "seconds_per_firing": secs / len(firings), "sequential_seconds_per_firing": seq_per_firing,
        "stacked_over_sequential_speedup": seq_per_firing / (secs / len(firings)),
        "loss": loss_total, "fd_check": fd,
    }
    return out


def bench_c4(a):
    """Precision form: 512^3 layered medium, FP16 map storage, FP32 arithmetic, two sensor sets."""
    import numpy as np
    import kwave_amd as K
    from kwave_amd.configs import layered_3d_fp16

    kg, med, src, sensors, kw = layered_3d_fp16(Nt=a.warmup + a.steps + a.profile_steps)
    if a.c4_maps == "fp32":
        kw["medium_storage"] = None   # comparison run: maps in the kernel precision
    res = {}
    for tag, sen in zip(("p_max_volume", "p_4096_points"), sensors):
        desc, keep, _ = K.build_desc(kg, med, src, sen, data_cast=np.float32, **kw)
        sim = K.Simulation(desc, keep)
        del keep
        sim.run(a.warmup)
        sim.synchronize()
        t1 = time.perf_counter()
        sim.run(a.steps)
        sim.synchronize()
        t2 = time.perf_counter()
        prof = sim.profile(a.profile_steps) if a.profile_steps > 0 else {}
        field = sen.record[0]
        v = sim.sensor(field)
        assert np.all(np.isfinite(v))
        sim.close()
        res[tag] = {"value": kg.Nx ** 3 * a.steps / (t2 - t1), "ms_per_step": (t2 - t1) * 1e3 / a.steps,
                    "kernels_ms": {k: round(x["ms"], 4) for k, x in prof.items()}}
    first = res["p_max_volume"]
    out = {
        "metric": METRIC, "value": first["value"], "unit": "voxel-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": first["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D layered {kg.Nx}^3 dx=0.2mm PML 24, FP16 c0/rho0 maps, B/A=6, alpha0=0.6 y=1.5, "
                               f"bowl source; sensor sets: full-volume p_max / 4096 points p",
                   "parallelism": "single GPU"},
        "sensor_sets": res,
    }
    print(json.dumps(out))


def cpu_baseline_elastic(n: int = 128, n_steps_max: int = 12, budget_s: float = 15.0):
    """Oracle elastic restatement (FP32, scipy.fft threads) on an n^3 cut of the workload."""
    import numpy as np
    from oracle import kwave_oracle as O
    from kwave_amd.configs import two_layer_elastic_3d

    model, visible, total = host_cpu()
    workers = min(visible, 16)
    O.set_fft_workers(workers)
    kg, med, src, sen, kw = two_layer_elastic_3d(N=n, Nt=200)
    stamps = []

    class Stop(Exception):
        pass

    def cb(t, Nt, p):
        stamps.append(time.perf_counter())
        if len(stamps) >= 2 and (stamps[-1] - stamps[0] > budget_s or len(stamps) > n_steps_max):
            raise Stop()

    try:
        O.pstd_elastic(kg, med, src, sen, data_cast=np.float32, progress_callback=cb, **kw)
    except Stop:
        pass
    k = len(stamps) - 1
    dt = stamps[-1] - stamps[0]
    return {"value": n ** 3 * k / dt, "unit": "voxel-updates/s", "cores": workers, "kind": "port",
            "sample": f"oracle FP32 elastic restatement on the {n}^3 cut of the workload: 1 warm-up step + {k} "
                      f"timed steps ({dt:.1f} s; scipy.fft with {workers} threads)",
            "cpu_model": model, "cpus_visible": visible, "cpus_in_host": total}


def bench_elastic(a):
    """Elastic stress-velocity PSTD (pstd_elastic_3d): 3-D two-layer 256^3 FP32 workload."""
    import numpy as np
    import kwave_amd as K
    from kwave_amd.configs import two_layer_elastic_3d, elastic_pass_bytes
    from kwave_amd.elastic import ElasticSimulation, build_elastic_desc

    N = a.N
    kg, med, src, sen, kw = two_layer_elastic_3d(N=N, Nt=a.warmup + a.steps + a.profile_steps)
    desc, keep, _ = build_elastic_desc(kg, med, src, sen, data_cast=np.float32, **kw)
    sim = ElasticSimulation(desc, keep)
    del keep
    sim.run(a.warmup)
    sim.synchronize()
    t1 = time.perf_counter()
    sim.run(a.steps)
    sim.synchronize()
    t2 = time.perf_counter()
    prof = sim.profile(a.profile_steps) if a.profile_steps > 0 else {}
    pm = sim.sensor("p_max")
    assert np.all(np.isfinite(pm)) and pm.max() > 0
    sim.close()
    vox = N ** 3
    ms_step = (t2 - t1) * 1e3 / a.steps
    table = elastic_pass_bytes(4)
    passes, step_bytes = {}, 0.0
    for name, (b, launches) in table.items():
        step_bytes += b
        if name in prof:
            ms = prof[name]["ms"]
            per_launch = b / launches * vox
            passes[name] = {"ms": round(ms, 4), "launches_per_step": launches,
                            "GB_s": round(per_launch / (ms * 1e-3) / 1e9, 1)}
    dom = max(passes, key=lambda k: passes[k]["ms"] * passes[k]["launches_per_step"]) if passes else None
    out = {
        "metric": "voxel-updates/s (elastic, FP32)", "value": vox * a.steps / (t2 - t1), "unit": "voxel-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D two-layer elastic {N}^3 dx=0.1mm PML 10 (fluid cp1540 / solid cp3000 cs1500), "
                               "p0 ball, z-plane sensor p + p_max", "parallelism": "single GPU"},
        "step_roofline": {"bound": "hbm", "bytes_per_voxel_step": step_bytes,
                          "achieved": round(step_bytes * vox / (ms_step * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_bytes * vox / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)},
        "dominant_pass": dom, "passes": passes,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_elastic()
    print(json.dumps(out))


def bench_fft(a):
    """FFT-component micro-benchmark, as run_benchmarks.jl:328-393 times its FFT plans.

    GPU: the line-FFT derivative along each axis (forward line transform, the
    staggered i k multiplier, inverse transform: the spectral_gradient! operator,
    fft_utils.jl:126-182) on FP32 device arrays through the allocation-free entry
    point, on torch's current stream, HIP events around `--steps` calls after 3
    untimed ones; per call the time, the algorithmic bytes (the field read and the
    derivative written, 8 B per voxel) over it, and 2 x 5 N log2 N flop per line.
    CPU: the reference's own sizes (1-D rfft single-threaded, 2-D and 3-D rfftn on
    every core) with scipy.fft in FP64 as the FFTW stand-in, median of repeated runs.
    """
    import ctypes as C
    import numpy as np
    import torch
    from kwave_amd import _lib

    lib = _lib.load()
    steps = a.steps if a.steps != 1200 else 20
    stream = torch.cuda.current_stream()
    gpu = []
    for shape in [(256,), (1024,), (4096,), (256, 256), (512, 512), (1024, 1024), (64, 64, 64), (128, 128, 128),
                  (256, 256, 256), (512, 512, 512)]:
        nd = len(shape)
        n = int(np.prod(shape))
        N = (C.c_int64 * 3)(*(list(shape) + [1] * (3 - nd)))
        dd = (C.c_double * 3)(*([1e-4] * 3))
        f = torch.randn(n, dtype=torch.float32, device="cuda")
        o = torch.empty_like(f)

        def call(axis):
            _lib.check(lib.kw_spectral_gradient_noalloc(_lib.context(0), 4, nd, N, dd, axis, 1,
                                                        C.c_void_p(f.data_ptr()), C.c_void_p(o.data_ptr()),
                                                        C.c_void_p(stream.cuda_stream)))
        for axis in range(nd):
            for _ in range(3):
                call(axis)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                call(axis)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1) / steps
            flops = 2 * 5 * n * np.log2(shape[axis])
            gpu.append({"shape": list(shape), "axis": axis, "us": round(ms * 1e3, 2),
                        "GB_s": round(8 * n / (ms * 1e-3) / 1e9, 1), "GFLOP_s": round(flops / (ms * 1e-3) / 1e9, 1)})
            print(f"[fft] GPU derivative {shape} axis {axis}: {ms * 1e3:8.2f} us  {gpu[-1]['GB_s']:7.1f} GB/s",
                  file=sys.stderr, flush=True)
        del f, o
        torch.cuda.empty_cache()
    cpu = []
    if not a.no_cpu:
        import scipy.fft as sf
        nproc = len(os.sched_getaffinity(0))
        sizes = [((n,), 1) for n in (128, 256, 512, 1024, 2048, 4096, 8192, 16384)]
        sizes += [((n, n), nproc) for n in (64, 128, 256, 512, 1024)]
        sizes += [((n, n, n), nproc) for n in (32, 64, 128, 256)]
        rng = np.random.default_rng(42)
        for shape, workers in sizes:
            x = rng.random(shape)
            sf.rfftn(x, workers=workers)
            reps = 200 if len(shape) == 1 else (20 if np.prod(shape) <= 1 << 20 else 5)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                sf.rfftn(x, workers=workers)
                ts.append(time.perf_counter() - t0)
            med = float(np.median(ts))
            nn = int(np.prod(shape))
            cpu.append({"shape": list(shape), "workers": workers, "median_us": round(med * 1e6, 2),
                        "GFLOP_s": round(5 * nn * np.log2(nn) / med / 1e9, 2)})
    big = [r for r in gpu if r["shape"] == [256, 256, 256]]
    val = sum(r["GB_s"] for r in big) / len(big)
    print(json.dumps({
        "metric": "line-FFT derivative throughput (FFT micro-benchmark)", "value": round(val, 1), "unit": "GB/s",
        "n_gpus": 1, "steps": steps, "warmup": 3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "spectral derivative (forward line FFT, staggered i k, inverse) of a 256^3 FP32 "
                               "field along x, y and z, mean over the axes; other sizes in gpu_rows",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(val, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(val / HBM_PEAK_GBS, 3), "traffic": None},
        "gpu_rows": gpu,
        "cpu_fft_rows": cpu,
        "cpu_fft_note": "scipy.fft.rfftn FP64 (FFTW stand-in) on the reference's micro-benchmark sizes",
    }))


def copy_probe():
    """A plain elementwise copy of known size (1 GiB read + 1 GiB written per launch), for
    calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE against algorithmic bytes."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 accuracy run")
    ap.add_argument("--accuracy-nt", type=int, default=1200,
                    help="steps of the FP32-vs-FP64 accuracy runs (the workload's Nt)")
    ap.add_argument("--config", choices=["c2", "batch", "adjoint", "c4", "elastic", "fft"], default="c2")
    ap.add_argument("--firings", type=int, default=8)
    ap.add_argument("--fd-pixels", type=int, default=32)
    ap.add_argument("--c4-maps", choices=["fp16", "fp32"], default="fp16")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the short batched and adjoint runs reported under 'forms'")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the multi-GPU slab decomposition "
              "is a design, not part of this build", file=sys.stderr)
        sys.exit(2)

    if a.copy_probe:
        return copy_probe()
    if a.config == "batch":
        if a.steps == 1200:
            a.steps = 1500
        print(json.dumps(bench_batch(a)))
        return
    if a.config == "adjoint":
        print(json.dumps(bench_adjoint(a)))
        return
    if a.config == "c4":
        if a.steps == 1200:
            a.steps = 200
        return bench_c4(a)
    if a.config == "elastic":
        if a.steps == 1200:
            a.steps = 200
        return bench_elastic(a)
    if a.config == "fft":
        return bench_fft(a)

    import numpy as np
    import kwave_amd as K
    from kwave_amd.configs import ellipsoids_3d, ellipsoids_3d_bytes_per_voxel_update

    t0 = time.perf_counter()
    nt = a.warmup + a.steps + a.profile_steps
    kg, med, src, sen, kw = ellipsoids_3d(N=a.N, Nt=nt)
    desc, keep, info = K.build_desc(kg, med, src, sen, pml_size=kw["pml_size"], pml_alpha=kw["pml_alpha"],
                                    smooth_p0=False, data_cast=np.float32)
    sim = K.Simulation(desc, keep)
    setup_s = time.perf_counter() - t0
    print(f"[bench] setup {setup_s:.1f}s; warmup {a.warmup} steps", file=sys.stderr, flush=True)
    sim.run(a.warmup)
    sim.synchronize()
    t1 = time.perf_counter()
    sim.run(a.steps)
    sim.synchronize()
    t2 = time.perf_counter()
    print(f"[bench] timed {a.steps} steps in {t2 - t1:.3f}s", file=sys.stderr, flush=True)
    prof = sim.profile(a.profile_steps) if a.profile_steps > 0 else {}
    sim.synchronize()
    pmax = sim.sensor("p_max")
    if not np.all(np.isfinite(pmax)):
        raise RuntimeError("non-finite sensor data")
    sim.close()

    fp64_check = None
    if not a.no_fp64:
        fp64_check = accuracy_check(a.N, a.accuracy_nt)
    forms = None
    if not a.no_forms:
        try:
            forms = other_forms()
        except Exception as e:   # the headline line stands on its own; say what failed
            print(f"[bench] forms failed: {e!r}", file=sys.stderr, flush=True)
            forms = {"error": repr(e)}

    N = a.N
    vox = N ** 3
    secs = t2 - t1
    value = vox * a.steps / secs
    step_b = ellipsoids_3d_bytes_per_voxel_update(4)
    ms_step = secs * 1e3 / a.steps
    # per kernel function: launches per step, mean launch time (HIP events on the sim
    # stream), algorithmic bytes per launch and the rate they imply
    kernels = {}
    for kname, classes in KERNELS.items():
        got = [prof[c]["ms"] for c in classes if c in prof]
        if not got:
            continue
        bpl = sum(kernel_bytes_per_voxel(c, N) for c in classes) / len(classes) * vox
        avg = sum(got) / len(got)
        kernels[kname] = {"launches_per_step": len(got), "avg_launch_ms": avg, "bytes_per_launch": bpl,
                          "GB_s": bpl / (avg * 1e-3) / 1e9, "share_of_step": len(got) * avg / ms_step,
                          "traffic_bytes_per_launch": None}
        t = [measured_traffic(c, N) for c in classes]
        if all(x is not None for x in t):
            kernels[kname]["traffic_bytes_per_launch"] = sum(t) / len(t)
    roof = None
    if kernels:
        # dominant kernel = largest share of the step time
        dname, d = max(kernels.items(), key=lambda kv: kv[1]["share_of_step"])
        roof = {"bound": "hbm", "kernel": dname, "achieved": d["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["GB_s"] / HBM_PEAK_GBS, "traffic": d["traffic_bytes_per_launch"],
                "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_launch_ms"],
                "share_of_step": d["share_of_step"]}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D heterogeneous absorbing nonlinear PSTD, {N}^3, dx=0.1mm, PML 20, "
                               f"8 seeded ellipsoids, alpha0=0.75 y=1.5 B/A=5, plane p source, "
                               f"z-plane sensor p/p_max/p_rms",
                   "grid": [N, N, N], "precision": "fp32", "parallelism": "single GPU"},
        "roofline": roof,
        "step_roofline": {"bytes_per_voxel_update": step_b, "achieved": step_b * value / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_b * value / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
        "fp32_vs_fp64_rel_l2": fp64_check,
        "setup_s": setup_s,
        "forms": forms,
    }
    if not a.no_cpu:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def other_forms():
    """Short runs of the batched and adjoint forms beside the headline step, so that one
    default run records their batched-over-sequential ratios: the 64-simulation uneven
    2-D batch (300 steps, default launch shapes; 50 steps per simulation one after another)
    and the 128-firing 512^2 ring gradient stacked (2 firings one after another)."""
    b = bench_batch(argparse.Namespace(warmup=20, steps=300), tune=False, seq_steps_max=50)
    d = bench_adjoint(argparse.Namespace(firings=128, fd_pixels=0), nseq_max=2)
    return {
        "batch_64_uneven_2d": {k: b[k] for k in ("value", "unit", "steps", "ms_per_step", "sequential_value",
                                                 "batched_over_sequential", "sequential_sample")},
        "adjoint_ring512_128_firings": {"seconds": d["seconds_per_firing"] * 128,
                                        "seconds_per_firing": d["seconds_per_firing"],
                                        "sequential_seconds_per_firing": d["sequential_seconds_per_firing"],
                                        "stacked_over_sequential_speedup": d["stacked_over_sequential_speedup"],
                                        "forward_voxel_steps_per_s": d["value"], "sequential_sample": "2 firings"},
    }


def accuracy_check(N: int, nt: int):
    """FP32 kernels against the same kernels compiled in FP64 (parity-tested against the
    FP64 oracle) on the full workload: rel-L2 of p, p_max and p_rms over all `nt` steps
    (the workload's Nt = 1200 by default, so the pulse launched at z = 21 has crossed the
    z = 200 sensor plane), untimed."""
    import numpy as np
    import kwave_amd as K
    from kwave_amd.configs import ellipsoids_3d

    print(f"[bench] accuracy: FP32 vs FP64 over {nt} steps", file=sys.stderr, flush=True)
    kg, med, src, sen, kw = ellipsoids_3d(N=N, Nt=nt)
    outs = {}
    for dc in (np.float32, np.float64):
        d, k, _ = K.build_desc(kg, med, src, sen, pml_size=kw["pml_size"], pml_alpha=kw["pml_alpha"],
                               smooth_p0=False, data_cast=dc)
        with K.Simulation(d, k) as s:
            s.run(nt)
            s.synchronize()
            outs[dc] = {f: s.sensor(f) for f in ("p", "p_max", "p_rms")}
    res = {f: float(np.linalg.norm(outs[np.float32][f] - outs[np.float64][f]) / np.linalg.norm(outs[np.float64][f]))
           for f in ("p", "p_max", "p_rms")}
    res["steps"] = nt
    res["sensor_peak_pa"] = float(np.abs(outs[np.float64]["p_max"]).max())
    return res


if __name__ == "__main__":
    main()
