"""Throughput of the fused collide-and-stream-pull step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line on stdout.  A step is one fused kernel launch over the
whole lattice (inputs resident in HBM).  `value` is MLUPS = cells * K / t
(wall clock of the K timed steps, device synchronised on both sides).  The
roofline block divides the algorithmic bytes per launch (2 Q sizeof(storage)
per cell, plus the alpha state for Newton ELBM) by the average kernel
duration measured with HIP events on the stream the kernels run on.  The CPU
baseline is the FP64 oracle (a restatement of the reference's CPU loop,
OpenMP over x) timed on a bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + % HBM roofline; entropic MLUPS learned-α vs Newton at equal residual"
PEAK_HBM_GBS = 8000.0

# name -> (description, config kwargs, init, bytes per storage value)
WORKLOADS = {
    "d3q19-bgk-256-f32": dict(
        desc="D3Q19 SRT BGK 256^3 periodic, SoA two-lattice pull, FP32 shifted storage, "
             "FP32 compute, tau 0.55, Guo + constant F=(1e-6,0,0), rho0=1, u0 noise +-1e-3 (seed 42)",
        cfg=dict(lattice="D3Q19", nx=256, ny=256, nz=256, tau=0.55, forcing_scheme="Guo",
                 force="Constant", amplitude=(1e-6, 0, 0), storage="F32", compute="F32",
                 shifted=True),
        init=("noise", 1e-3, 42), steps=200),
    "d3q19-bgk-256-f64": dict(
        desc="D3Q19 SRT BGK 256^3 periodic, FP64 storage and compute (reference order), "
             "tau 0.55, Guo + constant F=(1e-6,0,0), u0 noise +-1e-3 (seed 42)",
        cfg=dict(lattice="D3Q19", nx=256, ny=256, nz=256, tau=0.55, forcing_scheme="Guo",
                 force="Constant", amplitude=(1e-6, 0, 0), storage="F64", compute="F64",
                 shifted=False),
        init=("noise", 1e-3, 42), steps=200),
    "d3q27-bgk-384-f16": dict(
        desc="D3Q27 SRT BGK 384^3 Taylor-Green U0=0.05, tau 0.51, FP32 compute, shifted FP16 storage",
        cfg=dict(lattice="D3Q27", nx=384, ny=384, nz=384, tau=0.51, forcing_scheme="None_",
                 force="None_", storage="F16", compute="F32", shifted=True),
        init=("tgv", 0.05), steps=200),
    "d3q27-bgk-384-f32": dict(
        desc="D3Q27 SRT BGK 384^3 Taylor-Green U0=0.05, tau 0.51, FP32 compute, shifted FP32 storage",
        cfg=dict(lattice="D3Q27", nx=384, ny=384, nz=384, tau=0.51, forcing_scheme="None_",
                 force="None_", storage="F32", compute="F32", shifted=True),
        init=("tgv", 0.05), steps=200),
    "d3q27-bgk-384-f64": dict(
        desc="D3Q27 SRT BGK 384^3 Taylor-Green U0=0.05, tau 0.51, FP32 compute, shifted FP64 storage",
        cfg=dict(lattice="D3Q27", nx=384, ny=384, nz=384, tau=0.51, forcing_scheme="None_",
                 force="None_", storage="F64", compute="F32", shifted=True),
        init=("tgv", 0.05), steps=100),
    "d2q9-elbm-2048-newton": dict(
        desc="D2Q9 SRT Entropic 2048^2, nu=1e-5 (tau=0.50003), Kolmogorov k=4 A=1e-5, "
             "u0 noise +-1e-3 (seed 42), FP32 compute + FP64 Newton alpha solve",
        cfg=dict(lattice="D2Q9", collision="ELBM", nx=2048, ny=2048, nz=1, tau=0.50003,
                 forcing_scheme="Guo", force="Kolmogorov", amplitude=(1e-5, 0, 0),
                 wavelength=(512.0, 1, 1), storage="F32", compute="F32", shifted=True),
        init=("noise", 1e-3, 42), steps=200),
    "d2q9-elbm-2048-warm": dict(
        desc="D2Q9 SRT Entropic 2048^2 as d2q9-elbm-2048-newton, the learned solver "
             "without its network: the stored alpha, one Newton correction accepted by "
             "|G| <= 1e-10 or its Taylor certificate, else plain Newton to the same "
             "tolerance",
        cfg=dict(lattice="D2Q9", collision="ELBM", nx=2048, ny=2048, nz=1, tau=0.50003,
                 forcing_scheme="Guo", force="Kolmogorov", amplitude=(1e-5, 0, 0),
                 wavelength=(512.0, 1, 1), storage="F32", compute="F32", shifted=True,
                 alpha_mode="WARM_RESIDUAL", residual_tol=1e-10),
        init=("noise", 1e-3, 42), steps=200),
    "d2q9-elbm-2048-learned": dict(
        desc="D2Q9 SRT Entropic 2048^2 as d2q9-elbm-2048-newton, alpha from the MLP "
             "(9 normalised non-equilibrium moments -> 32 -> 32 -> 1, FP32 master weights, "
             "f16 matrix-core inference) trained on the GPU on (features, alpha) logged from "
             "the first 1000 Newton steps, then one Newton correction accepted by |G| <= "
             "1e-10 or its Taylor certificate where the reference Newton provably reaches "
             "the same root, else plain Newton to the same tolerance",
        cfg=dict(lattice="D2Q9", collision="ELBM", nx=2048, ny=2048, nz=1, tau=0.50003,
                 forcing_scheme="Guo", force="Kolmogorov", amplitude=(1e-5, 0, 0),
                 wavelength=(512.0, 1, 1), storage="F32", compute="F32", shifted=True,
                 alpha_mode="NEWTON", residual_tol=1e-10),
        init=("noise", 1e-3, 42), steps=200,
        learn=dict(log_steps=1000, fraction=1e-4, capacity=500000, iters=3000, batch=4096,
                   lr=2e-3)),
    "d2q9-elbm-2048-newton-residual": dict(
        desc="D2Q9 SRT Entropic 2048^2 as d2q9-elbm-2048-newton, plain Newton at the "
             "residual tolerance: from the stored alpha until |G| <= 1e-10 (at most 50 "
             "evaluations, alpha = 2 outside (1, alphaMax)), one launch",
        cfg=dict(lattice="D2Q9", collision="ELBM", nx=2048, ny=2048, nz=1, tau=0.50003,
                 forcing_scheme="Guo", force="Kolmogorov", amplitude=(1e-5, 0, 0),
                 wavelength=(512.0, 1, 1), storage="F32", compute="F32", shifted=True,
                 alpha_mode="NEWTON_RESIDUAL", residual_tol=1e-10),
        init=("noise", 1e-3, 42), steps=200),
    "d3q27-elbm-256-newton-residual": dict(
        desc="D3Q27 SRT Entropic 256^3 as d3q27-elbm-256-newton, plain Newton at the "
             "residual tolerance: from the stored alpha until |G| <= 1e-10, one launch",
        cfg=dict(lattice="D3Q27", collision="ELBM", nx=256, ny=256, nz=256, tau=0.50006,
                 forcing_scheme="None_", force="None_", storage="F32", compute="F32",
                 shifted=True, alpha_mode="NEWTON_RESIDUAL", residual_tol=1e-10),
        init=("spectral", 0.1 / 3 ** 0.5, 16, 4.0, 42), steps=50),
    "d3q27-elbm-256-newton": dict(
        desc="D3Q27 SRT Entropic 256^3 decaying turbulence, nu=2e-5 (tau=0.50006), no force, "
             "solenoidal random-phase u0 with E(k) ~ k^4 exp(-2 (k/4)^2) on shells 1..16 "
             "(seed 42), per-component rms 0.1 cs, FP32 compute + FP64 reference Newton alpha",
        cfg=dict(lattice="D3Q27", collision="ELBM", nx=256, ny=256, nz=256, tau=0.50006,
                 forcing_scheme="None_", force="None_", storage="F32", compute="F32",
                 shifted=True),
        init=("spectral", 0.1 / 3 ** 0.5, 16, 4.0, 42), steps=50),
    "d3q27-elbm-256-warm": dict(
        desc="D3Q27 SRT Entropic 256^3 as d3q27-elbm-256-newton, the learned solver "
             "without its network (stored alpha, up to two certified Newton steps, else "
             "plain Newton to the same tolerance)",
        cfg=dict(lattice="D3Q27", collision="ELBM", nx=256, ny=256, nz=256, tau=0.50006,
                 forcing_scheme="None_", force="None_", storage="F32", compute="F32",
                 shifted=True, alpha_mode="WARM_RESIDUAL", residual_tol=1e-10),
        init=("spectral", 0.1 / 3 ** 0.5, 16, 4.0, 42), steps=50),
    "d3q27-elbm-256-learned": dict(
        desc="D3Q27 SRT Entropic 256^3 as d3q27-elbm-256-newton, alpha from the MLP "
             "(27 normalised non-equilibrium moments -> 64 -> 64 -> 1, BF16 weights on the "
             "matrix cores, FP32 accumulate) trained on the GPU on (features, alpha) logged "
             "from the first 200 Newton steps, one certified correction kept where the "
             "reference Newton provably reaches the same root, else plain Newton to the same "
             "tolerance",
        cfg=dict(lattice="D3Q27", collision="ELBM", nx=256, ny=256, nz=256, tau=0.50006,
                 forcing_scheme="None_", force="None_", storage="F32", compute="F32",
                 shifted=True, alpha_mode="NEWTON", residual_tol=1e-10),
        init=("spectral", 0.1 / 3 ** 0.5, 16, 4.0, 42), steps=50,
        learn=dict(log_steps=200, fraction=1e-4, capacity=500000, iters=3000, batch=4096,
                   lr=2e-3)),
    "d2q9-bgk-turbulent2d-2048-f32": dict(
        desc="D2Q9 SRT BGK 2048^2, tau 0.6, EDM + Turbulent2D spectral force (injection "
             "amplitude 5e-4 on shells 5..7, energy removal 1e-6 on shells 1..2; the force "
             "field is a function of the last stored fields and is rebuilt when they change, "
             "by a direct transform of the 8 removal modes; fields stored every 10th step), "
             "FP32 shifted storage and compute, u0 noise +-1e-3",
        cfg=dict(lattice="D2Q9", nx=2048, ny=2048, nz=1, tau=0.6,
                 forcing_scheme="ExactDifferenceMethod", force="Turbulent2D",
                 amplitude=(5e-4, 5e-4, 5e-4), force_kmin=5, force_kmax=7,
                 removal_amplitude=(1e-6, 1e-6, 1e-6), removal_kmin=1, removal_kmax=2,
                 storage="F32", compute="F32", shifted=True),
        init=("noise", 1e-3, 42), steps=200, store_every=10),
    "d2q9-bgk-128-f64": dict(
        desc="D2Q9 SRT BGK 128^2 FP64, tau 0.6, Guo + Kolmogorov F_x=1e-5 sin(2 pi 4 y/128), "
             "u0 noise +-1e-3 (seed 42) -- cache-resident parity config",
        cfg=dict(lattice="D2Q9", nx=128, ny=128, nz=1, tau=0.6, forcing_scheme="Guo",
                 force="Kolmogorov", amplitude=(1e-5, 0, 0), wavelength=(32.0, 1, 1),
                 storage="F64", compute="F64", shifted=False),
        init=("noise", 1e-3, 42), steps=1000),
}
DEFAULT = "d3q19-bgk-256-f32"
SIZES = {"F64": 8, "F32": 4, "F16": 2}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Hip:
    """The few HIP runtime calls bench.py needs (events and one stream)."""

    def __init__(self):
        for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
            try:
                self.L = ctypes.CDLL(p)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")
        vp = ctypes.c_void_p
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        self.L.hipEventRecord.argtypes = [vp, vp]
        self.L.hipEventSynchronize.argtypes = [vp]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        self.L.hipEventDestroy.argtypes = [vp]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.L.hipStreamSynchronize.argtypes = [vp]
        self.L.hipStreamDestroy.argtypes = [vp]
        self.L.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def stream(self):
        s = ctypes.c_void_p()
        self.ok(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s

    def event(self):
        e = ctypes.c_void_p()
        self.ok(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e

    def record(self, e, s):
        self.ok(self.L.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self.ok(self.L.hipEventSynchronize(b), "hipEventSynchronize")
        t = ctypes.c_float()
        self.ok(self.L.hipEventElapsedTime(ctypes.byref(t), a, b), "hipEventElapsedTime")
        return float(t.value)

    def sync(self):
        self.ok(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def copy_calibration(self, nbytes: int, reps: int) -> float:
        """hipMemcpy device-to-device of nbytes, reps times (the runtime's own
        copy kernel): a plain copy of known size for scaling rocprofv3's
        FETCH_SIZE / WRITE_SIZE counters.  Returns GB/s."""
        vp = ctypes.c_void_p
        self.L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.L.hipFree.argtypes = [vp]
        self.L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        a, b = vp(), vp()
        self.ok(self.L.hipMalloc(ctypes.byref(a), nbytes), "hipMalloc")
        self.ok(self.L.hipMalloc(ctypes.byref(b), nbytes), "hipMalloc")
        self.ok(self.L.hipMemset(a, 0, nbytes), "hipMemset")
        self.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.ok(self.L.hipMemcpy(b, a, nbytes, 3), "hipMemcpy")   # 3: device to device
        self.sync()
        t = time.perf_counter() - t0
        self.L.hipFree(a)
        self.L.hipFree(b)
        return 2.0 * nbytes * reps / t / 1e9


def make_config(spec):
    import metalbm_amd as m
    kw = dict(spec)
    conv = {"lattice": m.LatticeType, "collision": m.CollisionType,
            "forcing_scheme": m.ForcingSchemeType, "force": m.ForceType,
            "equilibrium": m.EquilibriumType, "storage": m.Precision, "compute": m.Precision,
            "alpha_mode": m.AlphaMode}
    for k, e in conv.items():
        if k in kw and isinstance(kw[k], str):
            kw[k] = e[kw[k]]
    return m.Config(**kw)


def init(a, spec):
    kind = spec[0]
    if kind == "noise":
        a.init_noise(1.0, spec[1], spec[2])
    elif kind == "tgv":
        a.init_taylor_green(1.0, spec[1])
    elif kind == "spectral":
        a.init_spectral(1.0, spec[1], spec[2], spec[3], spec[4])
    else:
        raise ValueError(kind)


def bytes_per_cell(cfg, wl_cfg) -> int:
    q = cfg.q
    b = 2 * q * SIZES[wl_cfg["storage"]]
    # the spectral forces are a per-cell force field read by the step
    if wl_cfg.get("force") in ("ConstantShell", "EnergyRemoval", "Turbulent2D"):
        b += cfg.d * (8 if wl_cfg["compute"] == "F64" else 4)
    # Every entropic solver reads and writes the alpha state (Algorithm.h:103-106).
    if wl_cfg.get("collision") == "ELBM":
        b += 2 * (8 if wl_cfg["compute"] == "F64" else 4)   # alpha warm-start state r+w
    return b


def host_info() -> dict:
    """CPU model, logical CPUs, the CPUs this process may run on, physical cores."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "cpus_available": len(os.sched_getaffinity(0)),
            "physical_cores": len(cores) or None,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads() -> int:
    """Every CPU this process may use: its affinity set, within the host's CPU
    share when OMP_NUM_THREADS states one."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(wl, max_seconds=10.0):
    """The FP64 oracle, built -march=native on this host, on a bounded sample
    of the same workload: one thread (the reference's serial per-rank loop,
    Computation.h:102-119) and OpenMP over x on every CPU available."""
    import numpy as np

    from oracle import oracle as o
    o.lib_native()
    c = wl["cfg"]
    names = {"D2Q9": 3, "D3Q19": 9, "D3Q27": 10}
    cfg = o.Config(lattice=names[c["lattice"]],
                   collision=2 if c.get("collision") == "ELBM" else 1,
                   forcing_scheme={"None_": 1, "Guo": 2, "ShanChen": 3,
                                   "ExactDifferenceMethod": 4}[c.get("forcing_scheme", "Guo")],
                   force={"None_": 0, "Constant": 4, "Sinusoidal": 5, "Kolmogorov": 6,
                          "ConstantShell": 7, "EnergyRemoval": 8,
                          "Turbulent2D": 9}[c.get("force", "None_")],
                   equilibrium=3, nx=c["nx"], ny=c["ny"], nz=c.get("nz", 1), tau=c["tau"],
                   amplitude=c.get("amplitude", (0, 0, 0)),
                   wavelength=c.get("wavelength", (1, 1, 1)))
    spec = wl["init"]
    if spec[0] == "noise":
        u = o.noise_velocity(cfg, spec[2], spec[1])
    elif spec[0] == "spectral":
        u = o.spectral_velocity(cfg, spec[1], spec[2], spec[3], spec[4])
    else:
        u = o.taylor_green(cfg, spec[1])
    f = o.init_equilibrium(cfg, np.ones(cfg.cells), u)
    field_force = cfg.force >= o.FORCE_CONSTANT_SHELL

    def timed(threads):
        if field_force:
            # spectral force: numpy transforms of the field before every step
            fspec = dict(kmin=c["force_kmin"], kmax=c["force_kmax"],
                         removal_amplitude=c["removal_amplitude"],
                         removal_kmin=c["removal_kmin"], removal_kmax=c["removal_kmax"])
            every = wl.get("store_every", 1)
            ones = np.ones(cfg.cells)
            o.lib().ref_set_threads(threads)

            def steps(k):
                o.run_field_force(cfg, f, k, ones, u, fspec, store_every=every,
                                  threads=threads)
        else:
            def steps(k):
                o.run(cfg, f, k, threads=threads, native=True)
        t0 = time.perf_counter()
        steps(1)
        t1 = time.perf_counter() - t0
        n = int(max(1, min(50, (max_seconds - t1) / max(t1, 1e-6))))
        t0 = time.perf_counter()
        steps(n)
        t = time.perf_counter() - t0
        return round(cfg.cells * n / t / 1e6, 3), n

    threads = cpu_threads()
    par, npar = timed(threads)
    ser, nser = timed(1)
    where = f"same {cfg.nx}x{cfg.ny}x{cfg.nz} workload after 1 warm-up step" + (
        " (force field by numpy FFT before every step)" if field_force else "")
    return {"value": par, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle built -march=native, OpenMP over x on {threads} threads, "
                      f"{npar} steps of the {where}",
            "serial": {"value": ser, "unit": "MLUPS", "cores": 1,
                       "sample": f"the same oracle on one thread (the reference's serial "
                                 f"per-rank loop), {nser} steps of the {where}"},
            "host": host_info()}


def entropic_comparison(which: str, segments: int) -> dict:
    """Learned alpha vs Newton at equal residual on an entropic workload, all
    four solvers from the same state.  C4 (D2Q9 2048^2): the first 1000 steps
    run the reference Newton with the sample log on (the training data), the
    network is trained on the GPU, then every solver runs steps 1001..5000
    from that state in `segments` timed segments.  C5 (D3Q27 256^3): 200
    logged Newton steps, training, then 50 timed steps per solver."""
    import numpy as np

    import metalbm_amd as m
    name = {"c4": "d2q9-elbm-2048-learned", "c5": "d3q27-elbm-256-learned"}[which]
    wl = WORKLOADS[name]
    learn = wl["learn"]
    total = 4000 if which == "c4" else 50
    seg = total // segments
    cfg = make_config({**wl["cfg"], "alpha_mode": "NEWTON"})
    out = {"workload": name, "residual_tol": 1e-10, "timed_steps": total,
           "segment_steps": seg}
    with m.Algorithm(cfg) as a:
        init(a, wl["init"])
        t0 = time.perf_counter()
        a.sample_log_enable(learn["capacity"], learn["fraction"], 1)
        a.iterate(learn["log_steps"])
        a.sample_log_disable()
        a.synchronize()
        t1 = time.perf_counter()
        l0, l1 = a.mlp_train(learn["iters"], learn["batch"], learn["lr"], seed=1)
        t2 = time.perf_counter()
        out["training"] = {"logged_samples": len(a.sample_log()[1]),
                           "log_steps": learn["log_steps"], "log_seconds": round(t1 - t0, 3),
                           "train_iterations": learn["iters"],
                           "train_seconds": round(t2 - t1, 3), "loss_first": l0,
                           "loss_last": l1}
        f0, al0 = a.populations(), a.alpha()
        N = cfg.cells
        for mode in (m.AlphaMode.NEWTON, m.AlphaMode.NEWTON_RESIDUAL,
                     m.AlphaMode.WARM_RESIDUAL, m.AlphaMode.LEARNED):
            a.set_populations(f0)
            a.set_alpha(al0)
            a.set_alpha_mode(mode, 1e-10)
            rows, wall = [], 0.0
            for _ in range(segments):
                a.reset_elbm_stats()
                a.synchronize()
                t0 = time.perf_counter()
                a.iterate(seg)
                a.synchronize()
                t = time.perf_counter() - t0
                wall += t
                st = a.elbm_stats()
                rows.append([round(N * seg / t / 1e6, 1),
                             round(st["newton_iterations"] / max(1, st["newton"]), 4),
                             round(st["learned_accepted"] / max(1, st["newton"]), 4),
                             st["fallback"]])
            r = np.array(rows, dtype=float)
            out[mode.name] = {"MLUPS": round(N * seg * segments / wall / 1e6, 1),
                              "evaluations_per_solved_cell": round(float(r[:, 1].mean()), 4),
                              "accepted_by_correction": round(float(r[:, 2].mean()), 4),
                              "fallbacks": int(r[:, 3].sum()),
                              "segments": rows}
            log(f"entropic {which} {mode.name}: {out[mode.name]['MLUPS']} MLUPS, "
                f"{out[mode.name]['evaluations_per_solved_cell']} evaluations per solved cell")
    nres = out["NEWTON_RESIDUAL"]["MLUPS"]
    out["learned_vs_newton_residual"] = round(out["LEARNED"]["MLUPS"] / nres, 3)
    out["learned_vs_reference_newton"] = round(out["LEARNED"]["MLUPS"] / out["NEWTON"]["MLUPS"], 3)
    out["learned_vs_warm_residual"] = round(out["LEARNED"]["MLUPS"] /
                                            out["WARM_RESIDUAL"]["MLUPS"], 3)
    out["segment_columns"] = ["MLUPS", "evaluations_per_solved_cell", "accepted_by_correction",
                              "fallbacks"]
    return out


def accuracy_study(nsteps: int, every: int, n: int) -> dict:
    """Kinetic-energy decay of the D3Q27 Taylor-Green vortex (U0 = 0.05, tau =
    0.51, n^3) with FP32 compute and shifted FP16 / FP32 / FP64 storage against
    the FP64 reference-order build (bitwise equal to the FP64 oracle on every
    tested configuration).  E(t) = (1/N) sum 1/2 rho |u_h|^2 every `every` steps."""
    import numpy as np

    import metalbm_amd as m
    base = dict(lattice=m.LatticeType.D3Q27, nx=n, ny=n, nz=n, tau=0.51,
                forcing_scheme=m.ForcingSchemeType.None_, force=m.ForceType.None_)
    runs = {"f64-reference": dict(storage=m.Precision.F64, compute=m.Precision.F64, shifted=False),
            "f16-shifted": dict(storage=m.Precision.F16, compute=m.Precision.F32, shifted=True),
            "f32-shifted": dict(storage=m.Precision.F32, compute=m.Precision.F32, shifted=True),
            "f64-shifted": dict(storage=m.Precision.F64, compute=m.Precision.F32, shifted=True),
            "f32-unshifted": dict(storage=m.Precision.F32, compute=m.Precision.F32, shifted=False)}
    series = {}
    for name, prec in runs.items():
        t0 = time.perf_counter()
        with m.Algorithm(m.Config(**base, **prec)) as a:
            a.init_taylor_green(1.0, 0.05)
            e = []
            for _ in range(nsteps // every):
                a.iterate(every, store=True)
                e.append(a.reduce()[1])
        series[name] = np.array(e)
        log(f"accuracy {name}: E({nsteps}) = {e[-1]:.10e} ({time.perf_counter() - t0:.1f} s)")
    ref = series["f64-reference"]
    out = {"workload": f"D3Q27 BGK {n}^3 Taylor-Green U0=0.05 tau=0.51, {nsteps} steps, "
                       f"E every {every}", "E_reference_first_last": [ref[0], ref[-1]]}
    for name, e in series.items():
        if name == "f64-reference":
            continue
        rel = np.abs(e - ref) / np.abs(ref)
        decay_err = np.abs((e[-1] - e[0]) - (ref[-1] - ref[0])) / abs(ref[-1] - ref[0])
        out[name] = {"max_rel_E_error": float(rel.max()), "final_rel_E_error": float(rel[-1]),
                     "rel_error_of_E_decay": float(decay_err)}
    return out


def decay_study(nsteps: int, every: int, n: int) -> dict:
    """Energy and enstrophy of D3Q27 entropic decaying turbulence (n^3,
    tau = 0.50006, spectral initial field on shells 1..16 peaked at 4 with
    per-component rms 0.1 cs, seed 42) with FP32 compute and shifted FP32
    storage -- reference Newton, plain Newton at the residual tolerance,
    certified correction of the stored alpha, learned alpha -- against the FP64 reference-order build (equal to the FP64 oracle
    to <= 1e-12 on every tested configuration).  E and Omega every `every` steps
    (E = (1/N) sum 1/2 rho |u_h|^2, Omega = (1/N) sum 1/2 |curl u_h|^2 by
    central differences).  The learned run logs (features, alpha) over its
    first `every` steps with the Newton solver, trains, then switches; the
    retrained run also logs the Newton solves of the cells its correction
    defers and retrains on them every 5 intervals."""
    import numpy as np

    import metalbm_amd as m
    base = dict(lattice=m.LatticeType.D3Q27, collision=m.CollisionType.ELBM, nx=n, ny=n, nz=n,
                tau=0.50006, forcing_scheme=m.ForcingSchemeType.None_, force=m.ForceType.None_,
                residual_tol=1e-10)
    f32 = dict(storage=m.Precision.F32, compute=m.Precision.F32, shifted=True)
    runs = {"f64-reference": (dict(storage=m.Precision.F64, compute=m.Precision.F64,
                                   shifted=False), m.AlphaMode.NEWTON),
            "f32-newton": (f32, m.AlphaMode.NEWTON),
            "f32-newton-residual": (f32, m.AlphaMode.NEWTON_RESIDUAL),
            "f32-warm-residual": (f32, m.AlphaMode.WARM_RESIDUAL),
            "f32-learned": (f32, m.AlphaMode.LEARNED),
            "f32-learned-retrained": (f32, m.AlphaMode.LEARNED)}
    urms = 0.1 / 3 ** 0.5
    series, stats = {}, {}
    for name, (prec, mode) in runs.items():
        t0 = time.perf_counter()
        with m.Algorithm(m.Config(**base, **prec)) as a:
            a.init_spectral(1.0, urms, 16, 4.0, 42)
            out = []
            for k in range(nsteps // every):
                if mode == m.AlphaMode.LEARNED and k == 0:
                    # fractions chosen so the expected sample count fits the
                    # capacity: the kept samples, and so the network, do not
                    # depend on the order the wavefronts reach the log
                    a.sample_log_enable(500000, min(1.0, 2.5e-4 * (256 / n) ** 3), 1)
                    a.iterate(every, store=True)
                    a.sample_log_disable()
                    a.mlp_train(3000, 4096, 2e-3, seed=1)
                    a.set_alpha_mode(mode, 1e-10)
                    a.reset_elbm_stats()
                else:
                    if k == 0:
                        a.set_alpha_mode(mode, 1e-10)
                    if name == "f32-learned-retrained" and k % 5 == 1:
                        # log the deferred Newton solves, retrain every 5 intervals
                        if k > 1:
                            a.mlp_train(1000, 4096, 1e-3, seed=k)
                        # logging stays on for 5 intervals (500 steps)
                        a.sample_log_enable(500000, min(1.0, 5e-5 * (256 / n) ** 3), k)
                    a.iterate(every, store=True)
                out.append(a.reduce()[1:])
                if (k + 1) % 10 == 0:
                    log(f"decay {name}: step {(k + 1) * every}, E, Omega = {out[-1]}")
            stats[name] = a.elbm_stats()
        series[name] = np.array(out)
        log(f"decay {name}: E, Omega at {nsteps} = {out[-1]} ({time.perf_counter() - t0:.1f} s) "
            f"{stats[name]}")
    ref = series["f64-reference"]
    res = {"workload": f"D3Q27 SRT Entropic {n}^3 decaying turbulence tau=0.50006, {nsteps} "
                       f"steps, E and Omega every {every}",
           "reference_E_first_last": [ref[0, 0], ref[-1, 0]],
           "reference_Omega_first_last": [ref[0, 1], ref[-1, 1]]}
    for name, v in series.items():
        if name == "f64-reference":
            res[name] = {"elbm_stats": stats[name]}
            continue
        rel = np.abs(v - ref) / np.abs(ref)
        res[name] = {"max_rel_E_error": float(rel[:, 0].max()),
                     "max_rel_Omega_error": float(rel[:, 1].max()),
                     "final_rel_E_error": float(rel[-1, 0]),
                     "final_rel_Omega_error": float(rel[-1, 1]), "elbm_stats": stats[name]}
    return res


def diagnostics_bench(K: int, W: int, n: int, cpu: bool) -> dict:
    """Spectral diagnostics of the stored fields (lbm_vorticity + lbm_spectra,
    physical scaling) on the D3Q19 n^3 FP32 workload's fields: wall time per
    call pair, with the numpy oracle on the same fields as the CPU baseline."""
    import numpy as np

    import metalbm_amd as m
    wl = WORKLOADS[DEFAULT]
    cfg = make_config({**wl["cfg"], "nx": n, "ny": n, "nz": n})
    a = m.Algorithm(cfg)
    init(a, ("spectral", 0.02, 8, 4.0, 42))
    a.iterate(2, store=True)
    a.synchronize()
    for _ in range(W):
        a.vorticity("physical")
        a.spectra("physical")
    t0 = time.perf_counter()
    for _ in range(K):
        _, ens = a.vorticity("physical", field=False)
        e, f = a.spectra("physical")
    wall = (time.perf_counter() - t0) / K
    t0 = time.perf_counter()
    for _ in range(K):
        a.vorticity("physical")
    wall_field = (time.perf_counter() - t0) / K
    N, D = cfg.cells, 3
    M = n * n * (n // 2 + 1)
    # algorithmic bytes per launch of the library's own kernels: curl reads D
    # spectra and writes 3 (16 B per mode each); the shell bins read D spectra
    kbytes = {"k_curl_hat": (D + 3) * 16 * M, "k_shell_bins": D * 16 * M}
    out = {"metric": "spectral diagnostics: enstrophy (spectral curl) + energy/forcing "
                     "spectra per call pair, results on the host, vorticity field on the device",
           "value": round(wall * 1e3, 3), "unit": "ms", "higher_is_better": False,
           "steps": K, "warmup": W, "n_gpus": 1, "dtype": "f64",
           "data": "synthetic (spectral initial field, 2 FP32 steps)",
           "config": {"workload": f"d3q19-bgk-{n}-f32 stored fields", "cells": N,
                      "modes": M, "kernel_bytes": kbytes,
                      "vorticity_with_host_field_ms": round(wall_field * 1e3, 3),
                      "enstrophy": ens, "energy_sum": float(e.sum())}}
    rho, u, al, F = a.store_fields()
    a.close()
    if cpu:
        from oracle import spectral as sp
        shape = (n, n, n)
        t0 = time.perf_counter()
        sp.vorticity(u, shape, sp.PHYSICAL)
        sp.spectrum(u, shape, sp.PHYSICAL)
        sp.spectrum(F, shape, sp.PHYSICAL)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(t * 1e3, 1), "unit": "ms", "cores": 1,
                               "kind": "port",
                               "sample": f"one call pair of the numpy oracle (pocketfft, one "
                                         f"thread) on the same {n}^3 fields"}
    return out


def hdf5_bench(n: int) -> dict:
    """Field and distribution files in the reference's HDF5 layout (FP64,
    FFTW-padded last extent) for the D3Q19 n^3 FP32 workload: wall time of
    lbm_write_fields_hdf5 (first call and a later one), lbm_write_distribution_hdf5
    and lbm_read_distribution_hdf5 into a temporary directory (device-host
    copy, padding and libhdf5 included; the files are removed afterwards)."""
    import shutil
    import tempfile

    import metalbm_amd as m
    wl = WORKLOADS[DEFAULT]
    cfg = make_config({**wl["cfg"], "nx": n, "ny": n, "nz": n})
    d = tempfile.mkdtemp(prefix="metalbm_hdf5_")
    out = {"metric": "HDF5 field / distribution files, reference layout", "unit": "s",
           "config": {"workload": f"d3q19-bgk-{n}-f32 fields and populations", "cells": cfg.cells}}
    try:
        with m.Algorithm(cfg) as a:
            init(a, wl["init"])
            a.iterate(2, store=True)
            a.synchronize()
            # the first file call also loads libhdf5 and plans the vorticity
            # transforms; the second is the steady cost
            tf = time.perf_counter()
            a.write_fields(d, 2)
            first = time.perf_counter() - tf
            a.iterate(1, store=True)
            a.synchronize()
            t0 = time.perf_counter()
            fpath = a.write_fields(d, 3)
            t1 = time.perf_counter()
            dpath = a.write_distribution(d, 2)
            t2 = time.perf_counter()
            a.read_distribution(d, 2)
            t3 = time.perf_counter()
            fb, db = os.path.getsize(fpath), os.path.getsize(dpath)
        out.update({"write_fields_s": round(t1 - t0, 3), "first_write_fields_s": round(first, 3),
                    "field_file_bytes": fb,
                    "write_distribution_s": round(t2 - t1, 3),
                    "read_distribution_s": round(t3 - t2, 3), "distribution_file_bytes": db,
                    "write_distribution_GBps": round(db / (t2 - t1) / 1e9, 2),
                    "read_distribution_GBps": round(db / (t3 - t2) / 1e9, 2)})
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def stream_ceiling(K: int, W: int, name: str = DEFAULT) -> dict:
    """A BGK workload's step kernel (the headline by default) against the
    bandwidth probes of the library (lbm_bandwidth_probe) on the same
    lattices: the vector kernel's access pattern without its arithmetic
    (2 Q streams), its Q reads alone, its Q writes alone, and one contiguous
    copy of the same bytes.  Average launch time from HIP events on the
    stream the kernels run on."""
    import metalbm_amd as m
    wl = WORKLOADS[name]
    if wl["cfg"].get("collision", "BGK") != "BGK":
        raise SystemExit("--stream-ceiling takes a BGK workload")
    cfg = make_config(wl["cfg"])
    hip = Hip()
    a = m.Algorithm(cfg)
    s = hip.stream()
    a.set_stream(s.value)
    init(a, wl["init"])
    N, Q = cfg.cells, cfg.q
    pop = N * SIZES[wl["cfg"]["storage"]]

    def timed(run):
        run(W)
        a.synchronize()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, s)
        run(K)
        hip.record(e1, s)
        return hip.elapsed_ms(e0, e1) / K

    out = {"metric": "achieved HBM bandwidth of the step kernel vs probes of its access pattern",
           "unit": "GB/s", "steps": K, "warmup": W,
           "config": {"workload": name, "cells": N, "bytes_per_population": pop,
                      "step_kernel": m.kernel_shape(cfg)[0].name.lower()}}
    ms = timed(lambda k: a.iterate(k))
    out["step"] = {"us": round(ms * 1e3, 2), "bytes": 2 * Q * pop,
                   "GB/s": round(2 * Q * pop / ms / 1e6, 1)}
    for kind, name, nbytes in ((0, "pull_copy", 2 * Q * pop), (1, "read", Q * pop),
                               (2, "write", Q * pop), (3, "copy", 2 * Q * pop)):
        ms = timed(lambda k, kind=kind: a.bandwidth_probe(kind, k))
        out[name] = {"us": round(ms * 1e3, 2), "bytes": nbytes,
                     "GB/s": round(nbytes / ms / 1e6, 1)}
        log(f"probe {name}: {out[name]}")
    a.close()
    out["kernel_vs_pull_copy"] = round(out["step"]["GB/s"] / out["pull_copy"]["GB/s"], 4)
    out["kernel_vs_copy"] = round(out["step"]["GB/s"] / out["copy"]["GB/s"], 4)
    out["peak"] = PEAK_HBM_GBS
    return out


def traffic_of(name):
    """HBM bytes per launch of the workload's step kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic_<name>.txt,
    counters scaled by a copy of known size), or None when not measured."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.txt")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        for line in fh:
            if line.startswith("hbm_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-entropic", action="store_true",
                    help="skip the learned-vs-Newton entropic comparison of the default run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 1 GiB (counter calibration)")
    ap.add_argument("--accuracy", action="store_true",
                    help="D3Q27 Taylor-Green kinetic-energy decay, FP16/FP32/FP64 shifted "
                         "storage vs the FP64 build (prints its own JSON line)")
    ap.add_argument("--accuracy-n", type=int, default=384)
    ap.add_argument("--decay", action="store_true",
                    help="D3Q27 entropic decaying turbulence: energy and enstrophy of the FP32 "
                         "solvers vs the FP64 build (prints its own JSON line)")
    ap.add_argument("--decay-n", type=int, default=256)
    ap.add_argument("--diagnostics", action="store_true",
                    help="time lbm_vorticity + lbm_spectra on D3Q19 n^3 stored fields "
                         "(prints its own JSON line)")
    ap.add_argument("--diagnostics-n", type=int, default=256)
    ap.add_argument("--hdf5", action="store_true",
                    help="time the HDF5 field / distribution files of D3Q19 n^3 "
                         "(prints its own JSON line)")
    ap.add_argument("--hdf5-n", type=int, default=128)
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="the headline step kernel against bandwidth probes of its own "
                         "access pattern (prints its own JSON line)")
    args = ap.parse_args(argv)
    if args.stream_ceiling:
        print(json.dumps(stream_ceiling(args.steps or 100, args.warmup, args.workload)),
              flush=True)
        return 0
    if args.hdf5:
        print(json.dumps(hdf5_bench(args.hdf5_n)), flush=True)
        return 0
    if args.diagnostics:
        print(json.dumps(diagnostics_bench(args.steps or 20, args.warmup, args.diagnostics_n,
                                           not args.no_cpu_baseline)), flush=True)
        return 0
    if args.decay:
        res = decay_study(args.steps or 3000, 100, args.decay_n)
        print(json.dumps({"metric": "energy/enstrophy error vs FP64", **res}))
        return 0
    if args.accuracy:
        res = accuracy_study(args.steps or 2000, 100, args.accuracy_n)
        print(json.dumps({"metric": "kinetic-energy decay error vs FP64", **res}))
        return 0
    if args.copy_calibration:
        gbs = Hip().copy_calibration(1 << 30, 5)
        print(json.dumps({"copy_bytes_per_launch": 1 << 30, "launches": 5, "GB/s": round(gbs, 1)}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs one process-local domain "
              "on one MI355X (no halo exchange); run it with --gpus 1", file=sys.stderr)
        return 2
    wl = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else wl["steps"]
    W = args.warmup

    import metalbm_amd as m
    hip = Hip()
    cfg = make_config(wl["cfg"])
    log(f"workload {args.workload}: {wl['desc']}")
    a = m.Algorithm(cfg)
    s = hip.stream()
    a.set_stream(s.value)
    init(a, wl["init"])
    learn = wl.get("learn")
    learned = {}
    if learn:
        # Training data: (features, alpha) of the GPU's own Newton solves over
        # the first log_steps steps, then Adam on the GPU.
        t0 = time.perf_counter()
        a.sample_log_enable(learn["capacity"], learn["fraction"], 1)
        a.iterate(learn["log_steps"])
        a.sample_log_disable()
        a.synchronize()
        t1 = time.perf_counter()
        l0, l1 = a.mlp_train(learn["iters"], learn["batch"], learn["lr"], seed=1)
        t2 = time.perf_counter()
        n_samples = len(a.sample_log()[1])
        a.set_alpha_mode(m.AlphaMode.LEARNED, wl["cfg"].get("residual_tol", 1e-10))
        learned = {"logged_samples": n_samples, "log_steps": learn["log_steps"],
                   "log_seconds": round(t1 - t0, 3), "train_iterations": learn["iters"],
                   "train_seconds": round(t2 - t1, 3), "loss_first": l0, "loss_last": l1}
        log(f"learned alpha: {learned}")
    a.iterate(W)
    a.synchronize()
    if wl["cfg"].get("collision") == "ELBM":
        a.reset_elbm_stats()
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(e0, s)
    every = wl.get("store_every")
    if every:
        # storing steps feed the time-dependent spectral forces (Routine.h:121-136)
        for k in range(0, K, every):
            a.iterate(min(every, K - k), store=True)
    else:
        a.iterate(K)
    hip.record(e1, s)
    a.synchronize()
    hip.sync()
    wall = time.perf_counter() - t0
    ev_ms = hip.elapsed_ms(e0, e1)
    N = cfg.cells
    mlups = N * K / wall / 1e6
    kern_s = ev_ms / 1e3 / K
    bpc = bytes_per_cell(cfg, wl["cfg"])
    achieved = bpc * N / kern_s / 1e9
    log(f"wall {wall * 1e3:.2f} ms for {K} steps, events {ev_ms:.2f} ms, "
        f"{mlups:.1f} MLUPS, {achieved:.1f} GB/s algorithmic")
    stats = a.elbm_stats() if wl["cfg"].get("collision") == "ELBM" else None
    repeats = None
    if wl["cfg"].get("collision") != "ELBM" and not every:
        # four more timed runs of the same K steps, for the median the
        # headline configuration asks for (`value` stays the first run)
        runs = [mlups]
        for _ in range(4):
            a.synchronize()
            t0 = time.perf_counter()
            a.iterate(K)
            a.synchronize()
            runs.append(N * K / (time.perf_counter() - t0) / 1e6)
        repeats = {"mlups": [round(r, 1) for r in runs],
                   "median": round(sorted(runs)[len(runs) // 2], 1)}
        log(f"repeats: {repeats}")
    a.close()
    out = {
        "metric": METRIC, "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(wall * 1e3 / K, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"F64": "f64", "F32": "f32", "F16": "f16"}[wl["cfg"]["storage"]] + "-storage/"
                 + {"F64": "f64", "F32": "f32"}[wl["cfg"]["compute"]] + "-compute",
        "data": "synthetic (seeded initial field, see config.init)",
        "config": {"workload": args.workload, "description": wl["desc"],
                   "cells": N, "bytes_per_cell_update": bpc,
                   "init": list(wl["init"]), "parallelism": "single GPU, one domain"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic_of(args.workload),
                     "traffic_source": (f"profiles/traffic_{args.workload}.txt (rocprofv3 --pmc "
                                        f"FETCH_SIZE and WRITE_SIZE of the same kernel, scaled "
                                        f"by a copy of known size)"
                                        if traffic_of(args.workload) else None),
                     "kernel_avg_us": round(kern_s * 1e6, 2)},
    }
    if repeats:
        out["repeats"] = repeats
    if stats:
        out["elbm_stats"] = stats
        out["elbm_stats"]["evaluations_per_solved_cell"] = round(
            stats["newton_iterations"] / max(1, stats["newton"]), 4)
    if learned:
        out["learned_alpha"] = learned
    if args.workload == DEFAULT and not args.no_entropic:
        # the metric's second half: entropic MLUPS, learned alpha vs Newton at
        # equal residual (C4 over its 5000 steps, C5 over 50 steps)
        out["entropic"] = {"c4": entropic_comparison("c4", 8),
                           "c5": entropic_comparison("c5", 1)}
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(wl)
        except Exception as e:  # the baseline is a report, not the measurement
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
