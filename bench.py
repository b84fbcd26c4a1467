"""Benchmark of the Chebyshev moment engine on one MI355X.

Workload (the headline configuration): kpm_1d DOS moments of a disordered
graphene flake, 1024 x 1024 honeycomb cells x 2 orbitals (NH = 2,097,152,
nnz = 8,388,608 stored: 3 bonds per site each way + the disorder diagonal),
t = 1, on-site disorder U[-0.5, 0.5], rescaled by a fixed a = 3.6; NC = 2048,
NR = 64 unit random-phase probes.  One step = one full kpm_1d call (NC/2 =
1024 fused recurrence passes) with the probe block already resident in HBM,
returning what the DOS uses: the NC probe-averaged moments (kpm_1d's default
avg_output = true, as dos_moments calls it).  The metric is moments/s =
NC * NR * nnz / seconds, for the FP32-recurrence form (ComplexF32 vectors,
FP32 recurrence arithmetic, FP64 moment accumulation); the plain FP64 form is
timed beside it and the max moment error of FP32 against FP64 is reported,
and so is one call returning the NR x NC per-probe table (avg_output = false).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kpm.jl_amd"))

HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    """The host CPU's model string (/proc/cpuinfo), for the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def step_bytes(nh: int, nnz: int, nrp: int, elem: int, val_bytes: int) -> int:
    """Algorithmic HBM bytes of one fused pass: H once (values + int32 column
    indices + row pointers) and each vector row once: read v_n, read v_{n-1},
    write v_{n+1}.  Values are 8 / 16 bytes (real / complex); FP32 blocks of
    rows up to 8 words read 4 / 8-byte FP32 copies."""
    return nnz * (val_bytes + 4) + (nh + 1) * 4 + 3 * nh * nrp * elem


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1024, help="honeycomb cells per side")
    ap.add_argument("--nc", type=int, default=2048)
    ap.add_argument("--nr", type=int, default=64)
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline leg")
    ap.add_argument("--skip-fp64", action="store_true", help="omit the FP64 comparison leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--kernel-launches", type=int, default=100)
    ap.add_argument("--shape", default="",
                    help="fix the fused pass's launch shape instead of tuning it: "
                         "'unroll:chunk[:xmap]' (chunk 0: the persistent grid; xmap: chunk "
                         "order, 0 dispatch order, m >= 2 runs of m chunks per XCD), e.g. 2:4:4; "
                         "used to profile the timed calls alone")
    ap.add_argument("--workload", choices=["graphene", "ensemble", "conductivity", "core", "adjoint",
                                           "cpge", "cond_ensemble", "spmm"],
                    default="graphene",
                    help="graphene: the headline single-lattice DOS; ensemble: batched disorder "
                         "realisations (cubic Anderson, one fused launch per step); "
                         "conductivity: Haldane kpm_2d table; core: the reference's own "
                         "benchmark suite (benchmarks/core_bench.jl) beside its published times; "
                         "adjoint: DOS-loss gradient on a 512^2 disordered square lattice; "
                         "cpge: three-current response from a Haldane kpm_3d table; "
                         "cond_ensemble: one antithetic-pair task of the Anderson conductivity "
                         "ensemble pipeline; spmm: one tuned recurrence pass of --op at --nr "
                         "probes and --prec storage")
    ap.add_argument("--op", choices=["graphene", "haldane", "cubic"], default="haldane",
                    help="operator of the spmm workload (side --cells; cubic uses L = 128 "
                         "unless --cells is below 1024)")
    ap.add_argument("--prec", choices=["f64", "f32"], default="f64")
    ap.add_argument("--nreal", type=int, default=256)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU; multi-GPU runs are "
                                   "not part of this build (--gpus must be 1)"}))
        return 2

    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "conductivity":
        return bench_conductivity(args)
    if args.workload == "core":
        return bench_core(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "cpge":
        return bench_cpge(args)
    if args.workload == "cond_ensemble":
        return bench_cond_ensemble(args)
    if args.workload == "spmm":
        return bench_spmm(args)

    import numpy as np

    import kpmhip as K
    from kpmhip import models

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    ctx = K.get_context()
    L, NC, NR = args.cells, args.nc, args.nr
    print(f"[bench] building graphene {L}x{L}x2 ...", file=sys.stderr, flush=True)
    H = (models.graphene(L, L, disorder=0.5, seed=42) / 3.6).tocsc()
    NH, nnz = H.shape[0], H.nnz
    print(f"[bench] NH={NH} nnz={nnz}; probes NR={NR} ...", file=sys.stderr, flush=True)
    # probes drawn on the device from the same xoshiro256++ stream as
    # random_phase_vectors(Xoshiro256pp(42), NH, NR); one host copy feeds the
    # FP32 leg and the CPU baseline, so every leg sees identical probes
    psi = K.DeviceBlock(NH, NR, "f64").random_phase(42).download()
    Hd = K.to_device(H)

    def run_leg(precision: str):
        if args.shape:
            u, ch, *xm = (int(v) for v in args.shape.split(":"))
            if ch > 0:
                K.tune_set_chunked(Hd, NR, precision, unroll=u, chunk=ch,
                                   xcd_grouped=xm[0] if xm else 0)
            else:
                K.tune_set(Hd, NR, precision, unroll=u, grid=0)
            tuned = dict(K.tune_get(Hd, NR, precision), precision=precision, fixed=True)
        else:
            tuned = K.autotune(Hd, NR, precision)
        print(f"[bench] launch shape {precision}: {tuned}", file=sys.stderr, flush=True)
        blk = K.DeviceBlock.from_host(psi, precision)
        # the DOS call: kpm_1d(H, NC, NR; psi_in) with its default avg_output =
        # true (src/moment.jl:339-418, dos_moments at src/frontend.jl:254-256),
        # i.e. the NC probe-averaged moments
        mu = np.zeros(NC, dtype=np.complex128)
        lib = ctx.lib
        for _ in range(args.warmup):
            K._lib.check(lib.kpm_1d_block_avg(ctx.handle, Hd.handle, NC, blk.handle,
                                              K._lib.dptr(mu)), "kpm_1d_block_avg")
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.event_record(0)
        for _ in range(args.steps):
            K._lib.check(lib.kpm_1d_block_avg(ctx.handle, Hd.handle, NC, blk.handle,
                                              K._lib.dptr(mu)), "kpm_1d_block_avg")
        ctx.event_record(1)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        ev_ms = ctx.event_elapsed(0, 1)
        # the per-probe table (avg_output = false): one warm call, timed
        mu_all = np.zeros((NR, NC), dtype=np.complex128, order="F")
        K._lib.check(lib.kpm_1d_block(ctx.handle, Hd.handle, NC, blk.handle,
                                      K._lib.dptr(mu_all)), "kpm_1d_block")
        ctx.synchronize()
        tp = time.perf_counter()
        K._lib.check(lib.kpm_1d_block(ctx.handle, Hd.handle, NC, blk.handle,
                                      K._lib.dptr(mu_all)), "kpm_1d_block")
        per_probe_s = time.perf_counter() - tp
        # the dominant kernel alone: back-to-back fused probe-summed passes on
        # the same stream
        x = K.DeviceBlock.from_host(psi, precision)
        z = K.DeviceBlock.from_host(psi, precision)
        ms = ctypes.c_float(0.0)
        K._lib.check(lib.kpm_bench_step_mode(ctx.handle, Hd.handle, x.handle, z.handle,
                                             args.kernel_launches, 1, ctypes.byref(ms)),
                     "bench_step")
        cms = ctypes.c_float(0.0)
        K._lib.check(lib.kpm_bench_copy(ctx.handle, z.handle, x.handle, 20, ctypes.byref(cms)),
                     "bench_copy")
        copy_gbs = 2 * NH * K.moment._padded_nr(NR, x.prec) * (8 if x.prec else 16) / (cms.value / 20 * 1e-3) / 1e9
        print(f"[bench] {precision}: step kernel {ms.value / args.kernel_launches:.4f} ms, "
              f"copy kernel {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
        del x, z, blk
        tuned["copy_kernel_gbs"] = round(copy_gbs, 1)
        return (mu.copy(), wall / args.steps, ev_ms / 1e3 / args.steps,
                ms.value / args.kernel_launches, tuned, mu_all.copy(), per_probe_s)

    print("[bench] FP32-recurrence leg ...", file=sys.stderr, flush=True)
    mu32, wall32, ev32, kms32, tune32, all32, pp32 = run_leg("f32")
    mu64 = None
    if not args.skip_fp64:
        print("[bench] FP64 leg ...", file=sys.stderr, flush=True)
        mu64, wall64, ev64, kms64, tune64, all64, pp64 = run_leg("f64")

    units = NC * NR * nnz
    value = units / wall32
    nrp32 = K.moment._padded_nr(NR, 1)
    val_bytes = 8 if Hd.is_real else 16  # FP64 values (FP32 copies only for rows of <= 8 words)
    vb32 = val_bytes // 2 if nrp32 <= 16 else val_bytes
    b32 = step_bytes(NH, nnz, nrp32, 8, vb32)
    # the average launch of the timed calls themselves: event time of the
    # timed kpm_1d calls over their NC/2 fused passes (so it also carries the
    # call's probe copy, last reduction and moment download: an upper bound)
    call_ms32 = ev32 * 1e3 / (NC // 2)
    achieved32 = b32 / (call_ms32 * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved32, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved32 / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_step<FP32,LPR=32,UNR=2,KB=4,MODE_DBLA>", "bytes_per_launch": b32,
                "avg_launch_ms": round(call_ms32, 5),
                "launches_per_step": NC // 2,
                "kernel_only_avg_launch_ms": round(kms32, 5),
                "kernel_only_frac": round(b32 / (kms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "tuned": tune32}
    tfile = os.path.join(ROOT, "profiles", "traffic_graphene2M_fp32.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")

    extra = {"ms_per_step_event": round(ev32 * 1e3, 3),
             # the same FP32 recurrence returning the NR x NC per-probe table
             # (kpm_1d avg_output = false, per-workgroup partials of NR columns)
             "per_probe_value": units / pp32,
             "per_probe_ms_per_step": round(pp32 * 1e3, 3),
             "per_probe_frac": round(b32 / (pp32 / (NC // 2)) / 1e9 / HBM_PEAK_GBS, 4),
             "max_abs_avg_vs_per_probe_mean": float(np.max(np.abs(mu32 - all32.mean(axis=0))))}
    if mu64 is not None:
        nrp64 = K.moment._padded_nr(NR, 0)
        b64 = step_bytes(NH, nnz, nrp64, 16, val_bytes)
        avg32 = mu32
        avg64 = mu64
        extra.update({
            "fp64_value": units / wall64,
            "fp64_ms_per_step": round(wall64 * 1e3, 3),
            "fp32_speedup_vs_fp64": round(wall64 / wall32, 3),
            "fp64_kernel_gbs": round(b64 / (ev64 * 1e3 / (NC // 2) * 1e-3) / 1e9, 1),
            "fp64_kernel_frac": round(b64 / (ev64 * 1e3 / (NC // 2) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "fp64_kernel_only_frac": round(b64 / (kms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "fp64_tuned": tune64,
            "max_abs_dmu_fp32_vs_fp64": float(np.max(np.abs(avg32 - avg64))),
            "max_abs_dmu_per_probe": float(np.max(np.abs(all32 - all64))),
            "fp64_per_probe_ms_per_step": round(pp64 * 1e3, 3),
        })

    cpu = None
    if not args.skip_cpu:
        print("[bench] CPU baseline (oracle restatement) ...", file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        threads = min(avail or 1, 16)
        runner = oracle.StepRunner(H, psi)

        def cpu_rate(nthreads, seconds):
            oracle.set_threads(nthreads)
            runner.run(1)
            done, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < seconds:
                runner.run(1)
                done += 1
            tcpu = time.perf_counter() - t0
            return 2 * done * NR * nnz / tcpu, done, tcpu

        rate, done, tcpu = cpu_rate(threads, args.cpu_seconds)
        # 8 threads: the thread count of the reference's published CPU runs
        rate8, done8, tcpu8 = cpu_rate(8, max(args.cpu_seconds / 2, 5.0))
        cpu = {"value": rate, "unit": "moments/s", "cores": threads, "kind": "port",
               "sample": f"{done} doubling steps (2 moments each) of the same graphene "
                         f"NH={NH} NR={NR} workload, FP64, {threads} threads, {tcpu:.1f} s",
               "value_8_threads": rate8,
               "sample_8_threads": f"{done8} doubling steps, 8 threads, {tcpu8:.1f} s",
               "nproc": os.cpu_count(), "cpus_available": avail, "cpu_model": cpu_model()}

    out = {
        "metric": "moments/s (NC*NR*nnz/s), FP32-recurrence kpm_1d DOS",
        "value": value,
        "unit": "moments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall32 * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c32-storage/f64-accumulate",
        "data": "synthetic",
        "config": {"workload": f"graphene {L}x{L}x2 DOS kpm_1d", "NH": NH, "nnz": nnz, "NC": NC,
                   "NR": NR, "disorder": "U[-0.5,0.5]", "a": 3.6,
                   "output": "NC probe-averaged moments (kpm_1d avg_output=true, as dos_moments)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out))
    return 0


# Published times of benchmarks/core_bench.jl (seconds; BenchmarkTools minimum of
# <= 3 samples, 8 Julia threads): benchmarks/results/perf-core-mike-a100.json:7-13
# and perf-core-mike-cpu.json:7-13.  The A100 job passed host arrays to
# chebyshev_iter_single and chebyshev_action!, so those two ran CPU methods.
CORE_PUBLISHED = {
    "kpm_1d_dos_NH65536_NC1024_NR8": (0.056766432, 1.34261305),
    "kpm_1d_dos_NH262144_NC512_NR4": (0.058475161, 1.41826656),
    "kpm_2d_cond_NH16384_NC64_NR4_default": (0.010318251, 0.125779538),
    "kpm_2d_cond_NH16384_NC64_NR4_arr_size3": (0.079784355, 1.40411449),
    "kpm_2d_cond_NH16384_NC64_NR4_arr_size64": (0.006348302, 0.126314814),
    "chebyshev_iter_single_NH262144_NR8": (0.005135421, 0.003572621),
    "chebyshev_action_NH65536_NC1024_NR8_K2": (1.8632278, 1.60499539),
}


def bench_spmm(args) -> int:
    """One recurrence pass (k_step, MODE_DBL: out = 2 H x - z plus the two
    doubling dots) of a lattice operator at NR probes, after the autotuner
    has picked the launch shape; timed with HIP events over --kernel-launches
    back-to-back passes.  Operators: graphene (cells x cells x 2, disorder
    U[-0.5, 0.5], real H), haldane (cells x cells x 2, t2 = 0.1 e^{i pi/2},
    complex H, 10 stored entries per row), cubic (L^3, W = 2, real H).
    Algorithmic bytes: H once (values, column indices, row pointers) and three
    vector rows (read v_n, read v_{n-1}, write v_{n+1})."""
    import ctypes

    import numpy as np

    import kpmhip as K
    from kpmhip import models
    from kpmhip.device import DeviceBlock

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    ctx = K.get_context()
    L = args.cells
    if args.op == "graphene":
        H = (models.graphene(L, L, disorder=0.5, seed=1) / 3.6).tocsr()
    elif args.op == "haldane":
        H = (models.haldane_model(L, L, t=1.0, t2=0.1)[0] / 3.8).tocsr()
    else:
        L = L if L < 1024 else 128
        H = (models.cubic_model(L, W=2.0)[0] / 8.5).tocsr()
    NH, nnz, NR = H.shape[0], H.nnz, args.nr
    real = not (H.dtype.kind == "c" and np.any(H.data.imag != 0))
    pcode = 0 if args.prec == "f64" else 1
    NRp = K.moment._padded_nr(NR, pcode)
    vb = (8 if real else 16) // (2 if (pcode == 1 and NRp <= 16) else 1)  # FP32 values on rows <= 8 words
    byts = nnz * (vb + 4) + (NH + 1) * 4 + 3 * NH * NRp * (16 if pcode == 0 else 8)
    Hd = K.to_device(H)
    if args.shape:
        u, ch, *xm = (int(v) for v in args.shape.split(":"))
        if ch > 0:
            K.tune_set_chunked(Hd, NR, args.prec, unroll=u, chunk=ch, xcd_grouped=xm[0] if xm else 0)
        else:
            K.tune_set(Hd, NR, args.prec, unroll=u, grid=0)
        tuned = dict(K.tune_get(Hd, NR, args.prec), fixed=True)
    else:
        tuned = K.autotune(Hd, NR, args.prec)
    x, z = DeviceBlock(NH, NR, args.prec, ctx), DeviceBlock(NH, NR, args.prec, ctx)
    ms = ctypes.c_float(0.0)
    n = max(args.kernel_launches, 1)
    K._lib.check(ctx.lib.kpm_bench_step(ctx.handle, Hd.handle, x.handle, z.handle, 3, ctypes.byref(ms)),
                 "bench_step")
    K._lib.check(ctx.lib.kpm_bench_step(ctx.handle, Hd.handle, x.handle, z.handle, n, ctypes.byref(ms)),
                 "bench_step")
    t = ms.value / n * 1e-3
    gbs = byts / t / 1e9
    print(f"[bench] spmm {args.op} L={L} NR={NR} {args.prec}: {t * 1e6:.1f} us per pass, "
          f"{gbs:.0f} GB/s algorithmic, tuned {tuned}", file=sys.stderr, flush=True)
    res = {"metric": "recurrence pass time (k_step MODE_DBL)", "value": round(t * 1e6, 2), "unit": "us",
           "n_gpus": 1, "steps": n, "warmup": 3, "ms_per_step": round(t * 1e3, 5),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "c128" if pcode == 0 else "c32-storage/f64-accumulate", "data": "synthetic",
           "config": {"workload": f"{args.op} L={L} spmm", "NH": NH, "nnz": nnz, "NR": NR,
                      "real_H": bool(real)},
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                        "bytes_per_launch": int(byts), "avg_launch_ms": round(t * 1e3, 5)},
           "tuned": tuned}
    print(json.dumps(res))
    return 0


def bench_core(args) -> int:
    """The reference's benchmark suite (benchmarks/core_bench.jl:127-175) on
    the same workloads: periodic square lattices (hopping -1, no diagonal)
    rescaled by normalizeH, unit random-phase probes.  Each case is timed as
    the reference times it -- the whole call with host-resident H and probes
    in, host moments out -- and reported as the minimum of --steps samples
    after --warmup untimed calls.  chebyshev_iter_single is timed on
    device-resident blocks (one fused pass + synchronisation)."""
    import numpy as np

    import kpmhip as K
    from kpmhip import models
    from kpmhip.rng import Xoshiro256pp

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    ctx = K.get_context()
    nsamp = max(args.steps, 1)

    def timed(fn, sync=False):
        for _ in range(max(args.warmup, 1)):
            fn()
        if sync:
            ctx.synchronize()
        best = float("inf")
        for _ in range(nsamp):
            t0 = time.perf_counter()
            fn()
            if sync:
                ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    def norm_square(L, currents=False):
        lat = models.square_lattice(L, L, currents=currents)
        H = lat[0] if currents else lat
        a, Hn = K.normalizeH(H)
        return (Hn.tocsc(), lat[1], lat[2]) if currents else Hn.tocsc()

    times = {}
    H1 = norm_square(256)
    psi1 = K.random_phase_vectors(Xoshiro256pp(1234), 1 << 16, 8)
    times["kpm_1d_dos_NH65536_NC1024_NR8"] = timed(lambda: K.kpm_1d(H1, 1024, 8, 1 << 16, psi_in=psi1))
    print("[bench] core kpm_1d 256^2 done", file=sys.stderr, flush=True)
    H2 = norm_square(512)
    psi2 = K.random_phase_vectors(Xoshiro256pp(1234), 1 << 18, 4)
    times["kpm_1d_dos_NH262144_NC512_NR4"] = timed(lambda: K.kpm_1d(H2, 512, 4, 1 << 18, psi_in=psi2))
    Hc, Jx, Jy = norm_square(128, currents=True)
    psic = K.random_phase_vectors(Xoshiro256pp(1234), 1 << 14, 4)
    for suffix, kw in (("default", {}), ("arr_size3", {"arr_size": 3}), ("arr_size64", {"arr_size": 64})):
        times[f"kpm_2d_cond_NH16384_NC64_NR4_{suffix}"] = timed(
            lambda kw=kw: K.kpm_2d(Hc, Jx, Jy, 64, 4, 1 << 14, psi_in=psic, **kw))
    print("[bench] core kpm_2d done", file=sys.stderr, flush=True)
    Ha = H1
    Va = K.random_phase_vectors(Xoshiro256pp(1234), 1 << 16, 8)
    C = np.stack([K.fermi_coefficients(1.0, 0.0, 0.0, NC=1024),
                  K.fermi_coefficients(1.0, 0.0, 0.25, NC=1024)], axis=1)
    out = np.zeros((1 << 16, 8, 2), dtype=np.complex128, order="F")
    times["chebyshev_action_NH65536_NC1024_NR8_K2"] = timed(
        lambda: K.chebyshev_action_(out, Ha, Va, C, check_every=0))
    Hi = K.to_device(norm_square(512))
    Vpp = K.DeviceBlock.from_host(K.random_phase_vectors(Xoshiro256pp(1234), 1 << 18, 8))
    Vp = K.DeviceBlock.from_host(K.random_phase_vectors(Xoshiro256pp(99), 1 << 18, 8))
    times["chebyshev_iter_single_NH262144_NR8"] = timed(
        lambda: K.chebyshev_iter_single(Hi, Vpp, Vp), sync=True)
    # the CPU restatement of the same two kpm_1d calls on 8 host threads, the
    # thread count of the reference's published CPU runs
    # (perf-core-mike-cpu.json:10-11), as a fidelity check of the CPU baseline
    port8 = {}
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.set_threads(8)
        for name, (H, NC, psi) in (("kpm_1d_dos_NH65536_NC1024_NR8", (H1, 1024, psi1)),
                                   ("kpm_1d_dos_NH262144_NC512_NR4", (H2, 512, psi2))):
            best = float("inf")
            for _ in range(2):
                t0 = time.perf_counter()
                oracle.kpm_1d(H, NC, psi)
                best = min(best, time.perf_counter() - t0)
            port8[name] = round(best, 4)
            print(f"[bench] CPU port (8 threads) {name}: {best:.3f} s", file=sys.stderr, flush=True)
    cases = {}
    for name, t in times.items():
        a100, cpu = CORE_PUBLISHED[name]
        cases[name] = {"seconds": round(t, 6), "a100_seconds": a100, "cpu8_seconds": cpu,
                       "speedup_vs_a100": round(a100 / t, 2), "speedup_vs_cpu8": round(cpu / t, 2)}
        if name in port8:
            cases[name]["port_cpu8_seconds"] = port8[name]
        print(f"[bench] {name}: {t * 1e3:.3f} ms  (A100 {a100 * 1e3:.2f} ms, "
              f"x{a100 / t:.1f})", file=sys.stderr, flush=True)
    head = "kpm_1d_dos_NH65536_NC1024_NR8"
    res = {"metric": "seconds per kpm_1d call (KPM.jl core_bench kpm_1d_dos_NH65536_NC1024_NR8)",
           "value": round(times[head], 6), "unit": "s", "n_gpus": 1, "steps": nsamp,
           "warmup": max(args.warmup, 1), "ms_per_step": round(times[head] * 1e3, 3),
           "higher_is_better": False, "scaling": "weak",
           "vs_baseline": round(times[head] / CORE_PUBLISHED[head][0], 4),
           "dtype": "c128", "data": "synthetic",
           "config": {"workload": "KPM.jl benchmarks/core_bench.jl suite (periodic square lattices)",
                      "baseline": "perf-core-mike-a100.json (A100, Julia 1.12.6)"},
           "cases": cases,
           "cpu_port": {"threads": 8, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
                        "note": "oracle C restatement of kpm_1d! on the same inputs; "
                                "port_cpu8_seconds beside the reference's cpu8_seconds"}}
    print(json.dumps(res))
    return 0


def bench_adjoint(args) -> int:
    """Gradient of a DOS loss: periodic 512 x 512 square lattice, eps_i ~ U[-1, 1],
    t_ij ~ 1 + U[-0.1, 0.1] (all trainable), fixed a = 5.5 (Gershgorin bound
    4 * 1.1 + 1 = 5.4), NC = 1024, NR = 16, Jackson kernel; loss = sum over 256
    energies of (rho(E) - rho*(E))^2 with rho* the DOS of a second disorder
    seed.  One step = one forward pass that keeps every vector in HBM (69 GB
    here) and dots each order with the probes (the moments), the loss, the
    reverse Chebyshev gradient on the kept vectors, and the per-site /
    per-bond gradients, host H in.  Algorithmic bytes (FP64, per step):
    NC-1 forward passes (x and z read, out written; the moment dots of the
    first NC/2 use the same vectors), NC-1 reverse passes (x, z and the source
    read, out written) and the gradient terms: beta and alpha read once each
    per reverse step, the nnz gradient read and written once per 8-step
    window."""
    import numpy as np

    import kpmhip as K
    from kpmhip import models
    from kpmhip.rng import Xoshiro256pp

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    L, NC, NR, a = 512, 1024, 16, 5.5
    H = models.square_disordered(L, seed=42)
    Hn = (H / a).tocsc()
    Ht = (models.square_disordered(L, seed=43) / a).tocsc()
    NH, nnz = H.shape[0], H.nnz
    psi = K.random_phase_vectors(Xoshiro256pp(42), NH, NR)
    E = np.linspace(-0.9 * a, 0.9 * a, 256)
    mu_t = K.kpm_1d(Ht, NC, NR, psi_in=psi)
    x = E / a
    ns = np.arange(NC)
    w = K.JacksonKernel(ns, NC) * K.hn(ns)
    T = np.cos(ns[None, :] * np.arccos(x)[:, None])
    rho_t = (T * w[None, :] / (a * np.pi * np.sqrt(1 - x * x))[:, None]) @ mu_t

    Hd = K.to_device(Hn)  # the operator stays on the device across steps
    # launch shapes of the recurrence passes for this operator and probe width
    # (the forward, reverse and doubling passes all read the tuner's table)
    tuned = K.autotune(Hd, NR, "f64")

    def step():
        tape = K.kpm_1d_adjoint_forward(Hn, NC, psi, operator=Hd)
        loss, g = K.dos_loss(tape.mu, a, 0.0, E, rho_t)
        grad = tape.pullback(g)
        return loss, K.site_gradient(grad, a), K.bond_gradient(H, grad, a)

    for _ in range(max(args.warmup, 1)):
        step()
    ctx = K.get_context()
    ctx.synchronize()
    tf = time.perf_counter()
    for _ in range(args.steps):
        K.kpm_1d(Hn, NC, NR, psi_in=psi)
    t_fwd = (time.perf_counter() - tf) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, gs, gb = step()
    wall = (time.perf_counter() - t0) / args.steps
    nrp = K.moment._padded_nr(NR, 0)
    vec = NH * nrp * 16
    hb = nnz * (8 + 4) + (NH + 1) * 4
    windows = -(-(NC - 1) // 8)
    total = ((NC - 1) * (hb + 3 * vec) + (NC - 1) * (hb + 4 * vec) + (NC - 1) * 2 * vec
             + windows * (nnz * (4 + 32) + (NH + 1) * 4))
    out = {"metric": "DOS-loss gradient step time (forward + reverse Chebyshev)",
           "value": round(wall * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
           "warmup": max(args.warmup, 1), "ms_per_step": round(wall * 1e3, 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
           "data": "synthetic",
           "config": {"workload": "square 512x512, eps~U[-1,1], t~1+U[-0.1,0.1], DOS L2 loss at 256 E",
                      "NH": NH, "nnz": nnz, "NC": NC, "NR": NR, "a": a},
           "forward_ms": round(t_fwd * 1e3, 3), "gradient_over_forward": round(wall / t_fwd, 2),
           "tuned": tuned,
           "algorithmic_bytes": total, "hbm_gbs": round(total / wall / 1e9, 1),
           "hbm_frac": round(total / wall / 1e9 / HBM_PEAK_GBS, 4), "loss": float(loss),
           "grad_site_norm": float(np.linalg.norm(gs)), "grad_bond_norm": float(np.linalg.norm(gb.data))}
    print(json.dumps(out))
    return 0


def bench_conductivity(args) -> int:
    """Kubo-Bastin sigma_xy of the Haldane model (honeycomb 256 x 256 cells,
    NH = 131,072, t1 = 1, t2 = 0.1 e^{i pi/2}, M = 0, periodic): kpm_2d with
    NC = 512, NR = 16 (all 512 left and right slots resident, one MFMA
    contraction), FP64 vectors with the FP64 and FP32 contraction modes.
    One step = one kpm_2d call with H, Jx, Jy already resident on the device
    (probes uploaded inside the call, the table copied back); the quantised
    sigma_xy in the gap comes from the device Kubo-Bastin reconstruction.
    The roofline object is the contraction kernel alone (MFMA-bound), timed
    with HIP events over back-to-back launches of the same shape."""
    import ctypes

    import numpy as np

    import kpmhip as K
    from kpmhip import models
    from kpmhip.rng import Xoshiro256pp

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    ctx = K.get_context()
    Lc, NC, NR = args.cells if args.cells != 1024 else 256, 512, 16
    H, Jx, Jy, area = models.haldane_model(Lc, Lc, t=1.0, t2=0.1)
    NH = H.shape[0]
    a, b = 3.8, 0.0  # fixed rescaling above the Gershgorin bound 3 t1 + 6 t2 = 3.6
    Hn = (H / a).tocsc()
    psi = K.random_phase_vectors(Xoshiro256pp(42), NH, NR)
    Hd, Jxd, Jyd = K.to_device(Hn), K.to_device(Jx), K.to_device(Jy)
    tuned = K.autotune(Hd, NR, "f64")
    print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    out = {}
    for gemm in ("f64", "f32", "bf16x3"):
        K.kpm_2d(Hd, Jxd, Jyd, NC, NR, psi_in=psi, gemm=gemm)  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            mu = K.kpm_2d(Hd, Jxd, Jyd, NC, NR, psi_in=psi, gemm=gemm)
        wall = (time.perf_counter() - t0) / args.steps
        tr = time.perf_counter()
        sig = K.kubo_bastin_cond_raw(mu, a, 0.0, b=b, NH=NH, area=area)
        t_rec = time.perf_counter() - tr
        out[gemm] = {"ms_per_step": round(wall * 1e3, 2), "sigma_xy": sig,
                     "kubo_bastin_ms": round(t_rec * 1e3, 2)}
        print(f"[bench] kpm_2d gemm={gemm}: {wall * 1e3:.1f} ms, sigma_xy={sig:.6f} "
              f"(reconstruction {t_rec * 1e3:.1f} ms)", file=sys.stderr, flush=True)
    flops = 8.0 * NC * NC * NH * NR  # 4 real products per complex multiply-add
    Kc = NH * K.moment._padded_nr(NR, 0)
    rl = {}
    # the contraction kernels on slots of pseudo-random values, as kpm_2d
    # stores them: FP64 slots for the FP64 and FP32 modes, split-BF16 planes
    # (vec_prec 2) for the split-BF16 mode (k_contract_planes_t128)
    for gemm, code, peak in (("f64", 0, 78.6), ("f32", 1, 157.3), ("bf16x3", 2, 2500.0 / 3)):
        ms = ctypes.c_float(0.0)
        vp = 2 if code == 2 else 0
        K._lib.check(ctx.lib.kpm_bench_contract(ctx.handle, NC, NC, Kc, 0, vp, code, 3, ctypes.byref(ms)),
                     "bench_contract")
        t = ms.value / 3 * 1e-3
        # executed MFMA flops: the FP64 mode runs the 3M complex product
        # (3 real products per complex multiply-add), the others the 4M form
        executed = flops * 0.75 if code == 0 else flops
        rl[gemm] = {"bound": "mfma", "achieved": round(executed / t / 1e12, 2), "peak": peak,
                    "unit": "TFLOP/s", "frac": round(executed / t / 1e12 / peak, 4), "traffic": None,
                    "kernel": ("k_contract_planes_t128" if code == 2 else
                               f"k_contract<double,{'double' if code == 0 else 'float'}>"),
                    "complex_gemm_4m_equivalent_tflops": round(flops / t / 1e12, 2),
                    "avg_launch_ms": round(t * 1e3, 3)}
        print(f"[bench] contraction {gemm}: {t * 1e3:.2f} ms, {executed / t / 1e12:.1f} TF/s executed "
              f"({flops / t / 1e12:.1f} TF/s as a 4M complex GEMM)", file=sys.stderr, flush=True)
    res = {"metric": "kpm_2d moment-table time (Haldane sigma_xy)", "value": out["f64"]["ms_per_step"],
           "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": out["f64"]["ms_per_step"], "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128", "data": "synthetic",
           "config": {"workload": f"Haldane {Lc}x{Lc} kpm_2d", "NH": NH, "NC": NC, "NR": NR,
                      "a": a, "contraction_flops": flops},
           "roofline": rl["f64"], "roofline_fp32_gemm": rl["f32"], "roofline_bf16x3_gemm": rl["bf16x3"],
           "fp64_gemm": out["f64"], "fp32_gemm": out["f32"], "bf16x3_gemm": out["bf16x3"],
           "sigma_xy_fp64_minus_fp32": out["f64"]["sigma_xy"] - out["f32"]["sigma_xy"],
           "sigma_xy_fp64_minus_bf16x3": out["f64"]["sigma_xy"] - out["bf16x3"]["sigma_xy"]}
    print(json.dumps(res))
    return 0


def bench_cpge(args) -> int:
    """Three-current (CPGE) response of the Haldane model (honeycomb 64 x 64
    cells, NH = 8192, t1 = 1, t2 = 0.1 e^{i pi/2}, M = 0, periodic).

    The kpm_3d table mu[n,m,p] = Tr[Jx T_m Jy T_n Jy T_p]/D is computed on the
    device at NC = 128, NR = 8. Then cpge runs with omega1 = 0.2,
    omega2 = 0.1, E_f = 0, beta = inf, lambda = 0.01 (rescaled units).

    One step is one cpge call on the host table: kernel weighting, table
    upload, and the adaptive quadrature, with every refinement round's nodes
    contracted in one FP64-MFMA launch. The roofline object is the node
    contraction (k_trilinear) at the mean batch of those rounds, timed with
    HIP events over back-to-back launches. The CPU baseline times the
    oracle's numpy restatement of the same contraction on a bounded sample
    of nodes and scales it to the call's node count."""
    import ctypes

    import numpy as np

    import kpmhip as K
    from kpmhip import models
    from kpmhip.cpge import Table3, cpge_node_coefficients
    from kpmhip.rng import Xoshiro256pp

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    ctx = K.get_context()
    Lc, NC, NR = 64, 128, 8
    H, Jx, Jy, area = models.haldane_model(Lc, Lc, t=1.0, t2=0.1)
    NH = H.shape[0]
    a = 3.8
    Hn = (H / a).tocsc()
    psi = K.random_phase_vectors(Xoshiro256pp(42), NH, NR)
    # first call (cold: operator upload, workspace allocation, code loading),
    # then a second call on the same shapes (warm)
    t0 = time.perf_counter()
    mu3 = K.kpm_3d(Hn, Jx, Jy, Jy, NC, NR, psi_in=psi)
    t3d_cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    mu3 = K.kpm_3d(Hn, Jx, Jy, Jy, NC, NR, psi_in=psi)
    t3d = time.perf_counter() - t0
    print(f"[bench] kpm_3d NC={NC} NR={NR} NH={NH}: {t3d:.3f} s warm, {t3d_cold:.3f} s cold",
          file=sys.stderr, flush=True)
    w1, w2, lam = 0.2, 0.1, 0.01
    counts = []
    orig = Table3.forms_rec

    def counting(self, A, B, Cc):
        counts.append(A.shape[0])
        return orig(self, A, B, Cc)

    Table3.forms_rec = counting
    try:
        for _ in range(args.warmup):
            K.cpge(mu3, NC, w1, w2, E_f=0.0, lambda_=lam)
        counts.clear()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            val = K.cpge(mu3, NC, w1, w2, E_f=0.0, lambda_=lam)
        wall = (time.perf_counter() - t0) / args.steps
    finally:
        Table3.forms_rec = orig
    forms = sum(counts) // args.steps
    launches = len(counts) // args.steps
    nodes = forms // 3
    Qm = max(3, int(round(forms / max(launches, 1))))
    print(f"[bench] cpge: {wall * 1e3:.1f} ms per call, {nodes} nodes in {launches} launches "
          f"(mean batch {Qm} forms), value {val}", file=sys.stderr, flush=True)
    reps = 20
    kv = K.JacksonKernel(np.arange(NC), NC) * K.hn(np.arange(NC))
    with Table3(mu3, NC, kvec=kv) as tb:
        ms = ctypes.c_float(0.0)
        K._lib.check(ctx.lib.kpm_bench_trilinear(tb.handle, Qm, reps, ctypes.byref(ms)), "bench_trilinear")
    tk = ms.value / reps * 1e-3
    flops = 8.0 * Qm * NC ** 3
    roof = {"bound": "mfma", "achieved": round(flops / tk / 1e12, 2), "peak": 78.6, "unit": "TFLOP/s",
            "frac": round(flops / tk / 1e12 / 78.6, 4), "traffic": None,
            "kernel": "k_trilinear (+ k_trilinear_reduce)", "forms_per_launch": Qm,
            "avg_launch_ms": round(tk * 1e3, 4)}
    print(f"[bench] node contraction Q={Qm}: {tk * 1e3:.3f} ms, {flops / tk / 1e12:.1f} TF/s",
          file=sys.stderr, flush=True)
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from threadpoolctl import threadpool_info
        mt = K.mu3D_apply_kernel_and_h(mu3, NC, K.JacksonKernel)
        sample = 16
        th = np.linspace(0.3, 2.8, sample)
        A_, B_, C_ = cpge_node_coefficients(th, NC, w1, w2, lam, 0.0)
        O.cpge_forms(mt, A_[:3], B_[:3], C_[:3])  # warm the BLAS thread pool
        t0 = time.perf_counter()
        O.cpge_forms(mt, A_, B_, C_)
        tc = (time.perf_counter() - t0) / sample
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
        cpu = {"value": round(tc * nodes * 1e3, 1), "unit": "ms per cpge call (node contractions only)",
               "cores": threads, "kind": "port",
               "sample": f"{sample} nodes of the oracle's numpy _cpge_contract! restatement "
                         f"({tc * 1e3:.2f} ms/node), scaled to {nodes} nodes"}
    res = {"metric": "cpge call time (Haldane kpm_3d table)", "value": round(wall * 1e3, 2), "unit": "ms",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 2),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
           "data": "synthetic",
           "config": {"workload": f"Haldane {Lc}x{Lc} cpge from kpm_3d", "NH": NH, "NC": NC, "NR": NR,
                      "omega1": w1, "omega2": w2, "lambda": lam, "nodes": nodes, "launches": launches},
           "kpm_3d_s": round(t3d, 3), "kpm_3d_cold_s": round(t3d_cold, 3), "nodes_per_s": round(nodes / wall, 1), "cpge": [val.real, val.imag],
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(res))
    return 0


def bench_cond_ensemble(args) -> int:
    """One task of the Anderson-Seebeck conductivity ensemble
    (scripts/anderson_seebeck_ensemble.jl with the reference manifest's
    values: W = 12, a = 12.25, NR = 8, L = 16, NC = 1024). A task is the
    antithetic pair H(+eps), H(-eps) with Gamma-paired probes: two device
    kpm_2d tables and the task record written to a scratch directory. One
    step = one task (a fresh seed each step, so no record is skipped). The
    CPU baseline times the oracle's C kpm_2d on the same task at NC = 256 (a
    bounded sample; the contraction scales as NC^2, the recurrences as NC)."""
    import shutil
    import tempfile

    import numpy as np

    import kpmhip as K
    from kpmhip import ensemble as E
    from kpmhip.frontend import random_phase_vectors
    from kpmhip.rng import Xoshiro256pp

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    mf = E.validate_manifest({"model": {"W": 12.0, "t": 1.0, "a": 12.25, "b": 0.0},
                              "kpm": {"NR": 8, "arr_size": 16},
                              "ensemble": {"seed_base": 99532492, "sizes": [16], "NC": [1024],
                                           "pairs_per_size": 64}})
    tmp = tempfile.mkdtemp(prefix="kpm_ens_")
    try:
        quiet = lambda *a: None  # noqa: E731
        for t in range(1, args.warmup + 1):
            E.run_task(mf, t, tmp, log=quiet)
        t0 = time.perf_counter()
        for t in range(args.warmup + 1, args.warmup + args.steps + 1):
            path = E.run_task(mf, t, tmp, log=quiet)
        wall = (time.perf_counter() - t0) / args.steps
        rec = E.load_task(path)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    L, NC, NR = 16, 1024, 8
    NH = L ** 3
    print(f"[bench] cond ensemble task (L={L}, NC={NC}, NR={NR}): {wall:.3f} s, pair residual "
          f"{rec['meta']['pair_residual']:.2e}", file=sys.stderr, flush=True)
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        _, _, _, _, seed = E.task_point(mf, 1)
        rng = Xoshiro256pp(seed)
        eps = 12.0 * (rng.rand(NH) - 0.5)
        psi = random_phase_vectors(rng, NH, NR)
        Hp, Jx, _ = E.cubic_hamiltonian_and_current(L, eps)
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        NCs = 256
        t0 = time.perf_counter()
        oracle.kpm_2d((Hp / 12.25).tocsc(), Jx, Jx, NCs, psi, psi, NCs, NCs)
        tc = time.perf_counter() - t0
        cpu = {"value": round(2 * tc, 3), "unit": f"s per task at NC={NCs} (two tables)", "cores": threads,
               "kind": "port", "sample": f"one oracle kpm_2d (C, {threads} threads) at NC={NCs}, doubled "
                                         f"for the pair; the GPU task runs NC={NC}"}
    res = {"metric": "conductivity-ensemble task time (Anderson L=16 antithetic pair)",
           "value": round(wall, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(wall * 1e3, 2), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128", "data": "synthetic",
           "config": {"workload": "Anderson L=16 cubic, W=12, a=12.25, NC=1024, NR=8, two kpm_2d tables",
                      "NH": NH, "NC": NC, "NR": NR},
           "pair_residual": rec["meta"]["pair_residual"], "cpu_baseline": cpu}
    print(json.dumps(res))
    return 0


def bench_ensemble(args) -> int:
    """Disorder ensemble: nreal cubic Anderson lattices, L uniform in
    {16, 24, 32, 40, 48}, W uniform in [2, 20], eps_i ~ U[-W/2, W/2], each
    rescaled by its Gershgorin bound 6t + W/2 (+0.1 margin); NC = 1024,
    NR = 4 per realisation, ComplexF32 recurrence; the NC/2 steps of all
    realisations run as one fused launch per step, replayed from a HIP graph.
    One step = one batched call."""
    import numpy as np

    import kpmhip as K
    from kpmhip import models
    from kpmhip.rng import Xoshiro256pp

    if not K.activate(0):
        raise SystemExit("no HIP device visible")
    NC, NR = 1024, 4
    rng = Xoshiro256pp(2025, lanes=4)
    Ls = [(16, 24, 32, 40, 48)[int(5 * rng.rand())] for _ in range(args.nreal)]
    Ws = [2.0 + 18.0 * rng.rand() for _ in range(args.nreal)]
    print(f"[bench] building {args.nreal} cubic realisations ...", file=sys.stderr, flush=True)
    Hs = []
    for i, (L, W) in enumerate(zip(Ls, Ws)):
        H, *_ = models.cubic_model(L, W=W, seed=1000 + i)
        Hs.append((H / (6.0 + W / 2 + 0.1)).tocsc())
    nnz = sum(H.nnz for H in Hs)
    rows = sum(H.shape[0] for H in Hs)
    psis = [K.random_phase_vectors(Xoshiro256pp(5000 + i), H.shape[0], NR) for i, H in enumerate(Hs)]
    B = K.BatchOperator(Hs)
    B.upload(psis, NR, "f32")
    tuned = B.autotune()
    print(f"[bench] rows={rows} nnz={nnz} {B.info()} tuned {tuned}", file=sys.stderr, flush=True)
    for _ in range(max(args.warmup, 1)):
        B.run(NC, use_graph=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mu = B.run(NC, use_graph=True)
    wall = (time.perf_counter() - t0) / args.steps
    units = NC * NR * nnz
    out = {"metric": "moments/s (sum_r NC*NR*nnz_r / s), batched FP32 kpm_1d ensemble",
           "value": units / wall, "unit": "moments/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "c32-storage/f64-accumulate",
           "data": "synthetic",
           "config": {"workload": f"{args.nreal} cubic Anderson realisations, L in 16..48",
                      "rows": rows, "nnz": nnz, "NC": NC, "NR": NR, "graph": True},
           "tuned": tuned, "mu_finite": bool(np.all(np.isfinite(mu)))}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
