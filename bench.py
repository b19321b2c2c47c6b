#!/usr/bin/env python3
"""Benchmark of one time step of the Fourier-Chebyshev channel DNS on one MI355X.

Workload (the headline configuration): pressure-driven channel 128x129x128 (mx = my =
192, mz = 512: padded 192 x 257 x 192), Lx = 4 pi, Lz = 4 pi/3, Re = 3300, constant flow
rate, dt = 0.002, U = 1 - y^2, seeded wall-satisfying perturbation (amplitude 0.1,
|k| <= 16, degree <= 24, (1+|k|)^(-5/6), seed 0).  A step is nonlinear term +
Crank-Nicolson/AB2 solves + velocity recovery, all resident in HBM.

Prints one JSON line: ns per grid point per step (Nx*Ny*Nz = 2,113,536 points), the
roofline of the dominant kernel from HIP events recorded on the library's stream around
each of its launches in the timed steps, a per-kernel table from a separate untimed pass
with events around every launch, and the CPU baseline (the FP64 oracle, OpenMP, on this
host) on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-spectral-channel-dns_amd"))

import numpy as np  # noqa: E402

import rpcf  # noqa: E402

HBM_PEAK_GBS = 8000.0

C1 = dict(mx=192, my=192, mz=512, n_pi_x=4.0, n_pi_y=4.0 / 3.0, Re=3300.0, Ro=0.0, dt=0.002,
          u_profile=(1.0, 0.0, -1.0), flow_control=1)
C1_IC = (0.1, 16, 24, 5.0 / 6.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def kernel_bytes(s, prm):
    """algorithmic HBM bytes per launch of each kernel: the arrays it must read and write
    once.  FP64 form: complex FP64 (16 B) throughout; with an even base profile (the channel)
    the solves read the two parity blocks of each inverse (nz x (N/2 + 1) complex FP64
    entries) and the remainders of the v blocks (the same count of complex BF16 pairs, 4 B),
    otherwise the full inverses (nz x nz).  Mixed form (precision 1; channel profile, matrix-core
    Chebyshev transforms): the velocity and vorticity fields, the padded intermediates and
    the Lamb vector in complex FP32 (8 B), the state, nonlinear terms and operator outputs in
    FP64, and the solves with the parity blocks alone, the eta blocks as FP32 high and BF16
    low parts.  Precision 2: FP64 with the parity blocks alone.  (The mean-flow slot reads its two full inverses in every form: 0.2 MB
    more at 128x129x128, not counted.)"""
    P, nz, hnx, ny, pz = s["P"], s["nz"], s["hnx"], s["ny"], s["pz"]
    my, nxh = prm.my, s["nx"] // 2
    parity = prm.u_profile[1] == 0.0 and os.environ.get("RPCF_FULL_INVERSE", "0") == "0"
    mixed = prm.precision == 1 and 2 * (nz - 1) <= 384 and os.environ.get("RPCF_MIXED_MFMA", "1") != "0"
    h = 0.5 if prm.precision == 1 else 1.0  # FP32 intermediates of the mixed form
    S = P * nz * 16                    # one spectral field
    F = S * (0.5 if mixed else 1.0)    # one velocity / vorticity field (FP32 for the matrix-core form)
    Sp = hnx * ny * pz * 16 * h        # one field on the padded wall-normal points
    Wb = pz * my * nxh * 16 * h        # one field after the span transform (kx < nx/2)
    Zb = pz * my * hnx * 16 * h        # one Lamb component after the x transform
    np_ = (nz - 1) // 2 + 1            # parity pair columns
    M2 = 2 * nz * nz * 16              # a slot's two inverses
    if parity:  # their parity blocks: nz rows x (N/2 + 1) pair columns, FP64, and the v remainders;
        # mixed form: the eta blocks as FP32 high and BF16 low parts (12 B)
        M2 = {0: 2 * nz * np_ * 16 + nz * np_ * 4, 1: nz * np_ * (16 + 12)}.get(prm.precision, 2 * nz * np_ * 16)
    nslot = hnx * (ny // 2 + 1)        # matrix slots: pencils +-ky share one
    out = {
        "k_cheb_s2p": 6 * F + 6 * Sp,
        "k_span_inv": 6 * Sp + 6 * Wb,
        "k_lamb_phys": 6 * Wb + 3 * Zb,
        "k_span_fwd": 3 * Zb + 3 * Sp,
        "k_p2s_nl": 3 * Sp + 2 * S,
        "k_lin_tv": 2 * S + 5 * S,
        "k_lin_solve": nslot * M2 + 11 * S,
        "k_velocity": 2 * S + 6 * F,
    }
    return {k: int(v) for k, v in out.items()}


def read_traffic(kernel, precision=0):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE,
    committed in profiles/traffic_<kernel>_c1.txt (FP64) or traffic_<kernel>_c1_mixed.txt
    (line 'traffic_bytes_per_launch: N')."""
    tag = {0: "", 1: "_mixed"}.get(precision)
    if tag is None:
        return None
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_c1{tag}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prm, v0, e0, steps, ref_state):
    """the FP64 oracle (oracle/rpcf_oracle.c) on this host's cores: create (matrix
    build, not timed), bootstrap, then `steps` timed steps, each as its three parts
    (nonlinear term, linear step = CN/AB2 right-hand sides + per-mode solves + mean modes,
    velocity recovery) so that the linear part is reported on its own"""
    dll = rpcf.load(rpcf.ORACLE_LIB)
    dll.rpcf_oracle_threads.restype = ctypes.c_int
    threads = int(dll.rpcf_oracle_threads())
    t0 = time.perf_counter()
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as o:
        t_init = time.perf_counter() - t0
        o.set_state(v0, e0)
        o.bootstrap()
        a = o.get_state()
        parity = max(float(np.abs(ref_state[i] - a[i]).max() / np.abs(a[i]).max()) for i in (0, 1))
        t_parts = [0.0, 0.0, 0.0]
        for _ in range(steps):
            for j, fn in enumerate((o.nonlinear, o.linear, o.velocity)):
                t1 = time.perf_counter()
                fn()
                t_parts[j] += time.perf_counter() - t1
    return threads, t_init, [t / steps for t in t_parts], parity


def copy_probe(nbytes, reps):
    """a plain device-to-device copy of known size (torch's copy kernel): the calibration
    run for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, which count what one dispatch of this
    copy reads and writes against the nbytes it must move"""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def kernel_table(dll, h):
    """{kernel name: (summed ms, launches)} of the HIP-event timers of a context"""
    dll.rpcf_timing_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_long)]
    dll.rpcf_kernel_name.restype = ctypes.c_char_p
    dll.rpcf_kernel_name.argtypes = [ctypes.c_int]
    out = {}
    for kid in range(dll.rpcf_num_kernels()):
        tot, cnt = ctypes.c_double(), ctypes.c_long()
        dll.rpcf_timing_get(h, kid, ctypes.byref(tot), ctypes.byref(cnt))
        if cnt.value:
            out[dll.rpcf_kernel_name(kid).decode()] = (tot.value, cnt.value)
    return out


def mixed_leg(prm64, v0, e0, warmup, steps, ms64, final64, precision=1, table_steps=10):
    """the same workload in another form -- precision 1, the mixed form (FP32 transform
    intermediates, x/span FFTs and Lamb vector; Chebyshev transforms as split-FP16 products
    on the matrix cores; FP64 velocity recovery, operators and parity-block solves), or 2
    (FP64 with the parity-block solves): time per step, the relative field error against the
    FP64 run after the same steps, and the per-kernel times of a separate pass with events
    around every launch (run after the timed steps, so the error compares equal step counts)"""
    prm = rpcf.make_params(**C1)
    prm.precision = precision
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        sim.set_state(v0, e0)
        sim.bootstrap()
        sim.step(warmup)
        device_sync()
        t0 = time.perf_counter()
        sim.step(steps)
        device_sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        v, e = sim.get_state()
        dll = sim.dll
        dll.rpcf_timing_kernels.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        assert dll.rpcf_timing_kernels(sim.h, 0xFFFFFFFF) == 0
        sim.step(table_steps)
        table = {k: round(t / c, 5) for k, (t, c) in kernel_table(dll, sim.h).items()}
        assert dll.rpcf_timing_kernels(sim.h, 0) == 0
        nbytes = kernel_bytes(rpcf.sizes(dll, prm), prm)
    frac = {k: round(nbytes[k] / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for k, t in table.items() if k in nbytes}
    err = {"v": float(np.abs(v - final64[0]).max() / np.abs(final64[0]).max()),
           "eta": float(np.abs(e - final64[1]).max() / np.abs(final64[1]).max())}
    log(f"precision {precision}: {ms:.4f} ms/step ({ms64 / ms:.2f}x FP64), error after {1 + warmup + steps} steps {err}")
    log(f"precision {precision} kernels (ms): {table}")
    return {"ms_per_step": round(ms, 4), "speedup_vs_fp64": round(ms64 / ms, 3), "steps_compared": 1 + warmup + steps,
            "max_rel_error": err, "kernel_avg_ms": table, "kernel_hbm_frac": frac,
            "dtype": "f32 transforms (split-f16 MFMA Chebyshev) / f64 solves" if precision == 1
            else "f64, parity-block solves"}


def ensemble_leg(n, nsteps):
    """n independent simulations of mixed grid sizes in the mixed form (FP32 transforms,
    FP64 solves): one group per shape, the groups concurrently, against the same members
    run one at a time"""
    from rpcf import ensemble
    ms = ensemble.mixed_members(n, seed=0, precision=1)
    wall, runs = ensemble.run_grouped(ms, nsteps)
    alone_total, alone_each, _ = ensemble.run_alone(ms, nsteps)
    alone_g_total, alone_g_each, _ = ensemble.run_alone(ms, nsteps, graphs=True)
    per100_group = [r.seconds / nsteps * 100 for r in runs for _ in r.members]
    per100_alone = [t / nsteps * 100 for t in alone_each]
    gp = sum(ensemble.grid_points(m) for m in ms) * nsteps
    log(f"ensemble {n} members x {nsteps} steps: grouped {wall:.3f}s, one at a time {alone_total:.3f}s "
        f"(with graph replay {alone_g_total:.3f}s)")
    return {
        "members": n, "steps": nsteps, "dtype": "f32 transforms / f64 solves",
        "groups": {f"{r.shape[0]}x{r.shape[1]}x{r.shape[2]}": len(r.members) for r in runs},
        "group_s": {f"{r.shape[0]}x{r.shape[1]}x{r.shape[2]}": round(r.seconds, 4) for r in runs},
        "one_at_a_time_s": {f"{r.shape[0]}x{r.shape[1]}x{r.shape[2]}": round(sum(alone_each[i] for i in r.members), 4)
                            for r in runs},
        "steps_per_s": round(n * nsteps / wall, 1),
        "steps_per_s_one_at_a_time": round(n * nsteps / alone_total, 1),
        "speedup": round(alone_total / wall, 2),
        "one_at_a_time_graphs_s": round(alone_g_total, 4),
        "steps_per_s_one_at_a_time_graphs": round(n * nsteps / alone_g_total, 1),
        "speedup_vs_graphs_one_at_a_time": round(alone_g_total / wall, 2),
        "grid_point_steps_per_s": round(gp / wall, 1),
        "p50_s_per_100_steps": round(float(np.percentile(per100_group, 50)), 4),
        "p99_s_per_100_steps": round(float(np.percentile(per100_group, 99)), 4),
        "p99_s_per_100_steps_one_at_a_time": round(float(np.percentile(per100_alone, 99)), 4),
    }


def serving_leg(window, horizon, steps):
    """Poisson requests at 20 req/s over `window` seconds (grids 32x33x32 .. 256x129x256,
    `steps` = (min, max) step counts), continuous admission by free HBM, served until all
    finish or `horizon` seconds"""
    from rpcf import serving
    reqs = serving.poisson_requests(rate=20.0, duration=window, seed=0, steps=steps)
    srv = serving.Server()
    t0 = time.perf_counter()
    done = srv.run(reqs, horizon)
    wall = time.perf_counter() - t0
    out = serving.summarize(done, reqs, wall)
    out.update({"window_s": window, "horizon_s": horizon, "wall_s": round(wall, 2), "rate_per_s": 20.0,
                "steps_range": list(steps), "dtype": "f64",
                "groups": {f"{k[0]}x{k[1]}x{k[2]}": [len(g.slots) for g in v] for k, v in srv.groups.items()}})
    log(f"serving: {out['finished']}/{out['submitted']} finished in {wall:.1f}s")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed oracle steps for the CPU baseline")
    ap.add_argument("--table-steps", type=int, default=10,
                    help="steps of the untimed pass that times every kernel (the per-kernel table)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1, 2),
                    help="form of the headline leg (0: FP64, the metric; 1: mixed; 2: FP64 with parity-block "
                         "solves); with 1 or 2 the other legs are skipped (for profiling one form)")
    ap.add_argument("--ensemble", type=int, default=64, metavar="N",
                    help="members of the mixed-size ensemble leg (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--serve", type=float, default=10.0, metavar="SECONDS",
                    help="arrival window of the serving leg (Poisson, 20 req/s; 0: skip)")
    ap.add_argument("--serve-horizon", type=float, default=40.0, metavar="SECONDS",
                    help="wall-time limit of the serving leg")
    ap.add_argument("--serve-steps", type=int, nargs=2, default=(10, 200), metavar=("MIN", "MAX"),
                    help="step counts of the serving requests (uniform)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one simulation on one GPU; multi-GPU runs are not "
              "part of it (the pencil/plane slab decomposition is design only). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe, 3)
        return
    # the CPU baseline's OpenMP threads stay close to one another; their count is the
    # environment's OMP_NUM_THREADS, else this process's CPU share
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))

    prm = rpcf.make_params(**C1)
    prm.precision = args.precision
    if args.precision:
        args.no_mixed, args.ensemble, args.serve, args.no_cpu = True, 0, 0.0, True
    t0 = time.perf_counter()
    sim = rpcf.Simulation(rpcf.HIP_LIB, prm)
    t_create = time.perf_counter() - t0
    s = sim.s
    npts = s["nx"] * s["nz"] * s["ny"]
    v0, e0 = rpcf.seed_state(sim.dll, prm, 0, *C1_IC)
    sim.set_state(v0, e0)
    sim.bootstrap()
    state1 = sim.get_state()
    sim.step(args.warmup)
    log(f"create {t_create:.2f}s; grid {s}")

    dll = sim.dll
    dll.rpcf_timing_kernels.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    dll.rpcf_timing_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_long)]
    dll.rpcf_kernel_name.restype = ctypes.c_char_p
    dll.rpcf_kernel_name.argtypes = [ctypes.c_int]
    nbytes = kernel_bytes(s, prm)

    def kernel_times():
        out = {}
        for kid in range(dll.rpcf_num_kernels()):
            tot, cnt = ctypes.c_double(), ctypes.c_long()
            dll.rpcf_timing_get(sim.h, kid, ctypes.byref(tot), ctypes.byref(cnt))
            if cnt.value:
                out[dll.rpcf_kernel_name(kid).decode()] = (kid, tot.value, cnt.value)
        return out

    # per-kernel table: a separate pass with events around every kernel (they add ~60 us
    # per step, so the timed steps below bracket only the dominant kernel)
    assert dll.rpcf_timing_kernels(sim.h, 0xFFFFFFFF) == 0
    sim.step(args.table_steps)
    table = kernel_times()
    kernels = {}
    for name, (kid, tot, cnt) in table.items():
        avg = tot / cnt
        gbs = nbytes[name] / (avg * 1e-3) / 1e9
        kernels[name] = {"avg_ms": round(avg, 5), "launches": cnt, "bytes": nbytes[name],
                         "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "share": round(tot / sum(v[1] for v in table.values()), 4)}
    dom = max(table, key=lambda k: table[k][1])

    # the timed steps, with HIP events on the library's stream around the dominant kernel
    assert dll.rpcf_timing_kernels(sim.h, 1 << table[dom][0]) == 0
    device_sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    device_sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    _, dtot, dcnt = kernel_times()[dom]
    dgbs = nbytes[dom] / (dtot / dcnt * 1e-3) / 1e9
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(dgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dgbs / HBM_PEAK_GBS, 4), "traffic": read_traffic(dom, args.precision),
                "avg_ms": round(dtot / dcnt, 5), "launches": dcnt}
    assert dll.rpcf_timing_kernels(sim.h, 0) == 0
    final64 = sim.get_state()
    sim.close()
    # the mixed run takes the same steps as the FP64 run (bootstrap, warmup, table pass, timed)
    mixed = None if args.no_mixed else mixed_leg(prm, v0, e0, args.warmup + args.table_steps, args.steps, ms_step,
                                                 final64)
    fp64_parity = None if args.no_mixed else mixed_leg(prm, v0, e0, args.warmup + args.table_steps, args.steps,
                                                       ms_step, final64, precision=2)

    ens = ensemble_leg(args.ensemble, args.ensemble_steps) if args.ensemble else None
    srv = serving_leg(args.serve, args.serve_horizon, tuple(args.serve_steps)) if args.serve > 0 else None

    cpu = None
    parity = None
    if not args.no_cpu:
        threads, t_init, parts, parity = cpu_baseline(prm, v0, e0, args.cpu_steps, state1)
        t_cpu = sum(parts)
        cpu = {"value": round(t_cpu * 1e9 / npts, 2), "unit": "ns/gridpoint/step", "cores": threads,
               "kind": "port",
               "sample": f"FP64 oracle (OpenMP, {threads} threads, OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}) "
                         f"on the same 128x129x128 channel: matrix build {t_init:.1f}s untimed, bootstrap, then "
                         f"{args.cpu_steps} timed steps ({t_cpu * 1e3:.0f} ms/step)",
               "ms_per_step": round(t_cpu * 1e3, 2),
               "linear_ms_per_step": round(parts[1] * 1e3, 2),
               "nonlinear_ms_per_step": round(parts[0] * 1e3, 2),
               "velocity_ms_per_step": round(parts[2] * 1e3, 2),
               "cpu_model": cpu_model(), "nproc": os.cpu_count(), "cpu_share": len(os.sched_getaffinity(0))}
        log(f"cpu baseline {t_cpu * 1e3:.1f} ms/step (linear {parts[1] * 1e3:.1f}) on {threads} threads "
            f"of {cpu['cpu_model']}; GPU-vs-oracle parity after 1 step {parity:.2e}")

    out = {
        "metric": "ns/gridpoint/step at 128x129x128 (%s)" % ("FP64", "mixed", "FP64, parity-block solves")[args.precision],
        "value": round(ms_step * 1e6 / npts, 4),
        "unit": "ns/gridpoint/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != 1 else "f32 transforms (split-f16 MFMA Chebyshev) / f64 solves",
        "data": "synthetic (seeded wall-satisfying perturbation, seed 0)",
        "config": {"workload": "channel 128x129x128 %s, constant flow rate" % ("FP64", "mixed", "FP64 parity-block solves")[args.precision],
                   "grid": "128x129x128",
                   "padded": "192x257x192", "Re": 3300, "dt": 0.002, "Lx": "4pi", "Lz": "4pi/3",
                   "parallelism": "single GPU"},
        "steps_per_s": round(1e3 / ms_step, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_1step_vs_oracle": parity,
        "kernels": kernels,
        "mixed": mixed,
        "fp64_parity_solves": fp64_parity,
        "ensemble": ens,
        "serving": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
