#!/usr/bin/env python3
"""Benchmark of the nlchains hot path on one MI355X.

Workload (default): DNKG m=1, beta=1, N=1024 sites, 16384 copies, dt=0.05,
FP64 plain method, initial mode energies equal in the band k < N/8 at energy
density 0.5 (seed 2).  One bench step = one batch of --batch Yoshida-6 steps
(default 1000, the dump interval of this workload) over the whole ensemble
followed by the linear-mode energy pass and the entropies (the reference's
per-batch loop, DNKG_FPUT_Toda.cpp:108-143).

value  = site-steps/s over the timed region (whole job, inputs resident in HBM).
roofline: FP64 algorithmic FLOP of the step kernel (70 FLOP per DNKG
          site-step) / the kernel's mean duration from HIP events recorded on
          the context's stream, against the MI355X FP64 vector peak.
cpu_baseline: the oracle's release-flags build (OpenMP over copies) on the
          host cores, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "site-steps/s plain FP64 vs mixed/batched at stated energy drift; % of roofline"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# algorithmic FLOP per site-step (FMA = 2, exp/sincos = 0): 7 kicks + 7 drifts;
# FPUT: 17 FLOP per substep in the form DNKG_FPUT_Toda.cu:33-50 writes, 13 for its
# beta-only form when alpha = 0; DNLS: 7 x (two complex FFTs at 5 N log2 N / N
# each + 16 pointwise), N = 4096.
FLOP_PER_SITE_STEP = {"DNKG": 70, "dDNKG": 70, "FPUT": 119, "Toda": 56, "DNLS": 7 * (10 * 12 + 16)}


def flop_per_site_step(w) -> int:
    if w["model"] == "FPUT" and w["alpha"] == 0.0:
        return 91
    return FLOP_PER_SITE_STEP[w["model"]]

WORKLOADS = {
    # FPUT-beta, 64 sites x 1024 copies, dt=0.1, energy density 0.1 over k = 1..4 (seed 1),
    # energies + entropies every 100 steps: the batches are queued as the CLI queues them
    "c1": dict(model="FPUT", n=64, copies=1024, dt=0.1, m=0.0, alpha=0.0, beta=1.0, k_lo=1, k_hi=5, dens=0.1,
               seed=1, batch=100, pipelined=True),
    "dnkg": dict(model="DNKG", n=1024, copies=16384, dt=0.05, m=1.0, alpha=0.0, beta=1.0, k_lo=0, k_hi=128,
                 dens=0.5, seed=2, batch=1000),
    "fput": dict(model="FPUT", n=1024, copies=16384, dt=0.05, m=0.0, alpha=0.0, beta=1.0, k_lo=1, k_hi=128,
                 dens=0.5, seed=2, batch=1000),
    "toda": dict(model="Toda", n=1024, copies=16384, dt=0.05, m=0.0, alpha=1.0, beta=0.0, k_lo=1, k_hi=52,
                 dens=0.25, seed=4, batch=1000),
    # ragged Toda alpha=1: 65536 chains, lengths uniform over {32,...,2048}, harmonic energy
    # density U[0.05,0.25] in the lowest 5% of modes (at 0.5 the longest chains reach bond
    # stretches near 3 and leave the dt=0.05 stability region, in the oracle as on the GPU);
    # autotuned kernel, one hipGraph (move + energy pass + entropies) per batch
    "toda_ragged": dict(model="Toda", ragged=True, nchains=65536, choices=[32, 64, 128, 256, 512, 1024, 2048],
                        dt=0.05, m=0.0, alpha=1.0, beta=0.0, mode_fraction=0.05, e_lo=0.05, e_hi=0.25, seed=4,
                        batch=1000),
    # dDNKG beta=1, masses U[0.5,1.5] (seed 5), N=512, 8192 copies, dt=0.05, energy density 0.5 in
    # the lowest N/8 eigenmodes; energy pass = fused FP64 projection on the eigenbasis per batch
    "ddnkg": dict(model="dDNKG", n=512, copies=8192, dt=0.05, m=0.0, alpha=0.0, beta=1.0, m_lo=0.5, m_hi=1.5,
                  modes=64, dens=0.5, seed=5, batch=500),
    # DNLS beta=1, N=4096, 4096 copies, dt=0.01, norm density 1 on |k| < N/16
    "dnls": dict(model="DNLS", n=4096, copies=4096, dt=0.01, m=0.0, alpha=0.0, beta=1.0, kmax=256, dens=1.0,
                 seed=3, batch=1000),
}


class HipEvents:
    """hipEvent timing on an arbitrary stream (torch's events only see torch's stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        ev = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(ev)) == 0
        return ev

    def record(self, ev, stream):
        assert self.hip.hipEventRecord(ev, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)

    def destroy(self, ev):
        self.hip.hipEventDestroy(ev)


def host_cores() -> int:
    """Host cores this process may run on (its CPU affinity)."""
    return max(1, len(os.sched_getaffinity(0)))


def cpu_baseline(w, s0, seconds_budget: float):
    """Oracle release build, OpenMP over copies, on a bounded sample of the workload."""
    from oracle import oracle as O
    lib = O.fast()
    cores = host_cores()
    def move(state, steps):
        if w["model"] == "DNLS":
            return O.move_dnls(state, steps, w["dt"], w["beta"], threads=cores, lib=lib)
        model = {"DNKG": O.DNKG, "FPUT": O.FPUT, "Toda": O.TODA, "dDNKG": O.DDNKG}[w["model"]]
        return O.move_real(model, state, steps, w["dt"], m=w["m"], alpha=w["alpha"], beta=w["beta"],
                           mp2=w.get("mp2"), threads=cores, lib=lib)

    # calibrate on a small slice, then size the sample to the time budget
    probe_copies = min(w["copies"], 4 * cores if w["model"] == "DNLS" else 64 * cores)
    t = time.perf_counter()
    move(s0[:probe_copies], 4)
    rate = probe_copies * w["n"] * 4 / (time.perf_counter() - t)
    copies = w["copies"]
    steps = int(seconds_budget * rate / (copies * w["n"]))
    if steps < 10:  # keep at least 10 steps: shrink the copy count instead
        copies = max(cores, min(copies, int(seconds_budget * rate / (10 * w["n"]))))
        steps = 10
    steps = min(steps, 20 * w["batch"])  # the budget, not one batch, bounds the sample
    t = time.perf_counter()
    move(s0[:copies], steps)
    el = time.perf_counter() - t
    if el < 0.5 * seconds_budget:  # the small probe underestimates the rate: rescale once
        steps = max(steps, min(int(steps * seconds_budget / el), 10000 * w["batch"]))
        t = time.perf_counter()
        move(s0[:copies], steps)
        el = time.perf_counter() - t
    return {"value": copies * w["n"] * steps / el, "unit": "site-steps/s", "cores": cores, "host_cores": host_cores(),
            "kind": "port",
            "sample": f"oracle/liboracle_fast.so (-O3 -ffast-math, OpenMP over copies), {w['model']} "
                      f"N={w['n']} x {copies} copies x {steps} steps, {el:.1f} s"}


FP32C_METHOD = {
    "DNLS": "state on chip as FP32 pairs hi + lo (48 significant bits); every substep's increment (nonlinear "
            "phase, FFTs, Fourier multiplier) in FP32, added by a compensated sum; state in HBM, energies, norms "
            "and entropies in FP64",
    "FPUT": "q and p on chip as FP32 pairs hi + lo; bond forces and every drift and kick increment in FP32, added "
            "by Kahan's compensated sum; state in HBM, energies and entropies in FP64",
    "Toda": "q and p on chip as FP32 pairs hi + lo; the bond exponent and its exponential in FP64, every drift and "
            "kick increment rounded to FP32 and added by Kahan's compensated sum; state in HBM, energies and "
            "entropies in FP64",
}


def fp64_comparison(N, make_fp64, sites, w, s0, b, args, mode, pipelined, E0, nrm0, E1, nrm1, e_lin, ent):
    """The FP64 kernel on the same initial state for the same batches (timed
    the same way, for the speed ratio), and the FP32C run's norm and energy
    error against it: per-chain drift of the conserved norm and energy, the
    mode energies and entropies at the end of the run."""
    with make_fp64() as ref:
        ref.upload(s0)
        if mode == "batched":
            ref.autotune(max(1, b // 20))

        def batches(k):
            out = None
            if pipelined:  # queued three deep, as the FP32C run
                queued = 0
                for _ in range(k):
                    ref.batch_submit(b)
                    queued += 1
                    if queued == 3:
                        out = ref.batch_collect()
                        queued -= 1
                while queued:
                    out = ref.batch_collect()
                    queued -= 1
                return out
            for _ in range(k):
                if mode == "batched":
                    out = ref.batch(b)
                else:
                    ref.move(b)
                    out = ref.linenergies()
            return out
        batches(args.warmup)
        ref.sync()
        t0 = time.perf_counter()
        e64, s64 = batches(args.steps)
        ref.sync()
        el = time.perf_counter() - t0
        E64, nrm64 = ref.chain_energies()

    def drift(a, a0):
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.abs(a / a0 - 1)
        r = r[np.isfinite(r)]
        return {"max": float(r.max()), "mean": float(r.mean())} if r.size else None
    return {
        "kind": "fp32c",
        "method": FP32C_METHOD[w["model"]],
        "over_steps": (args.warmup + args.steps) * b,
        "norm_drift": drift(nrm1, nrm0) if w["model"] == "DNLS" else None, "energy_drift": drift(E1, E0),
        "fp64": {"value": sites * b * args.steps / el,
                 "norm_drift": drift(nrm64, nrm0) if w["model"] == "DNLS" else None,
                 "energy_drift": drift(E64, E0)},
        # the trajectories are chaotic: at the end of the run the FP32C and FP64
        # states differ chain by chain, the ensemble statistics do not
        "linenergies_vs_fp64": float(np.abs(e_lin - e64).max() / np.abs(e64).max()),
        "entropies_vs_fp64": [float(abs(ent[i] - s64[i]) / abs(s64[i])) for i in range(2)],
    }


def batched_form(ens, s0, sites, b, args, plain_value):
    """The batched form of the same workload and initial state, in the same
    process after the plain run: the autotuned kernel shape, every bench step
    one hipGraph replay (move + energy pass + entropies), timed the same way;
    with its energy drift beside the plain run's."""
    with ens:
        ens.upload(s0)
        report = ens.autotune(max(1, b // 20))
        ens.upload(s0)
        E0, _ = ens.chain_energies()
        for _ in range(args.warmup):
            ens.batch(b)
        ens.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ens.batch(b)
        ens.sync()
        el = time.perf_counter() - t0
        E1, _ = ens.chain_energies()
        desc = ens.describe()
    rel = np.abs(E1 - E0) / np.abs(E0)
    value = sites * b * args.steps / el
    return {"value": value, "ms_per_step": el / args.steps * 1e3, "vs_plain": value / plain_value,
            "kernel": desc, "autotune": report,
            "energy_drift": {"mean_rel": float(np.mean(rel[np.isfinite(rel)])) if np.isfinite(rel).any() else None,
                             "over_steps": (args.warmup + args.steps) * b,
                             "nonfinite_chains": int((~np.isfinite(rel)).sum())}}


def copy_calibration(mib: int) -> None:
    """Device-to-device copies of a known size (ROCm runs them as a blit
    kernel on the compute queue): the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), nbytes) == 0 and hip.hipMalloc(C.byref(b), nbytes) == 0
    assert hip.hipMemset(a, 1, nbytes) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def traffic_key(desc: str) -> str:
    """The kernel description without the fields that do not change the step
    kernel's HBM traffic (the state is read and written once per launch whatever
    the wave shape; the projection form belongs to the energy pass)."""
    drop = ("lanes_per_chain=", "sites_per_lane=", "items=", "projection=")
    return " ".join(t for t in desc.split() if not t.startswith(drop))


def traffic_bytes(desc: str):
    """HBM bytes per launch of the step kernel(s) for this kernel configuration:
    rocprofv3 FETCH_SIZE and WRITE_SIZE counts scaled by a copy kernel of known
    size, as recorded in profiles/traffic.json (None for an unprofiled one)."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    t = json.loads(f.read_text())["kernels"].get(traffic_key(desc))
    return t["hbm_bytes"] if t else None


def cpu_baseline_ragged(w, lengths, s0, seconds_budget: float):
    """Oracle release build on a sample of the ragged ensemble: the first chains
    of every length class, each class run as one oracle call."""
    from oracle import oracle as O
    lib = O.fast()
    cores = host_cores()
    offs = np.concatenate([[0], np.cumsum(lengths.astype(np.int64))])
    per_class = max(cores, 256)
    groups = []
    for n in sorted(set(int(x) for x in lengths)):
        idx = np.nonzero(lengths == n)[0][:per_class]
        groups.append(np.stack([s0[offs[i]:offs[i + 1]] for i in idx]))
    sites = sum(g.shape[0] * g.shape[1] for g in groups)
    t = time.perf_counter()
    for g in groups:
        O.move_real(O.TODA, g, 2, w["dt"], alpha=w["alpha"], threads=cores, lib=lib)
    rate = sites * 2 / (time.perf_counter() - t)
    steps = int(max(4, min(20 * w["batch"], seconds_budget * rate / sites)))

    def sample(steps):
        t = time.perf_counter()
        for g in groups:
            O.move_real(O.TODA, g, steps, w["dt"], alpha=w["alpha"], threads=cores, lib=lib)
        return time.perf_counter() - t
    el = sample(steps)
    if el < 0.5 * seconds_budget:  # the 2-step probe underestimates the rate: rescale once
        steps = max(steps, min(int(steps * seconds_budget / el), 10000 * w["batch"]))
        el = sample(steps)
    return {"value": sites * steps / el, "unit": "site-steps/s", "cores": cores, "host_cores": host_cores(),
            "kind": "port",
            "sample": f"oracle/liboracle_fast.so (-O3 -ffast-math, OpenMP over copies), Toda, first {per_class} "
                      f"chains of each length class ({sites} sites) x {steps} steps, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed bench steps (batches)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=None, help="Yoshida steps per bench step")
    ap.add_argument("--workload", default="dnkg", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="plain", choices=["plain", "batched"],
                    help="plain: move + energy pass per batch; batched: autotuned shape, one hipGraph per batch")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU-baseline time budget (0 = skip)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32c"],
                    help="fp32c: FP32 fields with compensated accumulation (DNLS, FPUT, Toda; reductions in FP64); "
                         "the line then also reports its energy (and DNLS norm) error against an FP64 run of the "
                         "same steps")
    ap.add_argument("--no-batched", action="store_true",
                    help="plain mode: skip the batched form of the same workload reported beside the plain one")
    ap.add_argument("--projection", default="auto", choices=["auto", "mfma", "mfma_wide", "valu", "bf16"],
                    help="dDNKG energy projection: FP64 MFMA (64 x 128 or 64 x 256 tiles), FP64 VALU or split BF16 "
                         "MFMA; auto = the autotuned FP64 form in --mode batched, else FP64 MFMA")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--sites", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run device-to-device copies of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "nlchains_amd runs on one GPU per process with no sharding; "
                                   "multi-GPU runs are not supported (bench.py --gpus 1)"}))
        sys.exit(2)

    import nlchains_amd as N
    w = dict(WORKLOADS[args.workload])
    if args.batch:
        w["batch"] = args.batch
    model = N.MODELS[w["model"]]
    b = w["batch"]
    ragged = w.get("ragged", False)
    pipelined = w.get("pipelined", False)
    mode = "batched" if (ragged or pipelined) else args.mode
    fp32c = args.precision == "fp32c"
    if fp32c and w["model"] not in ("DNLS", "FPUT", "Toda"):
        print(json.dumps({"error": "--precision fp32c is implemented for the DNLS, FPUT and Toda workloads"}))
        sys.exit(2)
    prec = N.FP32C if fp32c else N.FP64
    lengths = None
    if ragged:
        lengths = N.ic_ragged_lengths(w["nchains"], w["choices"], w["seed"])
        s0 = N.ic_ragged_plane_waves(lengths, w["m"], w["mode_fraction"], w["e_lo"], w["e_hi"], w["seed"])
        sites = float(lengths.sum())
        def make(precision):
            return N.Ensemble.ragged(model, lengths, w["dt"], m=w["m"], alpha=w["alpha"], beta=w["beta"],
                                     precision=precision)
        ens = make(prec)
        shape = f"{w['nchains']} chains, lengths uniform over {w['choices']}"
    else:
        n, copies = w["n"], w["copies"]
        masses = None
        if w["model"] == "DNLS":
            s0 = N.ic_dnls(n, copies, w["kmax"], w["dens"], w["seed"])
        elif w["model"] == "dDNKG":
            masses = N.ic_masses(n, w["m_lo"], w["m_hi"], w["seed"])
            w["mp2"] = masses + 2.0
            om, V = N.eigensystem(w["mp2"])
            s0 = N.ic_eigenmodes(n, copies, om, V, w["modes"], w["dens"], w["seed"])
        else:
            s0 = N.ic_plane_waves(n, copies, w["m"], w["k_lo"], w["k_hi"], w["dens"], w["seed"])
        sites = float(n) * copies
        proj = {"auto": N.PROJ_AUTO, "mfma": N.PROJ_MFMA_F64, "mfma_wide": N.PROJ_MFMA_F64_WIDE,
                "valu": N.PROJ_VALU_F64, "bf16": N.PROJ_SPLIT_BF16}[args.projection]
        def make(precision, shape_override=True):
            return N.Ensemble(model, n, copies, w["dt"], m=w["m"], alpha=w["alpha"], beta=w["beta"], masses=masses,
                              lanes_per_chain=args.lanes if shape_override else 0,
                              sites_per_lane=args.sites if shape_override else 0, projection=proj,
                              precision=precision)
        ens = make(prec)
        shape = f"N={n} copies={copies}"
        if w["model"] == "dDNKG":
            shape += f" masses U[{w['m_lo']},{w['m_hi']}] (seed {w['seed']}), {w['modes']} lowest eigenmodes excited"

    ev = HipEvents()
    tune = None
    with ens:
        ens.upload(s0)
        if mode == "batched":
            tune = ens.autotune(max(1, b // 20))
        desc = ens.describe()
        stream = ens.stream()
        E0, nrm0 = ens.chain_energies()

        def run_steps(k):
            """k bench steps; returns the last one's energies and entropies."""
            out = None
            if pipelined:
                # the CLI's pipeline: batches (one hipGraph each) queued up to three
                # deep, results collected in submission order
                queued = 0
                for _ in range(k):
                    ens.batch_submit(b)
                    queued += 1
                    if queued == 3:
                        out = ens.batch_collect()
                        queued -= 1
                while queued:
                    out = ens.batch_collect()
                    queued -= 1
                return out
            for _ in range(k):
                if mode == "batched":
                    out = ens.batch(b)  # one hipGraph: move + energy pass + entropies
                else:
                    ens.move(b)
                    out = ens.linenergies()  # energy pass + entropies, synchronous
            return out

        run_steps(args.warmup)
        ens.sync()
        t0 = time.perf_counter()
        e_lin, ent = run_steps(args.steps)
        ens.sync()
        elapsed = time.perf_counter() - t0
        E1, nrm1 = ens.chain_energies()
        # step-kernel duration from HIP events on the context stream (separate launches)
        starts = [ev.create() for _ in range(3)]
        stops = [ev.create() for _ in range(3)]
        for a_, z_ in zip(starts, stops):
            ev.record(a_, stream)
            ens.move(b)
            ev.record(z_, stream)
        kernel_ms = [ev.elapsed_ms(a_, z_) for a_, z_ in zip(starts, stops)]
        # energy pass alone (projection or FFT, fixed-order reduction, D2H of the sums)
        energy_ms = []
        for a_, z_ in zip(starts, stops):
            ev.record(a_, stream)
            ens.linenergies_raw()
            ev.record(z_, stream)
            energy_ms.append(ev.elapsed_ms(a_, z_))
        for e_ in starts + stops:
            ev.destroy(e_)
        desc = ens.describe()
        # host transfers of the whole shard over PCIe (a dump, an upload), for the
        # PCIe-inclusive rate of a run that writes the shard after every batch
        t = time.perf_counter()
        host_state = ens.download()
        download_ms = (time.perf_counter() - t) * 1e3
        t = time.perf_counter()
        ens.upload(host_state)
        upload_ms = (time.perf_counter() - t) * 1e3

    precision = None
    if fp32c:
        precision = fp64_comparison(N, lambda: make(N.FP64) if ragged else make(N.FP64, False), sites, w, s0, b,
                                    args, mode, pipelined, E0, nrm0, E1, nrm1, e_lin, ent)
    total_steps = (args.warmup + args.steps) * b
    rel = np.abs(E1 - E0) / np.abs(E0)
    drift = float(np.mean(rel[np.isfinite(rel)])) if np.isfinite(rel).any() else float("nan")
    value = sites * b * args.steps / elapsed
    k_ms = float(np.mean(kernel_ms))
    flop = flop_per_site_step(w) * sites * b
    achieved = flop / (k_ms * 1e-3) / 1e12
    kname = ("k_dnls_4096_r8" if "radix8" in desc else "k_dnls_4096_c" if fp32c and w["model"] == "DNLS"
             else "k_dnls_4096" if w["model"] == "DNLS"
             else "k_step_persistent" if "ragged_persistent" in desc
             else "k_step_wave32" if fp32c
             else "k_step_split" if "kernel=split" in desc else "k_step_wave")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if fp32c else "f64",
        "data": "synthetic (seeded initial conditions: " + ("random-phase Fourier modes" if w["model"] == "DNLS" else
                 "eigenmodes" if w["model"] == "dDNKG" else "plane waves") + ")",
        "config": {"workload": f"{w['model']} m={w['m']} alpha={w['alpha']} beta={w['beta']} {shape} dt={w['dt']}, "
                               f"{mode} {'FP32C (FP32 fields, compensated accumulation)' if fp32c else 'FP64'}, "
                               f"{b} steps + energy pass + entropies per bench step"
                               + (", batches queued three deep" if pipelined else ""),
                   "kernel": desc, "yoshida_steps_per_bench_step": b, "mode": mode},
        "energy_drift": {"mean_rel": drift, "over_steps": total_steps,
                         "nonfinite_chains": int((~np.isfinite(rel)).sum())},
        "entropies_final": [float(x) for x in ent[:2]],
        # compute-bound ("mfma" in the bound vocabulary): the step runs on the FP64
        # vector ALU, whose MI355X peak is the same 78.6 TFLOP/s as FP64 MFMA
        "roofline": {"bound": "mfma", "compute_unit": "fp64 vector ALU (same 78.6 TFLOP/s peak as fp64 MFMA)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_bytes(desc),
                     "kernel": kname, "kernel_ms": k_ms, "flop_per_launch": flop},
    }
    if fp32c:
        # the FP32 vector peak is the packed-FP32 rate; compiled without packed
        # FP32 ops, an FP32 instruction issues at the FP64 rate (78.6 TFLOP/s)
        out["roofline"].update({"compute_unit": "fp32 vector ALU (157.3 TFLOP/s with packed FP32; non-packed "
                                                "FP32, as compiled here, issues at 78.6 TFLOP/s)"
                                                + ("; the bond exponentials in fp64" if w["model"] == "Toda" else ""),
                                "peak": FP32_PEAK_TFLOPS, "frac": achieved / FP32_PEAK_TFLOPS,
                                "frac_of_nonpacked_fp32": achieved / FP64_PEAK_TFLOPS})
        out["precision"] = precision
        out["vs_fp64"] = value / precision["fp64"]["value"]
    if pipelined:
        out["us_per_batch"] = elapsed / args.steps * 1e6
    out["host_transfer"] = {"shard_bytes": int(host_state.nbytes), "download_ms": download_ms, "upload_ms": upload_ms,
                            "value_with_shard_download_every_bench_step":
                                sites * b / ((elapsed / args.steps * 1e3 + download_ms) * 1e-3)}
    e_ms = float(np.mean(energy_ms))
    out["energy_pass"] = {"ms": e_ms, "per_bench_step_fraction": e_ms / (elapsed / args.steps * 1e3)}
    if w["model"] == "dDNKG":
        # two products (copies x n) . (n x n) per pass, FMA = 2; the split BF16 form
        # executes six BF16 products per FP64 one, against the 2.5 PFLOP/s BF16 peak
        pflop = 4.0 * w["n"] ** 2 * w["copies"]
        bf16 = "split_bf16" in desc
        peak = 2500.0 if bf16 else 78.6
        hw = 6 * pflop if bf16 else pflop
        out["energy_pass"].update({"bound": "mfma", "projection": "split_bf16" if bf16 else "fp64", "flop": pflop,
                                   "achieved_tflops": pflop / (e_ms * 1e-3) / 1e12, "peak_tflops": peak,
                                   "frac": hw / (e_ms * 1e-3) / 1e12 / peak})
    if tune:
        out["config"]["autotune"] = tune
    if mode == "plain" and not fp32c and not args.no_batched:
        out["batched"] = batched_form(make(prec, False), s0, sites, b, args, value)
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = (cpu_baseline_ragged(w, lengths, s0, args.cpu_seconds) if ragged
                               else cpu_baseline(w, s0, args.cpu_seconds))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
