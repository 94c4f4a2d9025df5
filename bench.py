"""Benchmark: GFN1 energy+forces molecules/s of the ragged serving loop.

Workload: the serving stream of log-uniform 5..200-atom molecules (70 % CHNO
chains, 30 % water clusters), energy + analytic forces in FP64, seed 4.  A
"step" hands the next `--chunk` requests of the stream, all available at
once, to the continuous-batching serving loop, which admits them as streams
free up while the previous chunk's molecules finish (saturation throughput
of the loop; the timed region ends when every chunk is done).  The K steps
are timed `--repeats` times; the value is the median.  The JSON line also
carries

* the same chunks served one at a time, each drained before the next starts;
* the plain padded-batch drop-in on the same requests (arrival order,
  `--padded-batch` molecules per padded batch: 64, and 256, which fills the
  chip's 256 CUs) and the ragged/padded ratios;
* p50/p99 latency of the serving loop on the whole 20000-request stream under
  Poisson arrivals at 30/50/70/90 % of the measured saturation rate
  (`latency_curve`; the headline p50/p99 at 70 %);
* a roofline entry for the dominant kernel (the fused single-point kernel
  the serving loop launches: each molecule's integral pass, SCF and forces
  on its own CU, FP64): the SCF's algorithmic flops / duration from HIP
  events of one launch over requests 600-2999 (four chunks), alone on the
  GPU, with its HBM traffic and executed matrix-core flops from the
  committed rocprofv3 summaries of that launch; the SCF kernel alone on the
  same batch beside it, and the stage rates of the integral pass and of the
  SCF's Fock-build and Mulliken stages;
* the CPU baseline: the oracle (a PyTorch FP64 restatement of the upstream
  algorithm) timed on this host on a bounded prefix of the same stream.

    python bench.py [--steps K] [--warmup W] [--chunk R]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "GFN1 energy+forces molecules/s vs plain padded batch; p50/p99 latency @1e-8 Eh"
FP64_PEAK_TFLOPS = 78.6


def _requests(stream, start, count):
    from dxtb_amd.serving import Request

    out = []
    for i, (t, z, p) in enumerate(stream[start:start + count]):
        out.append(Request(rid=start + i, numbers=z, positions=p, arrival=t))
    return out


def cpu_baseline(stream, seconds: float) -> dict:
    import torch

    import oracle.gfn1 as og  # CPU restatement of the upstream algorithm

    opts = {"x_atol": 1e-10, "x_atol_max": 1e-10, "maxiter": 300}
    t0 = time.perf_counter()
    n = 0
    atoms = 0
    for _, z, p in stream:
        og.singlepoint(z, p, opts=opts, with_forces=True)
        n += 1
        atoms += len(z)
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    threads = torch.get_num_threads()
    # the same requests on one thread, for a bounded sample
    torch.set_num_threads(1)
    t1 = time.perf_counter()
    n1 = 0
    for _, z, p in stream[:n]:
        og.singlepoint(z, p, opts=opts, with_forces=True)
        n1 += 1
        if time.perf_counter() - t1 > seconds / 3:
            break
    dt1 = time.perf_counter() - t1
    torch.set_num_threads(threads)
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n / dt, "unit": "molecules/s", "cores": threads,
            "kind": "port", "single_thread_value": n1 / dt1,
            "sample": f"first {n} requests of the same stream ({atoms} atoms), energy+forces, "
                      f"sequential, {dt:.1f} s, torch {torch.__version__} CPU FP64, "
                      f"{threads} torch threads on {os.cpu_count()} logical CPUs ({model}); "
                      f"single thread: first {n1} requests, {dt1:.1f} s"}


def forces_cost(reqs, device, opts, repeats: int = 3) -> dict:
    """Wall time of one ragged batch for energy only and for energy+forces
    (median of `repeats`, after one warm-up of each)."""
    import torch

    from dxtb_amd.engine import BatchLayout, DeviceTables, Pipeline
    from dxtb_amd.param import SpeciesTables

    mols = [r.numbers for r in reqs]
    tb = SpeciesTables(np.unique(np.concatenate(mols)))
    pipe = Pipeline(BatchLayout(mols, tb, padded=False, chrg=np.zeros(len(mols))),
                    DeviceTables(tb, device), device)
    pos = torch.as_tensor(np.concatenate([r.positions for r in reqs]), dtype=torch.float64,
                          device=device).contiguous()
    out = {}
    for forces in (False, True):
        ts = []
        for k in range(repeats + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pipe.run(pos, opts, forces=forces)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        out["forces" if forces else "energy"] = float(np.median(ts[1:]))
    # parameter gradients of the same batch (after its last run)
    ts = []
    for k in range(repeats + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pipe.param_grad()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return {"energy_ms": round(out["energy"] * 1e3, 2), "energy_forces_ms": round(out["forces"] * 1e3, 2),
            "ratio": round(out["forces"] / out["energy"], 3), "molecules": len(reqs),
            "param_grad_ms": round(float(np.median(ts[1:])) * 1e3, 3)}


def md_warm_start(device, opts, steps: int = 8) -> dict:
    """Energy + forces along a short molecular-dynamics-like trajectory
    (constant drift of 0.01 bohr per step plus 0.002 bohr noise) of a
    66-water cluster and a 66-atom chain in one batch: per-step wall time and
    SCF iterations of cold starts vs warm starts (orbitals and potential
    kept from the previous step) and warm starts from the potential
    extrapolated from the last two steps, median over steps 2..steps-1."""
    import dataclasses

    import torch

    from dxtb_amd import workloads as W
    from dxtb_amd.engine import BatchLayout, DeviceTables, Pipeline
    from dxtb_amd.param import SpeciesTables

    rng = np.random.default_rng(17)
    mols = [W.water_cluster(rng, 66), W.chain(rng, 66)]
    p0 = np.concatenate([p for _, p in mols])
    v = 0.01 * rng.standard_normal(p0.shape)
    traj = [p0 + k * v + 0.002 * rng.standard_normal(p0.shape) for k in range(steps)]
    tb = SpeciesTables(np.unique(np.concatenate([z for z, _ in mols])))
    out = {}
    for mode in ("cold", "warm", "warm_extrapolated"):
        warm = mode != "cold"
        o = dataclasses.replace(opts, warm_start=warm, warm_extrapolate=mode == "warm_extrapolated")
        pipe = Pipeline(BatchLayout([z for z, _ in mols], tb, padded=False, chrg=np.zeros(len(mols))),
                        DeviceTables(tb, device), device)
        ts, its = [], []
        for p in traj:
            pos = torch.as_tensor(p, dtype=torch.float64, device=device).contiguous()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pipe.run(pos, o, forces=True)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            its.append(float(pipe.records()[:, 5].sum().item()))
        # steps 2.. (the extrapolation needs two kept potentials)
        out[mode] = (float(np.median(ts[2:])), float(np.median(its[2:])))
    return {"batch": "66 H2O + 66-atom chain (n = 528 + 506)", "steps": steps,
            "cold_ms": round(out["cold"][0] * 1e3, 2), "warm_ms": round(out["warm"][0] * 1e3, 2),
            "warm_extrapolated_ms": round(out["warm_extrapolated"][0] * 1e3, 2),
            "cold_iterations": out["cold"][1], "warm_iterations": out["warm"][1],
            "warm_extrapolated_iterations": out["warm_extrapolated"][1],
            "speedup": round(out["cold"][0] / out["warm"][0], 3),
            "speedup_extrapolated": round(out["cold"][0] / out["warm_extrapolated"][0], 3)}


def large_system(device, opts, repeats: int = 3) -> dict:
    """Config 3: one 600-water cluster (1800 atoms, 4800 orbitals), energy +
    analytic forces on the whole GPU (xtb_scf_large).  One warm-up run, then
    `repeats` timed runs (host clock around the whole call, synchronised);
    the median, with the SCF's own stage times from its result record."""
    import torch

    from dxtb_amd import _lib
    from dxtb_amd import workloads as W
    from dxtb_amd.large import LargeSystem

    z, p = W.config3()
    ls = LargeSystem(z, device, opts=opts)
    pos = torch.as_tensor(p, dtype=torch.float64, device=device)
    r = ls.run(pos, forces=True)
    times = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = ls.run(pos, forces=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    rec = r["records"]
    ms = sorted(times)[len(times) // 2] * 1e3
    return {"system": "600 H2O (1800 atoms, 4800 orbitals)", "ms": round(ms, 1),
            "repeats_ms": [round(t * 1e3, 1) for t in times],
            "energy": r["energy"], "iterations": r["iterations"], "status": r["status"],
            "scf_ms": round(float(rec[_lib.OUT_T_TOTAL]) * 1e-3, 1),
            "orthogonaliser_ms": round(float(rec[_lib.OUT_T_ORTHO]) * 1e-3, 1),
            "eigen_ms": round(float(rec[_lib.OUT_T_EIGEN]) * 1e-3, 1),
            "perturbative_ms": round(float(rec[_lib.OUT_T_PD]) * 1e-3, 1),
            "perturbative_steps": int(rec[_lib.OUT_PD_STEPS]),
            "jacobi_sweeps": int(rec[_lib.OUT_SWEEPS])}


def device_peaks(device) -> dict:
    """Achievable FP64 GEMM rate (torch.matmul, n=8192) and device copy
    bandwidth (1 GiB), for context next to the published peaks."""
    import torch

    a = torch.randn(8192, 8192, dtype=torch.float64, device=device)
    b = torch.randn_like(a)
    for _ in range(2):
        a @ b
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        a @ b
    torch.cuda.synchronize()
    gemm = 3 * 2 * 8192.0**3 / (time.perf_counter() - t0) / 1e12
    del a, b
    x = torch.empty(2**27, dtype=torch.float64, device=device)
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    bw = 5 * 2 * x.numel() * 8 / (time.perf_counter() - t0) / 1e9
    return {"dgemm_tflops": round(gemm, 2), "copy_gbs": round(bw, 1)}


# the roofline pass's workload: requests 600-2999 of the seed-4 stream (as
# many as the serving loop holds: 4 chunks of 600) in one launch.  Fixed, so
# that the committed rocprofv3 summaries (traffic, matrix-core counters) are
# of exactly the launch this run times, whatever --warmup or --chunk are.
ROOFLINE_START, ROOFLINE_COUNT = 600, 2400
HBM_PEAK_GBS = 8000.0
NCU = 256


def roofline_pass(stream, device, opts, start: int = ROOFLINE_START, count: int = ROOFLINE_COUNT) -> dict:
    """The SCF kernel alone in one launch: `count` requests of the stream
    (as many as the serving loop holds) in one ragged batch, one workgroup
    per molecule, largest first, so that the dispatcher fills CUs as they
    free up; one stream, HIP events around the k_scf launch on that stream,
    energy + forces settings; algorithmic flops from the measured iteration
    counts.  Also the integral pass (CN, S, H0, gamma) of the same batch
    with its algorithmic bytes (16 n^2 for S and H0, 8 m^2 for gamma), and
    the SCF kernel's Fock-build and Mulliken/potential stages from its
    device timers (24 n^2 and 16 n^2 + 8 m^2 bytes per pass)."""
    import torch

    from dxtb_amd import _lib
    from dxtb_amd.engine import BatchLayout, DeviceTables, Pipeline
    from dxtb_amd.param import SpeciesTables
    from dxtb_amd.serving import _process_streams

    reqs = _requests(stream, start, count)
    tb = SpeciesTables((1, 6, 7, 8))
    lay = BatchLayout([r.numbers for r in reqs], tb, padded=False, chrg=np.zeros(count))
    # one of the serving loop's process streams: every stream beyond the
    # hardware queues shares a queue, and later serving runs slowed down
    st = _process_streams(torch.device(device), 1)[0]
    dt = DeviceTables(tb, device)
    with torch.cuda.stream(st):
        pipe = Pipeline(lay, dt, device)
        pos = torch.as_tensor(np.concatenate([r.positions for r in reqs]), dtype=torch.float64).to(device)
        pipe.run(pos, opts, forces=True, stream=st, timing=True)
        ef = torch.cuda.Event(enable_timing=True)
        ef.record(st)
    st.synchronize()
    ev0, ev1 = pipe.scf_events
    ms = float(ev0.elapsed_time(ev1))
    forces_ms = float(ev1.elapsed_time(ef))
    ei0, ei1 = pipe.int_events
    int_ms = float(ei0.elapsed_time(ei1))
    fl = pipe.scf_flops()
    rec = pipe.records().cpu().numpy()
    bad = int((rec[:, _lib.OUT_STATUS] != 0).sum())
    achieved = fl / (ms * 1e-3) / 1e12
    # the same batch as ONE fused launch (xtb_singlepoint: each molecule's
    # integral pass, SCF and gradient on its own CU), as the serving loop runs it
    with torch.cuda.stream(st):
        fpipe = Pipeline(lay, dt, device)
        # the timed launch is the second: a kernel's first launch in a
        # process also sets it up, which the events would count
        for _ in range(2):
            fpipe.run(pos, opts, forces=True, stream=st, timing=True, fused=True)
    st.synchronize()
    fv0, fv1 = fpipe.scf_events
    fms = float(fv0.elapsed_time(fv1))
    frec = fpipe.records().cpu().numpy()
    fbusy = float(frec[:, _lib.OUT_T_TOTAL].sum()) * 1e-6
    fused = {"kernel": "xtb::k_singlepoint", "mean_launch_ms": fms, "achieved": fl / (fms * 1e-3) / 1e12,
             "cu_occupancy": fbusy / (NCU * fms * 1e-3),
             "longest_molecule_ms": float(frec[:, _lib.OUT_T_TOTAL].max()) * 1e-3,
             "integrals_workgroup_s": float(frec[:, _lib.OUT_T_INTEGRALS].sum()) * 1e-6,
             "forces_workgroup_s": float(frec[:, _lib.OUT_T_FORCES].sum()) * 1e-6,
             "same_energies": bool(np.array_equal(pipe.energies().cpu().numpy(), fpipe.energies().cpu().numpy())),
             "three_launches_ms": int_ms + ms + forces_ms}
    # the same flops over the workgroups' own device time (one workgroup per
    # molecule, one CU each): the kernel's efficiency while resident,
    # independent of how many CUs are busy
    per_cu_peak = FP64_PEAK_TFLOPS * 1e12 / NCU
    busy = float(rec[:, _lib.OUT_T_TOTAL].sum()) * 1e-6
    longest = float(rec[:, _lib.OUT_T_TOTAL].max()) * 1e-3
    n2 = lay.n.astype(np.float64) ** 2
    m2 = lay.nsh.astype(np.float64) ** 2
    int_bytes = float(16.0 * n2.sum() + 8.0 * m2.sum())
    # stages inside the SCF kernel: bytes per workgroup-microsecond; one
    # workgroup owns one CU, whose share of the HBM peak is peak / 256
    nf, tf = rec[:, _lib.OUT_N_FOCK], rec[:, _lib.OUT_T_FOCK]
    nm, tmu = rec[:, _lib.OUT_N_MULL], rec[:, _lib.OUT_T_MULL]
    steps = rec[:, _lib.OUT_ITER] + 1.0
    fock_bytes = float((24.0 * nf * n2).sum())
    # the last step's Mulliken pass reads P, S and H0 (nm = 1.5 units of
    # 16 n^2); the other steps take q from the density product's tiles and
    # read only gamma here
    mull_bytes = float((16.0 * nm * n2 + 8.0 * steps * m2).sum())
    fock_gbs = fock_bytes / (float(tf.sum()) * 1e-6) / 1e9
    mull_gbs = mull_bytes / (float(tmu.sum()) * 1e-6) / 1e9
    share = HBM_PEAK_GBS / NCU
    stages = {
        "integral_pass": {"ms": round(int_ms, 3), "bytes": int_bytes,
                                 "gbs": round(int_bytes / (int_ms * 1e-3) / 1e9, 1),
                                 "frac_of_hbm_peak": round(int_bytes / (int_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "bytes_rule": "16 n^2 (S, H0) + 8 m^2 (gamma) per molecule; CN and repulsion in the same launch"},
        "fock_build": {"passes": int(nf.sum()), "workgroup_s": round(float(tf.sum()) * 1e-6, 4),
                       "gbs_per_workgroup": round(fock_gbs, 2),
                       "frac_of_cu_share": round(fock_gbs / share, 4),
                       "share_of_scf_time": round(float(tf.sum()) * 1e-6 / busy, 4),
                       "bytes_rule": "24 n^2 per pass (read H0, S; write F)"},
        "mulliken_potential": {"passes": int(round(float(steps.sum()))),
                               "workgroup_s": round(float(tmu.sum()) * 1e-6, 4),
                               "gbs_per_workgroup": round(mull_gbs, 2),
                               "frac_of_cu_share": round(mull_gbs / share, 4),
                               "share_of_scf_time": round(float(tmu.sum()) * 1e-6 / busy, 4),
                               "bytes_rule": "8 m^2 (gamma) per step; + 24 n^2 (P, S, H0) in the last step, "
                                             "the others take q from the density product's tiles"},
        "cu_share_gbs": share,
    }
    stages["forces_ms"] = round(forces_ms, 3)
    return {"achieved": achieved, "launches": 1, "mean_launch_ms": ms, "flops_per_launch": fl,
            "molecules": count, "requests": f"{start}-{start + count - 1}", "nonconverged": bad,
            "resident_frac": fl / (busy * per_cu_peak), "cu_occupancy": busy / (NCU * ms * 1e-3),
            "longest_molecule_ms": longest, "stages": stages, "fused": fused}


def _profile_json(name: str, start: int, count: int) -> dict | None:
    """A committed rocprofv3 summary of the roofline pass's k_scf launch;
    None when it is missing or of another workload."""
    f = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(f) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("start") != start or t.get("count") != count or t.get("seed") != 4:
        return None
    return t


def measured_traffic(kernel: str, start: int = ROOFLINE_START, count: int = ROOFLINE_COUNT) -> float | None:
    """HBM bytes of the roofline pass's launch of `kernel` (k_singlepoint or
    k_scf), from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs (copy-calibrated)
    summarised in profiles/<kernel>_traffic.json."""
    t = _profile_json(f"{kernel}_traffic.json", start, count)
    return None if t is None else float(t["traffic_bytes_per_launch"])


def measured_mfma(kernel: str, ms: float, start: int = ROOFLINE_START,
                  count: int = ROOFLINE_COUNT) -> dict | None:
    """Executed FP64 matrix-core work of the same launch from the rocprofv3
    SQ_INSTS_VALU_MFMA_MOPS_F64 run (units of 512 flops) summarised in
    profiles/<kernel>_mfma.json, over this run's launch time: what the matrix
    cores executed, next to the fixed-convention flops of `achieved`."""
    t = _profile_json(f"{kernel}_mfma.json", start, count)
    if t is None:
        return None
    fl = float(t["mfma_f64_flops_per_launch"])
    out = {"executed_flops_per_launch": fl, "executed_tflops": round(fl / (ms * 1e-3) / 1e12, 3),
           "executed_frac": round(fl / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)}
    if "mfma_busy_frac" in t:
        out["mfma_busy_frac"] = t["mfma_busy_frac"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--lookahead", type=int, default=8,
                    help="chunks the serving loop holds (waiting or in flight) at most")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="HIP hardware queues of the process (sets GPU_MAX_HW_QUEUES, 1-16)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--merge-above", type=int, default=1200,
                    help="waiting requests above which a launch takes --merge-cap of them whatever "
                         "their size class, largest first (-1: class launches only)")
    ap.add_argument("--merge-cap", type=int, default=2400)
    ap.add_argument("--merge-oldest", action="store_true",
                    help="merged launches take the oldest waiting requests instead of the largest")
    ap.add_argument("--no-fused", action="store_true",
                    help="serving launches as integrals, SCF and forces kernels instead of one fused kernel")
    ap.add_argument("--chunk", type=int, default=600)
    ap.add_argument("--latency-requests", type=int, default=20000,
                    help="requests of the stream served under Poisson arrivals per rate")
    ap.add_argument("--latency-rates", default="0.3,0.5,0.7,0.9",
                    help="arrival rates as fractions of the measured saturation rate")
    ap.add_argument("--repeats", type=int, default=3,
                    help="timed repetitions of the K steps (value = median)")
    ap.add_argument("--padded-batch", default="64,256",
                    help="molecules per padded batch of the drop-in (first = headline ratio)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-padded", action="store_true")
    ap.add_argument("--no-barrier", action="store_true",
                    help="skip the comparison run that drains each chunk before the next")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--profile-run", action="store_true",
                    help="serve one chunk only (for rocprofv3 runs), print nothing else")
    ap.add_argument("--profile-serving", action="store_true",
                    help="only serve 8 chunks through the continuous-batching loop (kernel traces)")
    ap.add_argument("--no-large", action="store_true", help="skip the 600-water single system")
    ap.add_argument("--no-latency", action="store_true", help="skip the latency runs")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the forces-cost, MD and device-peak measurements")
    ap.add_argument("--profile-large", action="store_true",
                    help="only the 600-water single system, energy + forces once (kernel traces)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(molecules are independent: shard requests across GPUs by a router).",
              file=sys.stderr)
        return 2

    # more hardware queues let more micro-batches run side by side (one
    # stream per queue); it must be set before HIP initialises
    # each queue gets its own scratch reservation for the SCF kernel's
    # stack: 24 queues ran out of resources at launch
    if not 1 <= a.hw_queues <= 16:
        print("--hw-queues must be in 1..16", file=sys.stderr)
        return 2
    os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
    import gc

    import torch

    from dxtb_amd import workloads as W
    from dxtb_amd.engine import ScfOptions
    from dxtb_amd.serving import RaggedServer, padded_dropin

    log = (lambda *x: None) if a.quiet else (lambda *x: print(*x, file=sys.stderr, flush=True))
    dev = torch.device("cuda:0")
    if a.copy_probe:
        x = torch.ones(a.copy_probe * 2**20 // 8, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": a.copy_probe, "bytes_read": x.numel() * 8,
                          "bytes_written": y.numel() * 8}))
        return 0
    opts = ScfOptions(x_atol=1e-10, x_atol_max=1e-10, maxiter=300)
    merge = {"merge_above": a.merge_above if a.merge_above >= 0 else None, "merge_cap": a.merge_cap,
             "merge_largest": not a.merge_oldest, "fused": not a.no_fused}
    if a.profile_large:
        print(json.dumps({"profile_large": large_system(dev, opts, repeats=1)}))
        return 0
    total = max((a.warmup + a.steps) * a.chunk, a.latency_requests, ROOFLINE_START + ROOFLINE_COUNT)
    stream = W.config4(seed=4, count=total)
    log(f"stream: {total} requests generated")

    if a.profile_run:
        # the roofline pass alone (for rocprofv3 runs): its mean launch time is
        # the figure the kernel trace's average duration of k_scf must match
        rp = roofline_pass(stream, dev, opts)
        print(json.dumps({"profile_run": rp}))
        return 0
    server = RaggedServer(dev, opts=opts, **merge)
    if a.profile_serving:
        # the continuous-batching loop over 8 chunks, after one warm-up chunk
        server.run_stream(_requests(stream, 0, a.chunk), realtime=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        held, served = [], []
        for k in range(8):
            reqs = _requests(stream, (k + 1) * a.chunk, a.chunk)
            server.feed(reqs)
            held.append(reqs)
            served += reqs
            if len(held) >= a.lookahead:
                server.serve_until(held.pop(0))
        server.serve_until(served)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        print(json.dumps({"profile_serving": {"molecules": len(served), "seconds": sec,
                                              "molecules_per_s": len(served) / sec,
                                              "launches": server.stats.launches}}))
        return 0
    pos = 0
    for w in range(a.warmup):
        server.run_stream(_requests(stream, pos, a.chunk), realtime=False)
        pos += a.chunk
        log(f"warmup {w} done")
    # the start-up heap (the request stream, torch) is frozen so that Python's
    # full collections, ~0.1 s each over it, do not stall the host loop
    gc.collect()
    gc.freeze()
    # continuous batching: step k hands chunk k to the loop, which admits it
    # as streams free up while earlier chunks finish; once `lookahead` chunks
    # are held, the step ends when the oldest of them is done.  The timed
    # region ends when every chunk is done.  Repeated `--repeats` times on
    # the same chunks; the value is the median.
    ncu = torch.cuda.get_device_properties(dev).multi_processor_count
    runs = []
    for rep in range(a.repeats):
        torch.cuda.synchronize()
        busy0 = server.stats.scf_busy_s
        pos = a.warmup * a.chunk
        t0 = time.perf_counter()
        served, held = [], []
        for k in range(a.steps):
            reqs = _requests(stream, pos, a.chunk)
            server.feed(reqs)
            held.append(reqs)
            if len(held) >= a.lookahead:
                server.serve_until(held.pop(0))
            served += reqs
            pos += a.chunk
            if rep == 0:
                log(f"step {k} handed over")
        server.serve_until(served)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bad = sum(1 for r in served if server.results[r.rid].status != 0)
        # share of the CU-time the SCF kernel's workgroups held (one CU each)
        occ_r = (server.stats.scf_busy_s - busy0) / (ncu * (t1 - t0))
        runs.append((a.steps * a.chunk / (t1 - t0), (t1 - t0) * 1e3 / a.steps, occ_r, bad))
        log(f"repeat {rep}: ragged {runs[-1][0]:.1f} molecules/s")
        server.results.clear()
    order = sorted(range(len(runs)), key=lambda i: runs[i][0])
    value, ms_step, occupancy, bad = runs[order[len(runs) // 2]]
    bad = max(r[3] for r in runs)
    log(f"ragged: {value:.1f} molecules/s (median of {len(runs)})")
    # the same chunks one at a time, each drained before the next starts
    barrier = None
    if not a.no_barrier:
        bsrv = RaggedServer(dev, opts=opts, **merge)
        bsrv.run_stream(_requests(stream, 0, a.chunk), realtime=False)  # warm-up
        torch.cuda.synchronize()
        tb0 = time.perf_counter()
        for k in range(a.steps):
            bsrv.run_stream(_requests(stream, (a.warmup + k) * a.chunk, a.chunk), realtime=False)
        torch.cuda.synchronize()
        barrier = a.steps * a.chunk / (time.perf_counter() - tb0)
        bsrv.results.clear()
        log(f"ragged, chunks drained one at a time: {barrier:.1f} molecules/s")

    # roofline of the SCF kernel: one launch over the requests the loop holds
    rp = roofline_pass(stream, dev, opts)
    fz = rp["fused"]
    achieved = fz["achieved"]
    log(f"fused kernel: one launch of {rp['molecules']} molecules, {fz['mean_launch_ms']:.2f} ms, "
        f"{achieved:.3f} TFLOP/s, CU occupancy {fz['cu_occupancy']:.3f}, "
        f"longest molecule {fz['longest_molecule_ms']:.1f} ms (integrals + SCF + forces: three launches "
        f"{fz['three_launches_ms']:.2f} ms; same energies: {fz['same_energies']})")
    log(f"scf kernel alone: {rp['mean_launch_ms']:.2f} ms, {rp['achieved']:.3f} TFLOP/s, "
        f"CU occupancy {rp['cu_occupancy']:.3f}, longest molecule {rp['longest_molecule_ms']:.1f} ms")
    log(f"stages: {json.dumps(rp['stages'])}")

    # latency under Poisson arrivals at fractions of the saturation rate: the
    # first `--latency-requests` requests of the stream (20000: the whole
    # serving-stream configuration), the serving loop on the same process
    # streams, the heap frozen again as after start-up
    served = held = None
    gc.collect()
    gc.freeze()
    curve = []
    for frac in ([] if a.no_latency else [float(x) for x in a.latency_rates.split(",")]):
        r_ = frac * value
        lat_reqs = _requests(stream, 0, a.latency_requests)
        rng = np.random.default_rng(44)
        t = 0.0
        for r in lat_reqs:
            t += rng.exponential(1.0 / r_)
            r.arrival = t
        ls = RaggedServer(dev, opts=opts, **merge)
        lat = ls.run_stream(lat_reqs, realtime=True)
        log(f"latency @ {r_:.0f}/s: p50 {lat['p50_ms']:.1f} ms p99 {lat['p99_ms']:.1f} ms "
            f"({lat['launches']} launches, host admission {ls.stats.host_admit_s:.3f} s, "
            f"{lat['seconds']:.3f} s for {len(lat_reqs)} requests)")
        curve.append({"rate_frac": frac, "rate_per_s": round(r_, 1), "served_per_s": round(lat["molecules_per_s"], 1),
                      "p50_ms": round(lat["p50_ms"], 2), "p99_ms": round(lat["p99_ms"], 2)})
        ls.results.clear()
        del lat_reqs, ls
        gc.collect()
    head = min(curve, key=lambda c: abs(c["rate_frac"] - 0.7)) if curve else {
        "rate_per_s": None, "p50_ms": None, "p99_ms": None}

    padded = {}
    if not a.no_padded:
        pr = _requests(stream, a.warmup * a.chunk, a.chunk)
        for pb in [int(x) for x in a.padded_batch.split(",")]:
            padded[pb] = padded_dropin(pr, pb, dev, opts=opts)["molecules_per_s"]
            log(f"padded drop-in, batches of {pb}: {padded[pb]:.1f} molecules/s")
    pb0 = int(a.padded_batch.split(",")[0])

    fc = None if a.no_extras else forces_cost(_requests(stream, a.warmup * a.chunk, min(a.chunk, 128)), dev, opts)
    log(f"forces cost: {fc}")
    md = None if a.no_extras else md_warm_start(dev, opts)
    log(f"md warm start: {md}")
    big = None if a.no_large else large_system(dev, opts)
    log(f"600-water cluster: {big}")
    peaks = None if a.no_extras else device_peaks(dev)
    log(f"measured peaks: {peaks}")

    cpu = None if a.no_cpu else cpu_baseline(stream, a.cpu_seconds)
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3f} molecules/s")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "molecules/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "serving stream (log-uniform 5-200 atoms; 70% CHNO chains, 30% water "
                        "clusters), energy + analytic forces, FP64, SCF x_atol 1e-10",
            "requests_per_step": a.chunk, "seed": 4,
            "batching": f"continuous: step k hands its chunk to the serving loop, which holds "
                        f"at most {a.lookahead} chunks and admits each as streams free up",
            "streams": server.streams.__len__(), "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
            "repeats_molecules_per_s": [round(r[0], 1) for r in runs],
            "scf_cu_occupancy": round(occupancy, 3),
            "chunk_barrier_molecules_per_s": None if barrier is None else round(barrier, 3),
            "padded_dropin_molecules_per_s": padded.get(pb0),
            "ragged_over_padded": round(value / padded[pb0], 3) if pb0 in padded else None,
            "padded_batch": pb0,
            "padded_dropin_by_batch": {str(k): round(v, 2) for k, v in padded.items()},
            "ragged_over_padded_by_batch": {str(k): round(value / v, 2) for k, v in padded.items()},
            "padded_dropin_eigen": "padded layout: every matrix at the batch-maximum n; the same eigen "
                                   "options as the ragged loop (occupation-class and perturbative steps "
                                   "where a molecule's frontier gap allows, the padding levels at energy 0 "
                                   "taking part in its spectrum); one fused launch per batch",
            "latency_requests": a.latency_requests,
            "latency_rate_per_s": head["rate_per_s"], "p50_ms": head["p50_ms"],
            "p99_ms": head["p99_ms"], "latency_curve": curve, "nonconverged": bad,
            "forces_over_energy": fc, "md_warm_start": md,
            "single_large_system": big,
        },
        "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                     "traffic": measured_traffic("k_singlepoint"),
                     "kernel": "xtb::k_singlepoint",
                     "workload": f"one launch: requests {rp['requests']} of the stream, largest first, "
                                 "integral pass + SCF + forces per molecule (as the serving loop runs it)",
                     "launches": 1, "mean_launch_ms": round(fz["mean_launch_ms"], 3),
                     "flops_per_launch": rp["flops_per_launch"],
                     "flops_rule": "the SCF's fixed convention (9 1/3 n^3 per step + density + "
                                   "orthogonaliser + W); the integral pass and forces (FP64 VALU) add none",
                     "cu_occupancy": round(fz["cu_occupancy"], 4),
                     "longest_molecule_ms": round(fz["longest_molecule_ms"], 2),
                     "matrix_cores": measured_mfma("k_singlepoint", fz["mean_launch_ms"]),
                     "three_launches_ms": round(fz["three_launches_ms"], 3),
                     "integrals_workgroup_s": round(fz["integrals_workgroup_s"], 4),
                     "forces_workgroup_s": round(fz["forces_workgroup_s"], 4),
                     "same_energies_as_three_launches": fz["same_energies"],
                     "scf_kernel": {"kernel": "xtb::k_scf", "mean_launch_ms": round(rp["mean_launch_ms"], 3),
                                    "achieved": round(rp["achieved"], 4),
                                    "frac": round(rp["achieved"] / FP64_PEAK_TFLOPS, 5),
                                    "traffic": measured_traffic("k_scf"),
                                    "resident_frac": round(rp["resident_frac"], 4),
                                    "cu_occupancy": round(rp["cu_occupancy"], 4),
                                    "longest_molecule_ms": round(rp["longest_molecule_ms"], 2),
                                    "matrix_cores": measured_mfma("k_scf", rp["mean_launch_ms"])},
                     "stages": rp["stages"],
                     "measured_peaks": peaks},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
