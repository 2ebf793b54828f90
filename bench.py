"""H-matvec benchmark on one MI355X.

Workload (the single-GPU benchmark configuration): two disjoint icospheres of
radius 1, centres 3 apart, 81,920 flat triangles each (163,840 unknowns),
piecewise-constant collocation at centroids, kernel 1/(4 pi |x-y|) (centroid
quadrature off the diagonal, analytic self terms), bisection cluster tree
with leaf size 20, admissibility eta = 2, ACA+ tolerance 1e-5, x ~ U[-1,1]
from seed 1.  All data synthetic (generated on the host by libhmvgen).

One step = one FP64 H-matvec y = H x over the whole H-matrix with x, y and
the arena resident in HBM.  value = algorithmic bytes per matvec
(8*E + 16*N, E = stored entries) x steps / wall time of the timed steps.

Also reported: per-phase HIP-event timings and the roofline of the dominant
kernel, BiCGSTAB time-to-1e-8 in FP64 and with FP32 / BF16 storage under
FP64 iterative refinement (each solution's error against the oracle's FP64
solution), the CPU baseline (the oracle's restatement of the reference CPU
path, per-leaf GEMV, timed on this host's cores at the thread count of its
CPU share and at 1 thread, and its BiCGSTAB to 1e-8), and the leaf-rank /
leaf-size histograms.  Unless --no-extras: on one icosphere of 327,680
unknowns the per-precision apply rates and errors, FP64 vs mixed-precision
refined solves, the multi-RHS H-matmat (k = 8, 32, 128) and block BiCGSTAB
(FP64, and FP32 factors refined in FP64) at k = 8, 32, 128; and a batch of 48
uneven instances in one arena, solved in FP64 and with autotuned BF16
factors under per-instance FP64 refinement.
"""
import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H-matvec HBM GB/s + BiCGSTAB time-to-1e-8 residual, mixed vs FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_two_spheres_163840.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic(kernel_phase):
    """HBM bytes per launch of the given phase from the committed rocprofv3
    PMC summary (lines 'traffic_bytes_per_launch phase<k> <bytes>'), or None."""
    if not os.path.exists(PMC_FILE):
        return None
    for line in open(PMC_FILE):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel_phase:
            return float(parts[2])
    return None


# what the counter passes time: the two FP64 apply kernels and the
# calibration kernels of known traffic
_APPLY_KERNELS = {"phase1": "hmv_stream_kernel<double, 1, 0, 0,",
                  "phase2": "hmv_stream_kernel<double, 2, 0, 0,"}
CAL_READ_BYTES = 2 << 30   # hmv_read_kernel reads this many bytes
CAL_COPY_BYTES = 1 << 30   # hmv_copy_kernel writes this many bytes


def traffic_pass(H, subdiv):
    """Counter pass (run by measure_traffic under rocprofv3): the benchmark's
    arena, the calibration kernels and six applies, nothing timed."""
    L = H.hmv_lib()
    if L.hmv_init(0) != 0:
        raise SystemExit("no usable GPU: " + L.hmv_last_error().decode())
    mesh = H.merge(H.icosphere(subdiv), H.icosphere(subdiv, 1.0, (3.0, 0.0, 0.0)))
    hm = H.HMatrix(*mesh, H.baseline_params(aca_eps=1e-5))
    a64 = H.Arena(hm, "f64")
    dx = H.DeviceBuffer.from_array(H.uniform(1, hm.nd))
    dy = H.DeviceBuffer(8 * hm.nd)
    H.stream_test(0, CAL_READ_BYTES, reps=2)
    H.stream_test(1, CAL_COPY_BYTES, reps=2)
    for _ in range(6):
        a64.apply(dx.ptr, dy.ptr, 0.0)
    L.hmv_device_synchronize()


def measure_traffic(subdiv, log):
    """HBM bytes per launch of each apply kernel, measured in this run: two
    `rocprofv3 --pmc` passes over `bench.py --traffic-pass` (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch), each counter scaled by what it reads
    for the calibration kernel of known size in the same pass.  Returns
    ({phase: bytes}, {"fetch": f, "write": w}) or (None, reason)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    per = {}
    scale = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for counter, cal, cal_bytes in (("FETCH_SIZE", "hmv_read_kernel", CAL_READ_BYTES),
                                        ("WRITE_SIZE", "hmv_copy_kernel", CAL_COPY_BYTES)):
            d = os.path.join(td, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-pass", "--subdiv", str(subdiv)]
            log("traffic: rocprofv3 --pmc %s pass" % counter)
            # a pass takes about 10 s; a stuck profiler costs at most 2 minutes
            # and leaves the committed figure in place
            try:
                p = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
            except subprocess.TimeoutExpired:
                return None, "%s pass timed out" % counter
            path = os.path.join(d, "run_counter_collection.csv")
            if not os.path.exists(path):
                return None, "%s pass wrote no counters (exit %d)" % (counter, p.returncode)
            got = counter_bytes(path, counter, cal, cal_bytes)
            if isinstance(got, str):
                return None, "%s pass: %s" % (counter, got)
            scale[counter], pb = got
            for ph, b in pb.items():
                per[ph] = per.get(ph, 0.0) + b
    if any(v <= 0 for v in per.values()):
        return None, "no apply dispatch counted"
    return per, {"fetch": round(scale["FETCH_SIZE"], 4), "write": round(scale["WRITE_SIZE"], 4)}


def counter_bytes(path, counter, cal, cal_bytes):
    """From one rocprofv3 counter CSV: the scale that turns `counter` (KiB
    per dispatch) into bytes, from the calibration kernel `cal` of known
    traffic `cal_bytes`, and the scaled bytes per launch of each apply phase
    (the mean over dispatches of each kernel template, summed over the
    phase's templates).  Returns (scale, {phase: bytes}) or a reason."""
    vals = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    cal_kib = [v for n, vs in vals.items() if cal in n for v in vs]
    if not cal_kib or min(cal_kib) <= 0:
        return "no %s dispatch" % cal
    scale = cal_bytes / (1024.0 * float(np.median(cal_kib)))
    per = {ph: 1024.0 * scale * sum(float(np.mean(vs)) for n, vs in vals.items() if key in n)
           for ph, key in _APPLY_KERNELS.items()}
    return scale, per


FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


def cpu_threads():
    """Threads of this process's CPU share: OMP_NUM_THREADS when set (the
    host cores set aside for one GPU on a shared node), else the CPUs this
    process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cond_bound(nd):
    """cond * 2e-8, cond = 2 sqrt(N) + 1: the sphere single-layer operator's
    eigenvalues are R/(2l+1) over (L+1)^2 ~ N modes, so two solutions whose
    relative residuals are below 1e-8 may differ by up to this much."""
    return (2.0 * np.sqrt(nd) + 1.0) * 2e-8


def rel_err(u, u_ref):
    return float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref))


def median_solve(fn, runs=3, key=1):
    """One untimed call (graph capture, workspaces), then `runs` timed calls.
    Returns the call whose info (element `key` of the result) reports the
    median seconds; the solvers are deterministic, so every call returns the
    same iterates."""
    fn()
    res = [fn() for _ in range(runs)]
    res.sort(key=lambda r: r[key]["seconds"])
    return res[len(res) // 2]


def timed_median(fn, runs=20, warm=3):
    """Median wall time (s) of `runs` calls after `warm` untimed ones."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline_leg(O, hm, x, b, alg_bytes, log):
    """The reference CPU path restated by the oracle, on this host: the
    per-leaf GEMV apply (m_HACApK_solve.f90:203-253 with HACApK_MAGMA.c:114-199)
    at the CPU share's thread count and at 1 thread (the reference's own
    setting: OpenMP stubbed to one thread, sequential MKL), median of 20
    after 3 warm-ups; and its BiCGSTAB (m_HACApK_solve.f90:342-449) from 0
    to 1e-8 at the full thread count.  Returns (record, FP64 oracle
    solution)."""
    nthr = cpu_threads()
    t_all = timed_median(lambda: O.adot_lfmtx_hyp(hm, x, nthr))
    t_one = timed_median(lambda: O.adot_lfmtx_hyp(hm, x, 1))
    t0 = time.perf_counter()
    u_ref, nstp, hist = O.bicgstab(hm, b, None, mstep=3000, eps=1e-8, nthr=nthr)
    t_solve = time.perf_counter() - t0
    tr = rel_err(O.adot_lfmtx_hyp(hm, u_ref, nthr), b)
    rec = {"value": round(alg_bytes / t_all / 1e9, 2), "unit": "GB/s", "cores": nthr, "kind": "port",
           "sample": "median of 20 FP64 H-matvecs after 3 warm-ups of the same %d-unknown H-matrix "
                     "(oracle per-leaf GEMV over storage-balanced leaf ranges, %d threads); the same "
                     "at 1 thread; one BiCGSTAB to 1e-8" % (hm.nd, nthr),
           "ms_per_matvec": round(1e3 * t_all, 3),
           "one_thread": {"GBs": round(alg_bytes / t_one / 1e9, 2), "ms_per_matvec": round(1e3 * t_one, 3)},
           "bicgstab": {"seconds": round(t_solve, 3), "iterations": int(nstp),
                        "relres": float(hist[-1]), "true_relres": tr, "threads": nthr},
           "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
           "threads_note": "threads = this process's CPU share (OMP_NUM_THREADS or CPU affinity)"}
    log("CPU baseline (%s): %d threads %.1f GB/s (%.2f ms/matvec), 1 thread %.2f GB/s (%.1f ms); "
        "BiCGSTAB %d it in %.2f s, true relres %.2e"
        % (rec["cpu_model"], nthr, alg_bytes / t_all / 1e9, 1e3 * t_all, alg_bytes / t_one / 1e9,
           1e3 * t_one, nstp, t_solve, tr))
    return rec, u_ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--subdiv", type=int, default=6, help="icosphere subdivisions per sphere")
    ap.add_argument("--quick", action="store_true", help="matvec only (for profiler runs)")
    ap.add_argument("--hmatmat", action="store_true",
                    help="only the FP64 H-matmat (k = 8, 32, 128) on 327,680 unknowns (for profiler runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 327,680-unknown precision / multi-RHS study and the 48-instance batch")
    ap.add_argument("--construct", action="store_true",
                    help="only the device H-matrix fill and arena repack on 327,680 unknowns "
                         "(for profiler runs)")
    ap.add_argument("--applies", action="store_true",
                    help="only the single-vector applies on 327,680 unknowns in the four storage "
                         "precisions (for profiler runs)")
    ap.add_argument("--traffic-pass", action="store_true",
                    help="counter pass over the benchmark's applies and calibration kernels "
                         "(bench.py runs it under rocprofv3 to measure HBM traffic)")
    ap.add_argument("--no-pmc", action="store_true",
                    help="take roofline.traffic from the committed PMC summary instead of "
                         "measuring it with two rocprofv3 passes")
    args = ap.parse_args()
    if args.traffic_pass:
        import __graft_entry__ as G
        traffic_pass(G.load_package(), args.subdiv)
        return
    if args.applies:
        import __graft_entry__ as G
        print(json.dumps({"applies_327680": applies_only(G.load_package(), log)}), flush=True)
        return
    if args.hmatmat:
        import __graft_entry__ as G
        print(json.dumps({"hmatmat_fp64_327680": hmatmat_only(G.load_package(), log)}), flush=True)
        return
    if args.construct:
        import __graft_entry__ as G
        print(json.dumps({"construct_327680": construct_only(G.load_package(), log)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as g
    H = g.load_package()
    L = H.hmv_lib()
    if L.hmv_init(0) != 0:
        raise SystemExit("no usable GPU: " + L.hmv_last_error().decode())
    dev_name, ncu = H.device_info()
    log("device:", dev_name, "CUs", ncu)

    # ---- workload ---------------------------------------------------------
    t0 = time.time()
    mesh = H.merge(H.icosphere(args.subdiv), H.icosphere(args.subdiv, 1.0, (3.0, 0.0, 0.0)))
    hm = H.HMatrix(*mesh, H.baseline_params(aca_eps=1e-5))
    t_build = time.time() - t0
    nd = hm.nd
    tab = hm.leaf_table()
    lr, de = tab[tab[:, 0] == 1], tab[tab[:, 0] == 2]
    E_lr = int(((lr[:, 3] + lr[:, 5]) * lr[:, 1]).sum())
    E_d = int((de[:, 3] * de[:, 5]).sum())
    log("H-matrix: nd=%d leaves=%d (LR %d, dense %d) E=%d built in %.1fs"
        % (nd, hm.nlf, len(lr), len(de), E_lr + E_d, t_build))

    t0 = time.time()
    a64 = H.Arena(hm, "f64")
    t_arena = time.time() - t0
    st = a64.stats
    alg_bytes = 8 * (E_lr + E_d) + 16 * nd
    assert alg_bytes == st["alg_bytes"]
    log("arena: %.2f GB (payload %.2f GB), %d+%d workgroups, built in %.1fs"
        % (st["arena_bytes"] / 1e9, st["payload_bytes"] / 1e9, st["nwork_phase1"], st["nwork_phase2"], t_arena))

    x = H.uniform(1, nd)
    dx = H.DeviceBuffer.from_array(x)
    dy = H.DeviceBuffer(8 * nd)

    # ---- timed matvecs ----------------------------------------------------
    for _ in range(args.warmup):
        a64.apply(dx.ptr, dy.ptr, 0.0)
    L.hmv_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a64.apply(dx.ptr, dy.ptr, 0.0)
    L.hmv_device_synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = alg_bytes * args.steps / elapsed / 1e9
    log("H-matvec FP64: %.3f ms/step, %.1f GB/s algorithmic" % (ms_per_step, value))

    # per-phase HIP-event timing (events on the stream the kernels run on)
    ms1, ms2, msa = a64.apply_timed(dx.ptr, dy.ptr, max(20, args.steps))
    tslots = int(st["tslots"])
    bytes_p1 = 8 * int((lr[:, 5] * lr[:, 1]).sum()) + 8 * nd + 8 * tslots
    bytes_p2 = 8 * int((lr[:, 3] * lr[:, 1]).sum()) + 8 * E_d + 8 * tslots + 16 * nd
    phases = {"phase1": (ms1, bytes_p1, "hmv_stream_kernel<double, 1, F64, F64> (t = V^T x)"),
              "phase2": (ms2, bytes_p2, "hmv_stream_kernel<double, 2, F64, F64> (y = [A|U][x;t])")}
    dom = max(phases, key=lambda k: phases[k][0])
    dms, dbytes, dname = phases[dom]
    achieved = dbytes / (dms * 1e-3) / 1e9
    # HBM traffic per launch: measured by two counter passes of this workload
    # (or, with --no-pmc / --quick or when they fail, the committed summary)
    measured, how = (None, "not requested") if (args.quick or args.no_pmc) else \
        measure_traffic(args.subdiv, log)
    if measured:
        traffic = measured[dom]
        tsrc = ("measured in this run: rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over "
                "bench.py --traffic-pass (same workload), each scaled by the calibration "
                "kernels' counts (fetch x%.4g, write x%.4g)" % (how["fetch"], how["write"]))
        log("traffic: phase1 %.4g B, phase2 %.4g B per launch (%.3fx / %.3fx algorithmic)"
            % (measured["phase1"], measured["phase2"], measured["phase1"] / bytes_p1,
               measured["phase2"] / bytes_p2))
    else:
        traffic = load_traffic(dom)
        tsrc = ("profiles/pmc_traffic_two_spheres_163840.txt (rocprofv3 --pmc FETCH_SIZE and "
                "WRITE_SIZE of this kernel on this workload, scaled by the calibration kernels' "
                "counts; not measured in this run: %s)" % how)
    roofline = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": tsrc,
                "traffic_per_phase": ({k: round(v) for k, v in measured.items()} if measured else None),
                "alg_bytes_per_launch": dbytes, "avg_launch_ms": round(dms, 5),
                "phases_ms": {"phase1": round(ms1, 5), "phase2": round(ms2, 5), "apply": round(msa, 5)},
                "apply_event_GBs": round(alg_bytes / (msa * 1e-3) / 1e9, 1)}
    log("phases: p1 %.3f ms (%.1f GB/s), p2 %.3f ms (%.1f GB/s); dominant %s %.1f GB/s = %.1f%% of peak"
        % (ms1, bytes_p1 / ms1 / 1e6, ms2, bytes_p2 / ms2 / 1e6, dom, achieved, 100 * achieved / HBM_PEAK_GBS))

    # full-size parity of this very apply against the oracle (checker only)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    y = dy.download()
    nthr = cpu_threads()
    y_ref = O.adot_lfmtx_hyp(hm, x, nthr)
    parity = float(np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref))
    log("parity vs oracle: %.3e" % parity)

    # host-vector apply (what the reference-named zau += H*zu entry point does):
    # includes the x upload and y download over PCIe, into a caller-owned
    # host array as the reference's caller passes one; median of 30 calls
    # after 3; reported, never `value`
    yh = np.zeros(nd)
    th = []
    for i in range(33):
        t0 = time.perf_counter()
        a64.apply_host(x, yh)
        if i >= 3:
            th.append(1e3 * (time.perf_counter() - t0))
    ms_host = float(np.median(th))
    pcie = {"ms_per_matvec": round(ms_host, 4), "ms_min": round(min(th), 4),
            "alg_GBs": round(alg_bytes / ms_host / 1e6, 1), "calls": len(th)}
    log("host-vector apply (PCIe-inclusive): %.3f ms, %.1f GB/s algorithmic" % (ms_host, pcie["alg_GBs"]))

    # achievable streaming rates on this device (plain read / copy kernels)
    cal_bytes = 2 << 30
    ms_read = H.stream_test(0, cal_bytes, reps=10)
    ms_copy = H.stream_test(1, cal_bytes // 2, reps=10)
    calibration = {"read_GBs": round(cal_bytes / ms_read / 1e6, 1),
                   "copy_GBs": round(2 * (cal_bytes // 2) / ms_copy / 1e6, 1), "bytes": cal_bytes,
                   "mfma_f64_TFLOPs": round(H.mfma_rate(), 1)}
    log("calibration: plain read %.1f GB/s, plain copy %.1f GB/s, FP64 MFMA %.1f TFLOP/s"
        % (calibration["read_GBs"], calibration["copy_GBs"], calibration["mfma_f64_TFLOPs"]))

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (icosphere meshes generated on the host, x ~ U[-1,1] seed 1)",
        "config": {"workload": "two icospheres s=%d (r=1, centres 3 apart), %d unknowns, leaf 20, "
                               "eta 2, ACA+ 1e-5, 1/(4 pi r) centroid kernel" % (args.subdiv, nd),
                   "nd": nd, "leaves": hm.nlf, "lowrank_leaves": int(len(lr)), "dense_leaves": int(len(de)),
                   "entries": E_lr + E_d, "alg_bytes_per_matvec": alg_bytes,
                   "arena_bytes": int(st["arena_bytes"]), "device": dev_name},
        "roofline": roofline,
        "parity_rel_err_vs_oracle": parity, "host_vector_apply": pcie,
        "calibration": calibration,
        "construction_s": round(t_build, 2), "arena_build_s": round(t_arena, 2),
    }

    if not args.quick:
        # ---- CPU baseline: bounded sample on this host (and the oracle's
        # FP64 solution, against which every GPU solve's error is reported)
        b = H.uniform(1, nd)
        result["cpu_baseline"], u_oracle = cpu_baseline_leg(O, hm, x, b, alg_bytes, log)

        # ---- BiCGSTAB time to 1e-8: FP64 vs mixed precision -------------------
        solves = {"oracle_solution_bound": cond_bound(nd)}
        u64, i64, h64 = median_solve(lambda: a64.bicgstab(b, eps=1e-8, maxit=2000, check_every=8))
        solves["fp64"] = {"seconds": round(i64["seconds"], 4), "iterations": i64["nstp"],
                          "true_relres": i64["true_relres"], "status": i64["status_name"],
                          "rel_err_vs_oracle_solution": rel_err(u64, u_oracle)}
        log("BiCGSTAB FP64: %d it, %.3f s, true relres %.2e" % (i64["nstp"], i64["seconds"], i64["true_relres"]))
        for prec, ie in (("f32", 1e-4), ("bf16", 1e-2)):
            alow = H.Arena(hm, prec)
            u, inf, hist = median_solve(lambda: H.bicgstab_ir(a64, alow, b, eps=1e-8, inner_eps=ie,
                                                              max_outer=40, max_inner=2000))
            err = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
            solves["mixed_" + prec] = {
                "seconds": round(inf["seconds"], 4), "inner_eps": ie, "outer": inf["outer"],
                "inner_iterations": inf["inner_total"],
                "true_relres": inf["relres"], "status": inf["status_name"],
                "rel_err_vs_fp64_solution": err, "outer_residuals": [float(v) for v in hist],
                "time_ratio_vs_fp64": round(inf["seconds"] / i64["seconds"], 3),
                "rel_err_vs_oracle_solution": rel_err(u, u_oracle),
                "arena_bytes": int(alow.stats["arena_bytes"])}
            log("BiCGSTAB-IR %s: outer %d inner %d, %.3f s (%.2fx FP64), relres %.2e, err vs FP64 sol %.2e"
                % (prec, inf["outer"], inf["inner_total"], inf["seconds"], inf["seconds"] / i64["seconds"],
                   inf["relres"], err))
            if prec == "bf16":
                # refinement with GCR(32) inner solves on the same BF16 arena
                u, inf, hist = median_solve(lambda: H.gcr_ir(a64, alow, b, eps=1e-8, inner_eps=5e-3, m=32))
                err = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
                solves["gcr32_ir_bf16"] = {
                    "seconds": round(inf["seconds"], 4), "inner_eps": 5e-3, "outer": inf["outer"],
                    "inner_products": inf["inner_total"], "true_relres": inf["relres"],
                    "status": inf["status_name"], "rel_err_vs_fp64_solution": err,
                    "outer_residuals": [float(v) for v in hist],
                    "time_ratio_vs_fp64": round(inf["seconds"] / i64["seconds"], 3),
                    "rel_err_vs_oracle_solution": rel_err(u, u_oracle)}
                log("GCR(32)-IR bf16: outer %d, %d inner products, %.3f s (%.2fx FP64), relres %.2e, "
                    "err vs FP64 sol %.2e" % (inf["outer"], inf["inner_total"], inf["seconds"],
                                              inf["seconds"] / i64["seconds"], inf["relres"], err))
            alow.free()
        # refinement with GCR(32) inner solves on an all-BF16 arena
        alow = H.Arena(hm, "bf16_all")
        u, inf, hist = median_solve(lambda: H.gcr_ir(a64, alow, b, eps=1e-8, inner_eps=5e-3, m=32))
        err = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
        solves["gcr32_ir_bf16_all"] = {
            "seconds": round(inf["seconds"], 4), "inner_eps": 5e-3, "outer": inf["outer"],
            "inner_products": inf["inner_total"], "true_relres": inf["relres"],
            "status": inf["status_name"], "rel_err_vs_fp64_solution": err,
            "outer_residuals": [float(v) for v in hist],
            "time_ratio_vs_fp64": round(inf["seconds"] / i64["seconds"], 3),
            "rel_err_vs_oracle_solution": rel_err(u, u_oracle),
            "arena_bytes": int(alow.stats["arena_bytes"])}
        log("GCR(32)-IR bf16_all: outer %d, %d inner products, %.3f s (%.2fx FP64), relres %.2e, "
            "err vs FP64 sol %.2e" % (inf["outer"], inf["inner_total"], inf["seconds"],
                                      inf["seconds"] / i64["seconds"], inf["relres"], err))
        alow.free()
        solves["cpu_time_ratio_vs_fp64_gpu"] = round(
            result["cpu_baseline"]["bicgstab"]["seconds"] / i64["seconds"], 1)
        result["bicgstab"] = solves

        # ---- leaf histograms ---------------------------------------------------
        result["histograms"] = {
            "lowrank_rank": np.bincount(lr[:, 1]).tolist(),
            "leaf_rows_log2_bins": {"lowrank": np.bincount(np.log2(lr[:, 3]).astype(int)).tolist(),
                                    "dense": np.bincount(np.log2(de[:, 3]).astype(int)).tolist()},
        }
    else:
        result["cpu_baseline"] = None

    if not args.quick and not args.no_extras:
        result["sphere_327680"] = study_327680(H, O, log)
        result["instances_48"] = study_instances(H, O, log)

    print(json.dumps(result), flush=True)


def study_327680(H, O, log):
    """One icosphere s=7 (327,680 unknowns), ACA+ 1e-5: H-matvec rate and
    error per storage precision, BiCGSTAB to 1e-8 in FP64 vs FP64 iterative
    refinement with FP32 / BF16 inner solves (RHS = 1 and U[-1,1] seed 2,
    errors against the oracle's FP64 solution), the multi-RHS H-matmat and
    block BiCGSTAB in FP64 and with FP32 factors refined in FP64 (k RHS
    U[-1,1] seed 3)."""
    out = {}
    nthr = cpu_threads()
    t0 = time.perf_counter()
    hm = H.HMatrix(*H.icosphere(7), H.baseline_params(aca_eps=1e-5))
    nd = hm.nd
    out["construction_s"] = round(time.perf_counter() - t0, 2)
    arenas = {p: H.Arena(hm, p) for p in ("f64", "f32", "bf16", "bf16_all")}
    xh = H.uniform(1, nd)
    x = H.DeviceBuffer.from_array(xh)
    y = H.DeviceBuffer.from_array(np.zeros(nd))
    y_ref = O.adot_lfmtx_hyp(hm, xh, nthr)
    rates = {}
    for p, a in arenas.items():
        m1, m2, ma = a.apply_timed(x.ptr, y.ptr, 30)
        err = rel_err(y.download(), y_ref)
        rates[p] = {"ms": round(ma, 4), "storage_alg_GBs": round(a.stats["alg_bytes"] / ma / 1e6, 1),
                    "phase1_ms": round(m1, 4), "phase2_ms": round(m2, 4),
                    "arena_GB": round(a.stats["arena_bytes"] / 1e9, 3), "tuning": a.tuning,
                    "rel_err_vs_fp64_oracle": err}
        log("s=7 %s apply %.4f ms (p1 %.4f, p2 %.4f), %.1f GB/s at storage precision, rel err vs FP64 "
            "oracle %.2e" % (p, ma, m1, m2, rates[p]["storage_alg_GBs"], err))
    out["apply"] = rates
    solves = {"oracle_solution_bound": cond_bound(nd)}
    inner = {"f32": 1e-4, "bf16": 1e-2, "bf16_all": 1e-2}
    for rname, b in (("ones", np.ones(nd)), ("uniform_seed2", H.uniform(2, nd))):
        t_or = time.perf_counter()
        u_oracle, n_or, _ = O.bicgstab(hm, b, None, mstep=3000, eps=1e-8, nthr=nthr)
        t_or = time.perf_counter() - t_or
        a64 = arenas["f64"]
        u64, i64, _ = median_solve(lambda: a64.bicgstab(b, eps=1e-8, maxit=3000, check_every=8))
        row = {"oracle_cpu": {"seconds": round(t_or, 3), "iterations": int(n_or), "threads": nthr},
               "fp64": {"seconds": round(i64["seconds"], 4), "iterations": i64["nstp"],
                        "true_relres": i64["true_relres"],
                        "rel_err_vs_oracle_solution": rel_err(u64, u_oracle)}}
        ug, ig, _ = median_solve(lambda: a64.gcrm(b, eps=1e-8, maxit=3000, mreset=8, check_every=8,
                                                  history=False))
        row["gcrm8_fp64"] = {"seconds": round(ig["seconds"], 4), "iterations": ig["nstp"],
                             "true_relres": ig["true_relres"], "status": ig["status_name"],
                             "rel_err_vs_oracle_solution": rel_err(ug, u_oracle)}
        log("s=7 %s: oracle BiCGSTAB %d it in %.2f s (%d threads); FP64 BiCGSTAB %d it %.4f s; GCR(8) "
            "FP64 %.4f s (%d it)" % (rname, n_or, t_or, nthr, i64["nstp"], i64["seconds"],
                                     ig["seconds"], ig["nstp"]))
        up, ip, _ = median_solve(lambda: a64.bicgstab_pipe(b, eps=1e-8, maxit=3000, check_every=8,
                                                           history=False))
        row["pipelined_bicgstab_fp64"] = {"seconds": round(ip["seconds"], 4), "iterations": ip["nstp"],
                                          "true_relres": ip["true_relres"], "status": ip["status_name"],
                                          "rel_err_vs_oracle_solution": rel_err(up, u_oracle)}
        for p in ("f32", "bf16", "bf16_all"):
            u, inf, hist = median_solve(lambda: H.bicgstab_ir(a64, arenas[p], b, eps=1e-8,
                                                              inner_eps=inner[p], max_outer=60,
                                                              max_inner=3000))
            row["mixed_" + p] = {
                "seconds": round(inf["seconds"], 4), "inner_eps": inner[p], "outer": inf["outer"],
                "inner_iterations": inf["inner_total"], "true_relres": inf["relres"],
                "status": inf["status_name"], "time_ratio_vs_fp64": round(inf["seconds"] / i64["seconds"], 3),
                "rel_err_vs_oracle_solution": rel_err(u, u_oracle),
                "rel_err_vs_fp64_gpu_solution": rel_err(u, u64),
                "outer_residuals": [float(v) for v in hist]}
            log("s=7 %s: FP64 %.4f s vs %s BiCGSTAB-IR (inner %.0e) %.4f s (%.2fx), relres %.2e, err vs "
                "oracle %.1e" % (rname, i64["seconds"], p, inner[p], inf["seconds"],
                                 inf["seconds"] / i64["seconds"], inf["relres"], rel_err(u, u_oracle)))
            u, inf, hist = median_solve(lambda: H.gcr_ir(a64, arenas[p], b, eps=1e-8, inner_eps=5e-3,
                                                        m=32))
            row["gcr32_ir_" + p] = {
                "seconds": round(inf["seconds"], 4), "inner_eps": 5e-3, "outer": inf["outer"],
                "inner_products": inf["inner_total"], "true_relres": inf["relres"],
                "status": inf["status_name"], "time_ratio_vs_fp64": round(inf["seconds"] / i64["seconds"], 3),
                "rel_err_vs_oracle_solution": rel_err(u, u_oracle),
                "rel_err_vs_fp64_gpu_solution": rel_err(u, u64),
                "outer_residuals": [float(v) for v in hist]}
            log("s=7 %s: FP64 %.4f s vs %s GCR(32)-IR %.4f s (%.2fx), %d outer, %d inner products, "
                "relres %.2e, err vs oracle %.1e"
                % (rname, i64["seconds"], p, inf["seconds"], inf["seconds"] / i64["seconds"],
                   inf["outer"], inf["inner_total"], inf["relres"], rel_err(u, u_oracle)))
        solves[rname] = row
    out["bicgstab"] = solves
    # multi-RHS H-matmat
    a64 = arenas["f64"]
    st = a64.stats
    E = st["entries_lr"] + st["entries_dense"]
    single_ms = rates["f64"]["ms"]
    # measured rate of the 16x16x4 FP64 MFMA the kernel issues, on register
    # operands (the 4-block 4x4x4 FP64 MFMA reaches 69-75 TFLOP/s there)
    mfma_tf = H.mfma_rate()
    mm = {"mfma_f64_TFLOPs_measured": round(mfma_tf, 1), "mfma_measured_instruction": "16x16x4 f64",
          "fp64_peak_TFLOPs": FP64_PEAK_TFLOPS}
    rng = np.random.default_rng(3)
    for k in (8, 32, 128):
        X = H.DeviceBuffer.from_array(rng.uniform(-1, 1, nd * k))
        Y = H.DeviceBuffer.from_array(np.zeros(nd * k))
        # median of 5 rounds of 5 products (each round after one untimed)
        ms = float(np.median([a64.apply_block_timed(X.ptr, Y.ptr, k, 5) for _ in range(5)]))
        tf = 2.0 * E * k / ms / 1e9
        hbm_ms = (8.0 * E + 16.0 * nd * k) / (HBM_PEAK_GBS * 1e6)
        mfma_ms = 2.0 * E * k / (FP64_PEAK_TFLOPS * 1e9)
        floor = max(hbm_ms, mfma_ms)
        mm[str(k)] = {"ms": round(ms, 4), "ms_per_rhs": round(ms / k, 4),
                      "speedup_per_rhs_vs_single": round(single_ms * k / ms, 2),
                      "TFLOPs": round(tf, 2), "frac_of_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 3),
                      "frac_of_measured_mfma": round(tf / mfma_tf, 3),
                      "floor_ms": round(floor, 4), "bound": "hbm" if hbm_ms >= mfma_ms else "mfma",
                      "roofline_frac": round(floor / ms, 3)}
        log("s=7 H-matmat k=%d: %.3f ms (%.3f ms/RHS, %.1fx single), %.1f TFLOP/s = %.0f%% of the FP64 "
            "peak; floor %.3f ms (%s) -> %.0f%% of roofline"
            % (k, ms, ms / k, single_ms * k / ms, tf, 100 * tf / FP64_PEAK_TFLOPS, floor,
               mm[str(k)]["bound"], 100 * floor / ms))
        # the path autotuner (vector units vs MFMA, MFMA tiles per unit); the
        # block solves below run with what it picks
        dflt = a64.block_tuning(k)
        best, best_ms, dflt_ms = a64.autotune_block(k, reps=5)
        tuned_ms = a64.apply_block_timed(X.ptr, Y.ptr, k, 5)
        mm[str(k)]["autotune"] = {"default_path": dflt, "best_path": best,
                                  "default_ms": round(dflt_ms, 4), "best_ms": round(best_ms, 4),
                                  "tuned_ms": round(tuned_ms, 4),
                                  "tuned_roofline_frac": round(floor / tuned_ms, 3)}
        log("s=7 H-matmat k=%d autotuned: %s -> %s, %.3f ms (default path %.3f ms)"
            % (k, dflt, best, tuned_ms, dflt_ms))
        X.free()
        Y.free()
    out["hmatmat_fp64"] = mm
    # block BiCGSTAB: FP64, and FP32 factors refined in FP64, k = 8, 32, 128
    single_rate = 1.0 / solves["uniform_seed2"]["fp64"]["seconds"]
    blk = {"single_rhs_solves_per_s": round(single_rate, 2)}
    for k in (8, 32, 128):
        B = rng.uniform(-1, 1, (nd, k))
        a64.block_bicgstab(B, eps=1e-8, maxit=3000, check_every=8)
        U, cols, inf = a64.block_bicgstab(B, eps=1e-8, maxit=3000, check_every=8)
        blk["fp64_k%d" % k] = {
            "seconds": round(inf["seconds"], 4), "iterations": inf["nstp"], "status": inf["status_name"],
            "worst_true_relres": inf["true_relres"], "rhs_solves_per_s": round(k / inf["seconds"], 2)}
        log("s=7 block BiCGSTAB FP64 k=%d: %.3f s, %.1f RHS-solves/s (single: %.1f)"
            % (k, inf["seconds"], k / inf["seconds"], single_rate))
        H.block_bicgstab_ir(a64, arenas["f32"], B, eps=1e-8, inner_eps=1e-4, max_outer=40, max_inner=3000)
        U2, cols2, inf2 = H.block_bicgstab_ir(a64, arenas["f32"], B, eps=1e-8, inner_eps=1e-4,
                                              max_outer=40, max_inner=3000)
        errs = [rel_err(U2[:, j], U[:, j]) for j in (0, k - 1)]
        blk["f32_ir_k%d" % k] = {
            "seconds": round(inf2["seconds"], 4), "outer": inf2["outer"],
            "inner_iterations": inf2["inner_total"], "status": inf2["status_name"],
            "worst_true_relres": inf2["true_relres"], "rhs_solves_per_s": round(k / inf2["seconds"], 2),
            "rel_err_vs_fp64_block_solution_sampled": max(errs)}
        log("s=7 block BiCGSTAB FP32 factors + FP64 refinement k=%d: %.3f s (%d outer), %.1f RHS-solves/s,"
            " worst relres %.2e" % (k, inf2["seconds"], inf2["outer"], k / inf2["seconds"],
                                    inf2["true_relres"]))
    out["block_bicgstab"] = blk
    # the same H-matrix filled on the device, and its arena repacked there
    mesh = H.icosphere(7)
    t0 = time.perf_counter()
    dh = H.DeviceHMatrix(*mesh, H.baseline_params(aca_eps=1e-5))
    t1 = time.perf_counter()
    ad = dh.arena("f64")
    t2 = time.perf_counter()
    yd, yh = ad.matvec(xh), a64.matvec(xh)
    out["device_construction"] = {
        "blocks_host_s": round(dh.block_seconds, 3), "fill_device_s": round(dh.fill_seconds, 3),
        "construct_total_s": round(t1 - t0, 3), "arena_from_device_factors_s": round(t2 - t1, 3),
        "host_construction_s": out["construction_s"],
        "host_arena_build_s": round(a64.stats["build_seconds"], 3),
        "ranks_equal_host": bool(np.array_equal(dh.table, hm.leaf_table())),
        "apply_rel_diff_vs_host_arena": float(np.linalg.norm(yd - yh) / np.linalg.norm(yh))}
    log("s=7 device construction: blocks %.3f s (host), fill %.3f s (device), arena %.3f s; host: "
        "construction %.3f s, arena %.3f s; ranks equal host: %s, apply diff %.1e"
        % (dh.block_seconds, dh.fill_seconds, t2 - t1, out["construction_s"],
           a64.stats["build_seconds"], out["device_construction"]["ranks_equal_host"],
           out["device_construction"]["apply_rel_diff_vs_host_arena"]))
    ad.free()
    dh.free()
    for a in arenas.values():
        a.free()
    return out


def hmatmat_only(H, log, reps=5):
    """FP64 H-matmat timings (k = 8, 32, 128) on one icosphere s=7."""
    hm = H.HMatrix(*H.icosphere(7), H.baseline_params(aca_eps=1e-5))
    a = H.Arena(hm, "f64")
    st = a.stats
    E = st["entries_lr"] + st["entries_dense"]
    rng = np.random.default_rng(3)
    mfma_tf = H.mfma_rate()
    log("measured FP64 MFMA rate: %.1f TFLOP/s" % mfma_tf)
    out = {"mfma_f64_TFLOPs_measured": round(mfma_tf, 1)}
    for k in (8, 32, 128):
        X = H.DeviceBuffer.from_array(rng.uniform(-1, 1, hm.nd * k))
        Y = H.DeviceBuffer.from_array(np.zeros(hm.nd * k))
        ms = a.apply_block_timed(X.ptr, Y.ptr, k, reps)
        out[str(k)] = {"ms": round(ms, 4), "TFLOPs": round(2.0 * E * k / ms / 1e9, 2)}
        log("H-matmat k=%d: %.3f ms, %.1f TFLOP/s" % (k, ms, 2.0 * E * k / ms / 1e9))
        X.free()
        Y.free()
    a.free()
    # narrow-storage arenas (FP32 MFMA), k = 32
    for prec in ("f32", "bf16_all"):
        a = H.Arena(hm, prec)
        X = H.DeviceBuffer.from_array(rng.uniform(-1, 1, hm.nd * 32))
        Y = H.DeviceBuffer.from_array(np.zeros(hm.nd * 32))
        ms = a.apply_block_timed(X.ptr, Y.ptr, 32, reps)
        out[prec + "_32"] = {"ms": round(ms, 4)}
        log("H-matmat %s k=32: %.3f ms" % (prec, ms))
        X.free()
        Y.free()
        a.free()
    return out


def applies_only(H, log, reps=5):
    """Single-vector applies of one icosphere s=7 in the four storage
    precisions, one untimed and `reps` timed applies each."""
    if H.hmv_lib().hmv_init(0) != 0:
        raise SystemExit("no GPU: " + H.hmv_lib().hmv_last_error().decode())
    hm = H.HMatrix(*H.icosphere(7), H.baseline_params(aca_eps=1e-5))
    dx = H.DeviceBuffer.from_array(H.uniform(1, hm.nd))
    dy = H.DeviceBuffer(8 * hm.nd)
    out = {}
    for prec in ("f64", "f32", "bf16", "bf16_all"):
        a = H.Arena(hm, prec)
        a.apply(dx.ptr, dy.ptr)
        m1, m2, ma = a.apply_timed(dx.ptr, dy.ptr, reps)
        out[prec] = {"ms": round(ma, 4), "phase1_ms": round(m1, 4), "phase2_ms": round(m2, 4),
                     "storage_alg_GBs": round(a.stats["alg_bytes"] / ma / 1e6, 1)}
        log("%s apply %.4f ms (p1 %.4f, p2 %.4f)" % (prec, ma, m1, m2))
        a.free()
    dx.free()
    dy.free()
    return out


def construct_only(H, log, reps=2):
    """Device fill (hmv_construct) and arena repack of one icosphere s=7,
    ACA+ 1e-5, centroid kernel, `reps` times."""
    if H.hmv_lib().hmv_init(0) != 0:
        raise SystemExit("no GPU: " + H.hmv_lib().hmv_last_error().decode())
    mesh = H.icosphere(7)
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        dh = H.DeviceHMatrix(*mesh, H.baseline_params(aca_eps=1e-5))
        t1 = time.perf_counter()
        a = dh.arena("f64")
        t2 = time.perf_counter()
        out.append({"blocks_host_s": round(dh.block_seconds, 3), "fill_device_s": round(dh.fill_seconds, 3),
                    "arena_s": round(t2 - t1, 3), "total_s": round(t2 - t0, 3)})
        log("construct: blocks %.3f s (host), fill %.3f s (device), arena %.3f s"
            % (dh.block_seconds, dh.fill_seconds, t2 - t1))
        a.free()
        dh.free()
    return out


def study_instances(H, O, log):
    """48 icospheres with s in {4,5,6} and radii U[0.5,2] (seed 4), ACA+ 1e-4,
    RHS = 1: one FP64 arena with one BiCGSTAB per instance to 1e-8; and the
    mixed-precision configurations: BF16-factor arenas (the all-BF16 one
    with the tuning the autotuner picks over the whole batch) with
    per-instance FP64 refinement, BiCGSTAB (hmv_multi_bicgstab_ir) or GCR(32)
    (hmv_multi_gcr_ir) inner solves.  Errors of sampled instances (one of
    each size) against the oracle's FP64 solutions."""
    nthr = cpu_threads()
    rng = np.random.default_rng(4)
    subs = rng.choice([4, 5, 6], size=48)
    radii = rng.uniform(0.5, 2.0, size=48)
    t0 = time.perf_counter()
    hms = [H.HMatrix(*H.icosphere(int(s), float(r)), H.baseline_params(aca_eps=1e-4))
           for s, r in zip(subs, radii)]
    tcons = time.perf_counter() - t0
    m = H.MultiArena(hms, "f64")
    x = H.DeviceBuffer.from_array(H.uniform(1, m.nd))
    y = H.DeviceBuffer.from_array(np.zeros(m.nd))
    m1, m2, ma = m.apply_timed(x.ptr, y.ptr, 20)
    b = np.ones(m.nd)
    m.bicgstab_multi(b, eps=1e-8, maxit=3000, check_every=4)
    u, inst, secs, inf = m.bicgstab_multi(b, eps=1e-8, maxit=3000, check_every=4)
    steps = [d["nstp"] for d in inst]
    sample = sorted({int(np.flatnonzero(subs == s)[0]) for s in (4, 5, 6) if (subs == s).any()})
    oracle = {}
    for i in sample:
        lo, hi = m.offsets[i], m.offsets[i + 1]
        oracle[i] = O.bicgstab(hms[i], b[lo:hi], None, mstep=3000, eps=1e-8, nthr=nthr)[0]

    def errs(uu):
        return {str(i): rel_err(uu[m.offsets[i]:m.offsets[i + 1]], oracle[i]) for i in sample}

    out = {"instances": 48, "unknowns": int(m.nd),
           "sizes": {"5120": int((subs == 4).sum()), "20480": int((subs == 5).sum()),
                     "81920": int((subs == 6).sum())},
           "construction_s": round(tcons, 2), "arena_GB": round(m.stats["arena_bytes"] / 1e9, 2),
           "apply_all_ms": round(ma, 4), "apply_alg_GBs": round(m.stats["alg_bytes"] / ma / 1e6, 1),
           "fp64": {"solve_seconds": round(inf["seconds"], 4), "status": inf["status_name"],
                    "worst_true_relres": inf["true_relres"], "steps_min": min(steps),
                    "steps_max": max(steps), "unknowns_per_s": round(m.nd / inf["seconds"], 1),
                    "instance_seconds_p50": round(float(np.percentile(secs, 50)), 4),
                    "instance_seconds_p99": round(float(np.percentile(secs, 99)), 4),
                    "rel_err_vs_oracle_solution_sampled": errs(u)}}
    log("48 instances (%d unknowns): one apply %.3f ms; FP64 solve %.4f s, %.3e unknowns/s, p50 %.4f s "
        "p99 %.4f s" % (m.nd, ma, inf["seconds"], m.nd / inf["seconds"],
                        out["fp64"]["instance_seconds_p50"], out["fp64"]["instance_seconds_p99"]))
    # mixed precision: BF16 factors (all-BF16 arena autotuned over the whole
    # batch; BF16 low-rank + FP32 dense with its defaults), per-instance FP64
    # refinement with inner BiCGSTAB to 0.1 (the fastest of 1e-1 .. 1e-4
    # measured on this batch)
    for prec, tune in (("bf16_all", True), ("bf16", False)):
        t0 = time.perf_counter()
        if tune:
            best, best_ms, dflt_ms = H.autotune_multi(hms, prec, reps=5)
        else:
            best, best_ms, dflt_ms = None, None, None
        t_tune = time.perf_counter() - t0
        mlow = H.MultiArena(hms, prec, tuning=best)
        H.multi_bicgstab_ir(m, mlow, b, eps=1e-8, inner_eps=0.1, max_outer=60, max_inner=3000)
        u2, inst2, secs2, inf2 = H.multi_bicgstab_ir(m, mlow, b, eps=1e-8, inner_eps=0.1,
                                                     max_outer=60, max_inner=3000)
        rec = {
            "tuning": mlow.tuning, "autotuned": tune,
            "arena_GB": round(mlow.stats["arena_bytes"] / 1e9, 2),
            "inner_eps": 0.1, "outer": inf2["outer"], "inner_iterations": inf2["inner_total"],
            "solve_seconds": round(inf2["seconds"], 4), "status": inf2["status_name"],
            "worst_true_relres": inf2["true_relres"], "unknowns_per_s": round(m.nd / inf2["seconds"], 1),
            "instance_seconds_p50": round(float(np.percentile(secs2, 50)), 4),
            "instance_seconds_p99": round(float(np.percentile(secs2, 99)), 4),
            "time_ratio_vs_fp64": round(inf2["seconds"] / inf["seconds"], 3),
            "rel_err_vs_oracle_solution_sampled": errs(u2)}
        if tune:
            rec.update({"autotune_s": round(t_tune, 1), "apply_ms_tuned": round(best_ms, 4),
                        "apply_ms_default": round(dflt_ms, 4)})
        out["mixed_" + prec] = rec
        log("48 instances %s%s + FP64 refinement: %.4f s (%.2fx FP64), %d outer, %.3e unknowns/s, p50 %.4f s "
            "p99 %.4f s, worst relres %.2e" % (prec, " (autotuned %s)" % best if tune else "",
                                               inf2["seconds"], inf2["seconds"] / inf["seconds"], inf2["outer"],
                                               m.nd / inf2["seconds"], rec["instance_seconds_p50"],
                                               rec["instance_seconds_p99"], inf2["true_relres"]))
        # the same refinement with GCR(32) inner solves to 5e-3 per instance
        # (hmv_multi_gcr_ir; the single-instance defaults, also best here)
        u3, inst3, secs3, inf3 = median_solve(
            lambda: H.multi_gcr_ir(m, mlow, b, eps=1e-8, inner_eps=5e-3, m=32, max_outer=60,
                                   max_inner=3000), key=3)
        out["gcr32_ir_" + prec] = {
            "tuning": mlow.tuning, "autotuned": tune, "inner_eps": 5e-3, "m": 32,
            "outer": inf3["outer"], "inner_steps": inf3["inner_total"],
            "solve_seconds": round(inf3["seconds"], 4), "status": inf3["status_name"],
            "worst_true_relres": inf3["true_relres"], "unknowns_per_s": round(m.nd / inf3["seconds"], 1),
            "instance_seconds_p50": round(float(np.percentile(secs3, 50)), 4),
            "instance_seconds_p99": round(float(np.percentile(secs3, 99)), 4),
            "time_ratio_vs_fp64": round(inf3["seconds"] / inf["seconds"], 3),
            "rel_err_vs_oracle_solution_sampled": errs(u3)}
        log("48 instances %s + FP64 refinement, GCR(32) inner: %.4f s (%.2fx FP64), %d outer, %.3e "
            "unknowns/s, p50 %.4f s p99 %.4f s, worst relres %.2e"
            % (prec, inf3["seconds"], inf3["seconds"] / inf["seconds"], inf3["outer"],
               m.nd / inf3["seconds"], out["gcr32_ir_" + prec]["instance_seconds_p50"],
               out["gcr32_ir_" + prec]["instance_seconds_p99"], inf3["true_relres"]))
        mlow.free()
    m.free()
    return out


if __name__ == "__main__":
    main()
