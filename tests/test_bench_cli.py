"""bench.py's command line: the driver's contract for runs on more than one
GPU (this build measures one GPU; it must refuse N > 1 with a message and a
non-zero status, before touching a device)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_cli_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_counter_bytes_scales_by_the_calibration_kernel(tmp_path):
    """roofline.traffic's arithmetic: a counter in KiB per dispatch is scaled
    by what it reads for the calibration kernel of known traffic, averaged
    per kernel template and summed over a phase's templates; other kernels
    (the k-RHS vector kernel's names contain "stream_kernel" too) are not
    counted.  Figures: FETCH_SIZE of the FP64 apply kernels on two spheres
    of 163,840 unknowns (hmv_read_kernel over 2 GiB reads 1,048,588 KiB)."""
    B = _bench_module()
    head = "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
    rows = []
    for i in range(2):
        rows.append((i, "void (anonymous namespace)::hmv_read_kernel(double2 const*, unsigned long, double*)", 1048588))
    for i in range(3):
        rows.append((10 + i, "void (anonymous namespace)::hmv_stream_kernel<double, 2, 0, 0, 256>(char const*)", 700399 + 2 * (i - 1)))
        rows.append((20 + i, "void (anonymous namespace)::hmv_stream_kernel<double, 1, 0, 0, 256>(char const*)", 472000))
        rows.append((30 + i, "void (anonymous namespace)::hmv_stream_kernel<double, 1, 0, 0, 64>(char const*)", 776))
        rows.append((40 + i, "void (anonymous namespace)::hmv_kstream_kernel<double, 2, 0, 0, 256, 8, true>(char const*)", 9e9))
        rows.append((50 + i, "void (anonymous namespace)::hmv_stream_kernel<double, 2, 0, 0, 256>(char const*)", -1))
    p = tmp_path / "run_counter_collection.csv"
    with open(p, "w") as f:
        f.write(head)
        for i, (d, n, v) in enumerate(rows):
            cn = "WRITE_SIZE" if d >= 50 else "FETCH_SIZE"  # another counter's rows are ignored
            f.write('%d,"%s",%s,%r\n' % (d, n, cn, float(v)))
    scale, per = B.counter_bytes(str(p), "FETCH_SIZE", "hmv_read_kernel", 2 << 30)
    assert abs(scale - (2 << 30) / (1024.0 * 1048588)) < 1e-12
    assert abs(per["phase2"] - 700399 * 1024 * scale) < 1e-3 * 1024
    assert abs(per["phase1"] - (472000 + 776) * 1024 * scale) < 1e-3 * 1024
    # 2.0 scale within 1e-5: the committed summary's factor
    assert abs(scale - 2.0) < 1e-4
    assert isinstance(B.counter_bytes(str(p), "FETCH_SIZE", "hmv_copy_kernel", 1 << 30), str)
