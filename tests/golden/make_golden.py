"""Writes the golden fixtures under tests/golden/ from the reference's data files.

Runs only where the reference tree exists at /root/reference; the tests read
the committed outputs.  Nothing from the reference's source code is stored: only its mesh
data files (converted to .npz) and the figures its recorded run printed.

  input_sample_mesh.npz  input_sample.pbf            (304 nodes, 600 triangles)
  input_1ts_mesh.npz     bem-bb_inputs/input_1ts.pbs  (502 nodes, 1,000 triangles)
  input_10ts_mesh.npz    bem-bb_inputs/input_10ts.pbs (5,202 nodes, 10,400 triangles)
  recorded_run_10ts.json figures of ivb/job_cpu.sh.o78632 and ivb/log0000.txt
  input_sample_pbf.json  size and SHA-256 of input_sample.pbf with its CRLF line
                         ends turned into LF (what a .pbf writer must reproduce)
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"


def main():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as g
    H = g.load_package()
    meshes = {
        "input_sample_mesh.npz": "input_sample.pbf",
        "input_1ts_mesh.npz": "bem-bb_inputs/input_1ts.pbs",
        "input_10ts_mesh.npz": "bem-bb_inputs/input_10ts.pbs",
    }
    for out, src in meshes.items():
        nodes, faces, para = H.read_pbf(os.path.join(REF, src))
        np.savez_compressed(os.path.join(HERE, out), nodes=nodes, faces=faces, dble_para=para)
        print("wrote", out, nodes.shape, faces.shape)

    raw = open(os.path.join(REF, "input_sample.pbf"), "rb").read().replace(b"\r\n", b"\n")
    with open(os.path.join(HERE, "input_sample_pbf.json"), "w") as f:
        json.dump({"source": "input_sample.pbf, line ends CRLF -> LF", "bytes": len(raw),
                   "sha256": hashlib.sha256(raw).hexdigest()}, f, indent=1)
    print("wrote input_sample_pbf.json")

    log = open(os.path.join(REF, "ivb", "job_cpu.sh.o78632"), encoding="utf-8", errors="replace").read()
    log0 = open(os.path.join(REF, "ivb", "log0000.txt"), encoding="utf-8", errors="replace").read()
    rec = {"source": "ivb/job_cpu.sh.o78632, ivb/log0000.txt",
           "mesh": "bem-bb_inputs/input_10ts.pbs", "tolerance": 1e-6, "leaf_size": 15}
    rec["nclst"] = int(re.search(r"No\. of cluster=\s*(\d+)", log).group(1))
    m = re.search(r"No\. of nsmtx\s+(\d+)\s+(\d+)\s+(\d+)", log)
    rec["nlfkt"], rec["nlfdense"], rec["nhblocks"] = (int(m.group(i)) for i in (1, 2, 3))
    rec["ncpc"] = int(re.search(r"ncpc=\s*(\d+)", log).group(1))
    rec["ktmax"] = int(re.search(r"ktmax=\s*(\d+)", log0).group(1))
    rec["mem_mbyte"] = float(re.search(r"Memory of the H-matrix=\s*([0-9.Ee+-]+)", log).group(1))
    # first solve (Fortran BiCGSTAB): iteration, time, log10(||r||/||b||)
    block = log.split("HACApK_bicgstab_lfmtx_hyp start", 1)[1].split("time_HACApK_solve", 1)[0]
    hist = []
    for line in block.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0].isdigit():
            hist.append(float(parts[2]))
    rec["log10_relres_history"] = hist
    with open(os.path.join(HERE, "recorded_run_10ts.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote recorded_run_10ts.json", rec)


if __name__ == "__main__":
    main()
