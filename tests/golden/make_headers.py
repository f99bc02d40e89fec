#!/usr/bin/env python3
"""Collect the text headers the reference writes for each golden case into
tests/golden/headers.npz (data only): for case <name>, the keys
<name>:SeisHeader_sem2d.hdr, <name>:FltXX_sem2d.hdr and <name>:FltXX_init_sem2d.tab
hold the files' text as written by the reference's list-directed WRITE statements
(REC_init, SRC/receivers.f90:214-221; BC_DYNFLT_init, SRC/bc_dynflt.f90:446-500).

Runs oracle/_ref/ref_harness (the reference built from /root/reference by
oracle/build_ref.py) on every case directory under tests/golden/cases/.

Usage: python tests/golden/make_headers.py   (needs /root/reference)
"""
import glob
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "headers.npz")


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    texts = {}
    for src in sorted(glob.glob(os.path.join(HERE, "cases", "*", "Par.inp"))):
        name = os.path.basename(os.path.dirname(src))
        tmp = tempfile.mkdtemp(prefix="hdr_")
        try:
            for f in os.listdir(os.path.dirname(src)):
                if not f.endswith(".py"):
                    shutil.copy(os.path.join(os.path.dirname(src), f), tmp)
            r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {exe} 1"], cwd=tmp, capture_output=True,
                               text=True)
            if "ref_harness: done" not in r.stdout:
                raise RuntimeError(name + r.stdout[-1000:])
            for pat in ("SeisHeader_sem2d.hdr", "Flt??_sem2d.hdr", "Flt??_init_sem2d.tab"):
                for path in glob.glob(os.path.join(tmp, pat)):
                    with open(path) as fh:
                        texts[f"{name}:{os.path.basename(path)}"] = np.array(fh.read())
        finally:
            shutil.rmtree(tmp)
    np.savez_compressed(OUT, **texts)
    print(OUT, len(texts), "files")


if __name__ == "__main__":
    main()
