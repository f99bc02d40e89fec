#!/usr/bin/env python3
"""Build the reference SEM2DPACK solver (test infrastructure, never shipped).

This recipe compiles the reference Fortran sources where they lie
(``<ref>/SRC/*.f90``) with AMD ``flang -O2`` in a throw-away temporary
directory and copies only the resulting executables into ``oracle/_ref/``:

* ``oracle/_ref/sem2dsolve``  -- the unmodified reference program (``SRC/main.f90``),
  used as the CPU baseline and to produce the float32 output files;
* ``oracle/_ref/ref_harness`` -- the reference modules linked with our own
  driver ``oracle/ref_harness.f90`` that dumps FP64 arrays (numbering, mass,
  element forces, per-step fields) for pinning the C restatement.

One run-time incompatibility of the reference with flang's runtime is patched in
the temporary copy (never in the repository): three ``DEALLOCATE`` statements of
pointers that were obtained through ``TRANSFER`` (``fem_grid.f90:161``,
``spec_grid.f90:782``, ``spec_grid.f90:795``) abort under flang with
"DEALLOCATE of a pointer that is not the whole content of a pointer ALLOCATE".
They only free linked-list cells during initialisation; commenting them out
leaks a few bytes and changes no arithmetic.

The compile order is derived from the ``use`` statements of the sources (the
reference's own Makefile and dependency file are not used).

Usage: python3 oracle/build_ref.py [--ref /root/reference] [--out oracle/_ref]
"""
import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

# (file, 1-based line, expected statement) -- the leak patch described above
P1 = [
    ("fem_grid.f90", 161, "deallocate( Vertex%P )"),
    ("spec_grid.f90", 782, "deallocate(BC_corner%P)"),
    ("spec_grid.f90", 795, "deallocate(BC_node%P)"),
]

FLANG = os.environ.get("FLANG", "flang")
FFLAGS = ["-O2"]


def _modules(path):
    prov, used = set(), set()
    with open(path, errors="replace") as fh:
        for line in fh:
            s = line.split("!")[0].strip().lower()
            m = re.match(r"module\s+(\w+)\s*$", s)
            if m and m.group(1) != "procedure":
                prov.add(m.group(1))
            m = re.match(r"use\s*(?:,\s*intrinsic\s*::)?\s*(\w+)", s)
            if m:
                used.add(m.group(1))
    return prov, used


def _order(srcdir, files):
    prov_of, used_of = {}, {}
    for f in files:
        p, u = _modules(os.path.join(srcdir, f))
        for m in p:
            prov_of[m] = f
        used_of[f] = u
    deps = {f: {prov_of[m] for m in used_of[f] if m in prov_of and prov_of[m] != f} for f in files}
    out, done = [], set()

    def visit(f, stack=()):
        if f in done:
            return
        if f in stack:
            raise RuntimeError("module cycle at " + f)
        for d in sorted(deps[f]):
            visit(d, stack + (f,))
        done.add(f)
        out.append(f)

    for f in sorted(files):
        visit(f)
    return out


def _patch(tmp):
    for fname, lineno, stmt in P1:
        p = os.path.join(tmp, fname)
        with open(p) as fh:
            lines = fh.readlines()
        got = lines[lineno - 1].strip()
        if got != stmt:
            raise RuntimeError(f"{fname}:{lineno}: expected '{stmt}', found '{got}'")
        lines[lineno - 1] = "!" + lines[lineno - 1]
        with open(p, "w") as fh:
            fh.writelines(lines)


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("failed: " + " ".join(cmd))


def build(ref="/root/reference", out=os.path.join(HERE, "_ref"), quiet=False):
    src = os.path.join(ref, "SRC")
    if not os.path.isdir(src):
        raise FileNotFoundError(src)
    os.makedirs(out, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="sem2d_ref_")
    try:
        files = sorted(f for f in os.listdir(src) if f.endswith(".f90"))
        for f in files:
            shutil.copy(os.path.join(src, f), tmp)
        _patch(tmp)
        order = _order(tmp, files)
        objs = []
        for f in order:
            _run([FLANG] + FFLAGS + ["-c", f], tmp)
            objs.append(f[:-4] + ".o")
        lib_objs = [o for o in objs if o != "main.o"]
        _run([FLANG] + FFLAGS + ["-o", "sem2dsolve"] + objs, tmp)
        shutil.copy(os.path.join(HERE, "ref_harness.f90"), tmp)
        _run([FLANG] + FFLAGS + ["-c", "ref_harness.f90"], tmp)
        _run([FLANG] + FFLAGS + ["-o", "ref_harness", "ref_harness.o"] + lib_objs, tmp)
        for exe in ("sem2dsolve", "ref_harness"):
            shutil.copy(os.path.join(tmp, exe), os.path.join(out, exe))
            os.chmod(os.path.join(out, exe), 0o755)
        if not quiet:
            print("built", ", ".join(os.path.join(out, e) for e in ("sem2dsolve", "ref_harness")))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=os.path.join(HERE, "_ref"))
    a = ap.parse_args()
    build(a.ref, a.out)
