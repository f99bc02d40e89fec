"""Build the C restatement of the time step (the oracle; test infrastructure only).

Output: oracle/lib/libsem2d_oracle.so.  -ffp-contract=off keeps every product and
sum separately rounded, as in the reference's Fortran built for x86-64.  OpenMP
parallelises only work whose result does not depend on the thread count (the
per-element forces, then a sequential assembly in element order; the pointwise
predictor and corrector).

The perturbation ensemble (the same source, each build changing one thing that a
correct implementation of the reference's algorithm may legitimately do differently)
measures how far the reference arithmetic itself moves under round-off; the parity
tests use its spread as the bar of quantities the rate-and-state Newton solve makes
ill-conditioned:

  ulp_p, ulp_m   every exp/log/sinh/cosh/asinh result moved by +1 / -1 ulp
  ulp_r1..ulp_r4 every exp/log/sinh/cosh/asinh/pow result moved by -1, 0 or +1 ulp,
                 the direction drawn per call (four seeds)
  rev            element forces assembled in reverse element order
  col            element forces assembled colour by colour (greedy colouring)
  fma            products and sums contracted into fused multiply-adds
  fmacol         both of the last two (the arithmetic pattern of the GPU kernels)
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
OUT = os.path.join(LIBDIR, "libsem2d_oracle.so")

BASE = ["gcc", "-O2", "-std=c99", "-fopenmp", "-fPIC", "-shared", "-Wall"]
NOCONTRACT = ["-ffp-contract=off"]
CONTRACT = ["-ffp-contract=fast", "-mfma"]

VARIANTS = {
    "": NOCONTRACT,
    "ulp_p": NOCONTRACT + ["-DORACLE_ULP_PERTURB=1"],
    "ulp_m": NOCONTRACT + ["-DORACLE_ULP_PERTURB=-1"],
    "ulp_r1": NOCONTRACT + ["-DORACLE_ULP_RANDOM=1"],
    "ulp_r2": NOCONTRACT + ["-DORACLE_ULP_RANDOM=2"],
    "ulp_r3": NOCONTRACT + ["-DORACLE_ULP_RANDOM=3"],
    "ulp_r4": NOCONTRACT + ["-DORACLE_ULP_RANDOM=4"],
    "rev": NOCONTRACT + ["-DORACLE_ORDER=1"],
    "col": NOCONTRACT + ["-DORACLE_ORDER=2"],
    "fma": CONTRACT,
    "fmacol": CONTRACT + ["-DORACLE_ORDER=2"],
}
ENSEMBLE = [k for k in VARIANTS if k]


def lib_path(variant=""):
    return os.path.join(LIBDIR, "libsem2d_oracle%s.so" % ("_" + variant if variant else ""))


def build():
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(HERE, "sem2d_oracle.c")
    for name, extra in VARIANTS.items():
        out = lib_path(name)
        subprocess.run(BASE + extra + ["-o", out + ".tmp", src, "-lm"], check=True)
        os.replace(out + ".tmp", out)
    return OUT


if __name__ == "__main__":
    print("built", build())
