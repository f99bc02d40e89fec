"""Build the C restatement of the time step (the oracle; test infrastructure only).

Output: oracle/lib/libsem2d_oracle.so.  -ffp-contract=off keeps every product and
sum separately rounded, as in the reference's Fortran built for x86-64.  A second
build, oracle/lib/libsem2d_oracle_ulp.so (-DORACLE_ULP_PERTURB), moves every libm
result by one ulp: the parity tests use it to measure the reference arithmetic's own
sensitivity to round-off.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "lib", "libsem2d_oracle.so")
OUT_ULP = os.path.join(HERE, "lib", "libsem2d_oracle_ulp.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    for out, extra in ((OUT, []), (OUT_ULP, ["-DORACLE_ULP_PERTURB"])):
        subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-fPIC", "-shared", "-Wall"] + extra +
                       ["-o", out, os.path.join(HERE, "sem2d_oracle.c"), "-lm"], check=True)
    return OUT


if __name__ == "__main__":
    print("built", build())
