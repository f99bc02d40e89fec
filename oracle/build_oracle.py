"""Build the C restatement of the time step (the oracle; test infrastructure only).

Output: oracle/lib/libsem2d_oracle.so.  -ffp-contract=off keeps every product and
sum separately rounded, as in the reference's Fortran built for x86-64.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "lib", "libsem2d_oracle.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
                    "-o", OUT, os.path.join(HERE, "sem2d_oracle.c"), "-lm"], check=True)
    return OUT


if __name__ == "__main__":
    print("built", build())
