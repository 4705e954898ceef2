"""Build recipe of the C oracle (test infrastructure): oracle/lpm_oracle.c and oracle/lpm_oracle_amg.c ->
oracle/liblpm_oracle.so (gcc, OpenMP).
The reference itself is pure Python / Numba, so there is no reference source to compile into oracle/_ref."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblpm_oracle.so")


def build():
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-o", tmp,
           os.path.join(HERE, "lpm_oracle.c"), os.path.join(HERE, "lpm_oracle_amg.c"), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("gcc failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
