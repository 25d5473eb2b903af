"""Builds the oracle's C++ transport restatement into oracle/_build/liboracle.so
(test infrastructure only; see transport_oracle.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
SO = os.path.join(OUT, "liboracle.so")


def build():
    os.makedirs(OUT, exist_ok=True)
    tmp = SO + ".tmp"
    cmd = ["g++", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-o", tmp,
           os.path.join(HERE, "transport_oracle.cpp")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, SO)
    return SO


if __name__ == "__main__":
    print(build())
