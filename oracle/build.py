"""Build the CPU checker (oracle/libsiminf_oracle.so) with make/gcc.
Test infrastructure only."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsiminf_oracle.so")


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB
