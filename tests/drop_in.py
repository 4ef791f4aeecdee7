"""Builds tests/native/test_rdme_dropin: code written against MarkovKit's resolvent interface
compiled against include/markovkit_amd/markovkit.hpp and linked to the in-tree HIP library."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "test_rdme_dropin.cpp")
BIN = os.path.join(HERE, "native", "test_rdme_dropin")


def build():
    tmp = BIN + ".tmp"
    subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{ROOT}/include", SRC,
                    f"-L{ROOT}/markovkit_amd", "-lmarkovkit_amd",
                    "-Wl,-rpath,$ORIGIN/../../markovkit_amd", "-o", tmp], check=True)
    os.replace(tmp, BIN)
