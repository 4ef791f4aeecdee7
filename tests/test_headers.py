"""The drop-in C++ wrappers compile (CPU): the header is self-contained C++20."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_wrapper_header_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "markovkit_amd/markovkit.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                    f"-I{ROOT}/include", str(src)], check=True)


def test_c_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "markovkit_amd.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                    f"-I{ROOT}/include", str(src)], check=True)
