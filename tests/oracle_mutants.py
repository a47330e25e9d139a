"""Copies of the CPU oracle with one planted slip, for the tests that show a
pin catches that slip.  The slip is planted in a copy of the oracle's
sources under a temporary directory and compiled there with g++; the
oracle under oracle/ is never modified."""
import os
import shutil
import subprocess
import tempfile

import oracle as O

ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))
SOURCES = ("shell_oracle.cpp", "shell_oracle.h", "oracle_geom.h")


def build(old, new, tag):
    """an oracle library handle with the text `old` of shell_oracle.cpp
    replaced by `new` (which must occur exactly once)"""
    d = tempfile.mkdtemp(prefix="oracle_mut_")
    try:
        for f in SOURCES:
            shutil.copy(os.path.join(ORACLE_DIR, f), d)
        p = os.path.join(d, "shell_oracle.cpp")
        src = open(p).read()
        assert src.count(old) == 1, f"planted slip '{tag}': the text to replace occurs {src.count(old)} times"
        open(p, "w").write(src.replace(old, new))
        so = os.path.join(d, f"liboracle_{tag}.so")
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                        "-I", d, "-o", so, p], check=True)
        return O.load(so)
    finally:
        # the loaded library stays mapped after its file is removed
        shutil.rmtree(d, ignore_errors=True)
