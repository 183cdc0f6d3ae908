"""Build the MEX gateways (staq_dic_amd/mex) against the test stand-in MEX
API (tests/mexshim) into tests/mexshim/_build/<name>.so, one library per
gateway.  The staq_* symbols are left to the process, which loads
libstaqdic.so globally first (a MATLAB build links -lstaqdic instead)."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
MEX = os.path.join(ROOT, "staq_dic_amd", "mex")
OUT = os.path.join(HERE, "_build")
GATEWAYS = {
    "ba_interp2": "ba_interp2_mex.cpp",
    "staq_local_icgn_mex": "staq_local_icgn_mex.cpp",
    "staq_subpb1_mex": "staq_subpb1_mex.cpp",
    "staq_integer_search_mex": "staq_integer_search_mex.cpp",
}


def path(name: str) -> str:
    return os.path.join(OUT, name + ".so")


def build(force: bool = True) -> dict:
    """Compile every gateway with g++ (always when force)."""
    cxx = shutil.which("g++") or "g++"
    os.makedirs(OUT, exist_ok=True)
    deps = [os.path.join(HERE, "fake_mx.cpp"), os.path.join(HERE, "mex.h"),
            os.path.join(MEX, "staq_mex.h"), os.path.join(ROOT, "include", "staq_dic.h")]
    out = {}
    for name, src in GATEWAYS.items():
        so = path(name)
        srcs = [os.path.join(MEX, src), os.path.join(HERE, "fake_mx.cpp")]
        fresh = os.path.exists(so) and all(os.path.getmtime(so) >= os.path.getmtime(d)
                                           for d in srcs + deps)
        if force or not fresh:
            cmd = [cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I", HERE, "-I", MEX,
                   "-I", os.path.join(ROOT, "include"), "-o", so + ".tmp", *srcs]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"building gateway {name} failed:\n{r.stderr}")
            os.replace(so + ".tmp", so)
        out[name] = so
    return out


if __name__ == "__main__":
    for k, v in build().items():
        print(k, v)
