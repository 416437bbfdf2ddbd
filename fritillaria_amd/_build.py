"""Builds the native libraries in-tree.

* ``libfrhip.so``   -- HIP kernels for gfx950 + the C ABI of include/frhip.h
* ``libfrsynth.so`` -- synthetic BGZF/BAM writer (host C, zlib)

Both land next to this file so they travel with the source tree. hipcc
cross-compiles gfx950 code objects without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["inflate.hip", "bam.hip", "pipeline.hip", "fastq.hip", "soa.hip", "noalloc.hip", "api.cpp", "bam_host.cpp", "pipeline_host.cpp",
               "fastq_host.cpp", "serve_host.cpp", "region.hip", "index_host.cpp", "text.hip", "fasta.hip", "text_host.cpp", "deflate.hip", "writer_host.cpp",
               "bcf_host.cpp", "mixed_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed fp32 ops are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def frhip_lib_path() -> str:
    return os.path.join(HERE, "libfrhip.so")


def synth_lib_path() -> str:
    return os.path.join(HERE, "libfrsynth.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r.stdout


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(OBJ, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return obj

    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = frhip_lib_path() + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, frhip_lib_path())
    cc = shutil.which("gcc") or "cc"
    tmp = synth_lib_path() + ".tmp"
    _run([cc, "-O2", "-fPIC", "-shared", "-pthread", "-Wall", "-o", tmp,
          os.path.join(CSRC, "synth.c"), "-lz", "-lm"])
    os.replace(tmp, synth_lib_path())


if __name__ == "__main__":
    build(verbose=True)
