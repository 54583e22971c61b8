"""Build-time guard over this repository's own gfx950 code: compiles every
HIP source to AMD assembly and fails if any instruction writes through the
scalar data cache (scalar stores, scalar atomics, scalar cache write-back or
discard) or any packed-FP32 arithmetic appears.  Runs in build() on the
build host only.

usage: python check_isa.py [libpetar_hip.so]   (the argument is informational)
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["pth_kernels.hip", "pth_gtree.hip", "pth_noalloc.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fno-slp-vectorize", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-S", "--cuda-device-only", "-o", "-"]
FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_scratch_store|s_atomic_|s_buffer_atomic_|"
                       r"s_dcache_wb|s_dcache_discard|v_pk_fma_f32|v_pk_mul_f32|v_pk_add_f32)", re.M)


def main():
    bad = []
    for src in SOURCES:
        asm = subprocess.run(["/opt/rocm/bin/hipcc"] + FLAGS + [os.path.join(HERE, src)], capture_output=True,
                             text=True, check=True).stdout
        n_ins = len(re.findall(r"^\s+[sv]_\w+", asm, re.M))
        hits = FORBIDDEN.findall(asm)
        if hits:
            bad.append((src, sorted(set(hits))))
        print(f"check_isa: {src}: {n_ins} instructions scanned, forbidden forms: {sorted(set(hits)) or 'none'}")
    if bad:
        sys.exit(f"check_isa: forbidden instructions in {bad}")


if __name__ == "__main__":
    main()
