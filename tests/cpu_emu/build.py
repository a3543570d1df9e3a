"""Build the library for the CPU emulation of the HIP runtime (test infrastructure).

The library's own sources (dmff_amd/csrc, include/admp_hip.h) are compiled with
g++ against tests/cpu_emu/hip/hip_runtime.h and tests/cpu_emu/emu.cpp: kernels run
on fibres (one per GPU thread; barriers and wave operations switch between
them) and every device buffer ends at a guard page, so an access past a buffer
faults at once and the handler names the kernel, block and thread.  The one
textual change to the sources: dynamic LDS declared as an extern array becomes
a pointer to a guarded per-launch buffer.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def build(out_dir):
    src = os.path.join(out_dir, "dmff_amd", "csrc")
    inc = os.path.join(out_dir, "include")
    os.makedirs(src, exist_ok=True)
    os.makedirs(inc, exist_ok=True)
    csrc = os.path.join(REPO, "dmff_amd", "csrc")
    for name in os.listdir(csrc):
        if name.endswith((".hip", ".h")):
            with open(os.path.join(csrc, name)) as f:
                text = f.read()
            text = text.replace("extern __shared__ unsigned char smem_raw[];",
                                "unsigned char* smem_raw = (unsigned char*)emu::dyn_smem();")
            text = text.replace("extern __shared__ unsigned int srt[];",
                                "unsigned int* srt = (unsigned int*)emu::dyn_smem();")
            with open(os.path.join(src, name), "w") as f:
                f.write(text)
    shutil.copy(os.path.join(REPO, "include", "admp_hip.h"), inc)
    lib = os.path.join(out_dir, "libadmp_emu.so")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0",
           "-fno-strict-aliasing", "-w", "-I" + HERE, "-x", "c++", os.path.join(src, "admp_api.hip"),
           os.path.join(HERE, "emu.cpp"), "-o", lib]
    subprocess.run(cmd, check=True)
    return lib
