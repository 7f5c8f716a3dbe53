"""Build the gfx950 kernels for the host against the CPU HIP emulation
(tests/emu/hip/hip_runtime.h): every GPU thread is a fiber, blocks run one at
a time, device memory is malloc'd under AddressSanitizer.  Test
infrastructure only; the product library is contactmechanics_amd/libcmhip.so.

The kernel source is used as is, except for three GPU-only spellings that the
host compiler lacks: the global address-space qualifier (dropped), the clang
2-vector type (replaced by a 2-element struct) and an inline s_waitcnt
(dropped; blocks run one at a time here).
"""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "contactmechanics_amd", "csrc", "cm_kernels.hip")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libcmemu.so")


def _host_source():
    s = open(SRC).read()
    s = s.replace('#include "../../include/cmhip.h"', '#include "%s"' % os.path.join(ROOT, "include", "cmhip.h"))
    import re
    s = re.sub(r'extern __shared__ __align__\(16\) unsigned char (\w+)\[\];',
               r'unsigned char* \1 = emu::g_dyn;', s)
    s = s.replace("__attribute__((address_space(1)))", "")
    s = s.replace("using Gv2 = T __attribute__((ext_vector_type(2)));",
                  "struct Gv2 { T v_[2]; T& operator[](int i) { return v_[i]; } "
                  "const T& operator[](int i) const { return v_[i]; } };")
    s = s.replace('asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");', "(void)0;")
    return s


def build(force=False):
    srcs = [SRC, os.path.join(ROOT, "contactmechanics_amd", "csrc", "cm_fft.hpp"),
            os.path.join(HERE, "hip", "hip_runtime.h"), os.path.join(HERE, "emu_rt.cpp")]
    if not force and os.path.exists(LIB) and all(os.path.getmtime(LIB) > os.path.getmtime(p) for p in srcs):
        return LIB
    os.makedirs(OUT_DIR, exist_ok=True)
    gen = os.path.join(OUT_DIR, "cm_kernels_host.cpp")
    with open(gen, "w") as fh:
        fh.write(_host_source())
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-I" + HERE, "-I" + os.path.join(ROOT, "contactmechanics_amd", "csrc"),
           "-DCM_BIS_PER_BLOCK=160",  # small grids: the bisection rounds on several blocks
           "-o", LIB + ".tmp", gen, os.path.join(HERE, "emu_rt.cpp")]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("emulated build failed:\n" + proc.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
