"""Build libpwdft_hip.so in-tree with hipcc for gfx950, then check its device code for
packed FP32 arithmetic (v_pk_fma/mul/add_f32), which this library must not contain: the
build passes -packed-fp32-ops to the device compiler (the host pass warns that it does not
know the feature, which is harmless), and check_no_packed_fp32() disassembles every gfx950
code object of the built library and fails the build if any such instruction is there."""
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
SOURCES = ["pw_capi.hip", "pw_fft.hip", "pw_gemm.hip", "pw_forces.hip", "pw_batch.hip", "pw_dense.hip", "pw_scf.hip", "pw_proj.hip", "pw_eig.hip", "pw_lobpcg.hip", "pw_small.hip", "pw_noalloc.hip"]
OUT = os.path.join(HERE, "libpwdft_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wno-unused-value", "-Wno-unused-result",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# rocSOLVER for the small dense Cholesky / eigen problems of the eigensolvers
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build_library(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + srcs + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    n = check_no_packed_fp32(tmp)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} ({n} gfx950 code objects checked: no packed FP32 instructions)")
    return OUT


LLVM = "/opt/rocm/llvm/bin"
_PACKED = re.compile(r"v_pk_(fma|mul|add)_f32")


def check_no_packed_fp32(lib):
    """Disassemble every gfx950 code object bundled in `lib` (one per translation unit) and
    raise if any packed FP32 instruction appears.  Returns the number of code objects."""
    with tempfile.TemporaryDirectory() as td:
        fat = os.path.join(td, "fat.bin")
        r = subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fat}", lib,
                            os.path.join(td, "copy.so")], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("llvm-objcopy: " + r.stderr)
        data = open(fat, "rb").read()
        magic = b"__CLANG_OFFLOAD_BUNDLE__"
        offs = [m.start() for m in re.finditer(re.escape(magic), data)]
        if not offs:
            raise RuntimeError("no offload bundle in " + lib)
        found = []
        for k, o in enumerate(offs):
            e = offs[k + 1] if k + 1 < len(offs) else len(data)
            b, co = os.path.join(td, f"b{k}.bin"), os.path.join(td, f"c{k}.o")
            with open(b, "wb") as f:
                f.write(data[o:e])
            r = subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                                "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={b}", f"--output={co}"],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("clang-offload-bundler: " + r.stderr)
            asm = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--mcpu=gfx950", co],
                                 capture_output=True, text=True, check=True).stdout
            found += [m.group(0) for m in _PACKED.finditer(asm)]
        if found:
            raise RuntimeError(f"packed FP32 instructions in the device code: {sorted(set(found))} ({len(found)})")
        return len(offs)


if __name__ == "__main__":
    build_library(verbose=True)
