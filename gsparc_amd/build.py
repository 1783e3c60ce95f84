"""Build recipe for the native libraries (run: python -m gsparc_amd.build).

  lib/libsptc_gen.so  host-only workload generator (g++)
  lib/libsptc_hip.so  gfx950 kernels + C ABI (hipcc, --offload-arch=gfx950)
  bin/gsparc          the reference-compatible command line (g++)
  bin/main_block      the reference's main.cu SpTC block compiled unchanged
                      against include/gsparc/ (only where the reference
                      checkout is present)

Every call recompiles everything (no incremental state), in parallel.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

from ._lib import LIB_DIR, PKG_DIR, REPO_DIR

CSRC = os.path.join(PKG_DIR, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_SOURCES = ["sort.hip", "segsort.hip", "kernels.hip", "batch.hip", "result.hip", "noalloc.hip",
               "tdense.hip", "bucket.hip",
               "device.cpp", "engine.cpp", "staticpart.cpp",
               "batch_engine.cpp", "serve.cpp", "api.cpp"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # products are rounded to FP32 before they are summed, on every path (an
    # FMA of product and running sum would make the dense and ESC
    # accumulators round differently)
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(REPO_DIR, "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_gen():
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libsptc_gen.so")
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", out,
          os.path.join(CSRC, "gen.cpp")])
    return out


def build_hip(jobs=8):
    os.makedirs(LIB_DIR, exist_ok=True)
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                print(out)
    lib = os.path.join(LIB_DIR, "libsptc_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_oracle():
    """The CPU checker (oracle/liboracle.so) and, when the reference sources
    are present, the reference CPU SpTC binary (oracle/_ref/ttt)."""
    odir = os.path.join(REPO_DIR, "oracle")
    _run(["make", "-s", "-B", "-C", odir, "all"])
    if os.path.isdir("/root/reference/baselines/Sparta/src"):
        _run(["make", "-s", "-C", odir, "ref"])


def build_cli():
    """bin/gsparc: the reference-compatible command line (host C++ over the
    C ABI; links lib/libsptc_hip.so through an $ORIGIN-relative rpath)."""
    bindir = os.path.join(PKG_DIR, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "gsparc")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I",
          os.path.join(REPO_DIR, "include"), "-o", out,
          os.path.join(PKG_DIR, "cli", "gsparc_main.cpp"), "-L", LIB_DIR, "-lsptc_hip",
          "-Wl,-rpath,$ORIGIN/../lib"])
    return out


REF_MAIN = "/root/reference/main.cu"


def extract_main_block(path, dst):
    """Writes the reference driver's SpTC iteration body -- from its
    `if (tensor_x->nbits > 64)` to the line before `// delete SZ;` -- to dst
    (the build tree) for tests/dropin/main_block.cpp to include."""
    lines = open(path).read().splitlines(keepends=True)
    b = next(i for i, l in enumerate(lines) if "if (tensor_x->nbits > 64)" in l)
    e = next(i for i, l in enumerate(lines) if "// delete SZ;" in l)
    with open(dst, "w") as f:
        f.write("".join(lines[b:e]))
    return b + 1, e


def build_main_block(out=None, ref_main=REF_MAIN):
    """bin/main_block: the reference's own main.cu SpTC block compiled
    unchanged against include/gsparc/ and lib/libsptc_hip.so (needs the
    reference checkout; the block itself stays in the build tree)."""
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    extract_main_block(ref_main, os.path.join(objdir, "main_block.inc"))
    out = out or os.path.join(PKG_DIR, "bin", "main_block")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-Wno-unknown-pragmas", "-I",
          os.path.join(REPO_DIR, "include"), "-I", objdir, '-DMAIN_BLOCK="main_block.inc"',
          "-o", out, os.path.join(REPO_DIR, "tests", "dropin", "main_block.cpp"), "-L", LIB_DIR,
          "-lsptc_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    return out


def build_all():
    build_gen()
    build_hip()
    build_cli()
    build_oracle()
    if os.path.exists(REF_MAIN):
        build_main_block()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gen": build_gen, "hip": build_hip, "cli": build_cli,
         "oracle": build_oracle, "main_block": build_main_block}[w]()
    print("built:", ", ".join(what))
