"""Builds libclb.so (HIP for gfx950) in place next to this package."""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                       # complab3d-gems_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(ROOT, "libclb.so")

SOURCES = ["clb_lattice.hip", "clb_equilibrium.hip", "clb_aux.hip", "clb_host.cpp", "clb_geometry.cpp"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
OBJ = os.path.join(ROOT, "_obj")


def build(verbose=False):
    """Compiles every source to an object in parallel, then links libclb.so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + CSRC]
    jobs, objs = [], []
    for s in SOURCES:
        o = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmd = [hipcc] + FLAGS + inc + ["-c", os.path.join(CSRC, s), "-o", o]
        jobs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for s, proc in jobs:
        out, _ = proc.communicate()
        if proc.returncode != 0:
            errors.append("%s:\n%s" % (s, out))
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    # the C++ programs that drive a case through clb_driver.hh (the set-up and
    # whole steps) and through clb_processors.hh (the processor form) alone
    exes = []
    for name in ("clb_run_case", "clb_run_processors"):
        exe = os.path.join(ROOT, name)
        cmd = ["g++", "-O2", "-std=c++11", "-Wall", "-I" + os.path.join(REPO, "include"),
               os.path.join(CSRC, name + ".cpp"), "-o", exe + ".tmp", "-L" + ROOT, "-lclb", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s failed to build:\n%s%s" % (name, r.stdout, r.stderr))
        os.replace(exe + ".tmp", exe)
        exes.append(exe)
    if verbose:
        print("built", LIB, "and", ", ".join(exes))
    return LIB
