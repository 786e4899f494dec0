"""Builds libcuspx.so in-tree with hipcc for gfx950, then the twelve per-set
libraries libspx_<set>.so (csrc/perset.c with gcc: the api.h names bound to
one parameter set, as the reference builds one library per set).

    python -m cuspx_amd.build            # incremental
    python -m cuspx_amd.build --force    # rebuild everything

Each parameter set is its own translation unit (csrc/set_*.hip) so the
twelve instantiations compile in parallel.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcuspx.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUSPX_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I" + os.path.join(ROOT, "include"),
]


def _deps_mtime():
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    return max(os.path.getmtime(h) for h in hs)


def _compile(src, obj, force):
    if not force and os.path.exists(obj):
        if os.path.getmtime(obj) >= max(os.path.getmtime(src), _deps_mtime()):
            return obj, None
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (os.path.basename(src), out))
    return obj, out or None


def build(force=False, jobs=None, verbose=False):
    """Compile every kernel translation unit and link libcuspx.so."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 1)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile, s, os.path.join(BUILD, os.path.basename(s) + ".o"), force)
                for s in srcs]
        for f in futs:
            obj, msg = f.result()
            objs.append(obj)
            if msg and verbose:
                print(msg)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(o) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
        os.replace(tmp, LIB)
    build_perset(force)
    return LIB


def perset_names():
    """The parameter sets, from the per-set units: set_sha2_128f.hip -> sha2-128f."""
    names = []
    for src in sorted(glob.glob(os.path.join(CSRC, "set_*.hip"))):
        stem = os.path.basename(src)[len("set_"):-len(".hip")]
        names.append(stem.replace("_", "-", 1))
    return names


def perset_path(name):
    return os.path.join(HERE, "libspx_%s.so" % name.replace("-", "_"))


def build_perset(force=False):
    """libspx_<set>.so for every set: csrc/perset.c with CUSPX_SET_NAME, linked
    to libcuspx.so beside it."""
    src = os.path.join(CSRC, "perset.c")
    dep = max(os.path.getmtime(src), os.path.getmtime(LIB),
              os.path.getmtime(os.path.join(ROOT, "include", "cuspx.h")))
    for name in perset_names():
        out = perset_path(name)
        if not force and os.path.exists(out) and os.path.getmtime(out) >= dep:
            continue
        cmd = [os.environ.get("CC", "gcc"), "-O2", "-fPIC", "-shared", "-Wall", "-Wextra",
               '-DCUSPX_SET_NAME="%s"' % name, "-I" + os.path.join(ROOT, "include"), src,
               "-o", out + ".tmp", "-L" + HERE, "-lcuspx", "-ldl", "-Wl,-rpath,$ORIGIN"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0 or p.stderr.strip():
            raise RuntimeError("per-set library %s failed:\n%s" % (name, p.stdout + p.stderr))
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose="-v" in sys.argv)
    print(LIB)
