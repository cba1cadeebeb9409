"""Build step of the package: compiles the HIP sources under csrc/ for gfx950
and links the C-ABI library ``nfft_amd/libnfft3_hip.so``.

    python -m nfft_amd.build           # or nfft_amd.build.build()

Objects go to ``nfft_amd/_build/`` and are recompiled only when a source or
header is newer than the object.  Every HIP unit is compiled with packed-FP32
instructions disabled (-packed-fp32-ops) and hardware float atomics enabled.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnfft3_hip.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result", "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"),
]


def _sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _headers():
    return glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))


def _compile(src, obj, force):
    newest_dep = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in _headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest_dep:
        return src, "up to date", ""
    cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), out))
    return src, "compiled", out


def build(force=False, jobs=None, verbose=True):
    """Compile every HIP unit for gfx950 and link libnfft3_hip.so."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o, force) for s, o in zip(srcs, objs)]
        for fu in cf.as_completed(futs):
            src, status, out = fu.result()
            if verbose:
                print("[nfft_amd.build] %-24s %s" % (os.path.basename(src), status), flush=True)
                if out.strip():
                    print(out, flush=True)
    newest = max(os.path.getmtime(o) for o in objs)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
        os.replace(tmp, LIB)
        if verbose:
            print("[nfft_amd.build] linked %s" % LIB, flush=True)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
