"""Foreign callers of the plan API linked against libnfft3_hip.so.

* tests/c/mv_caller.c: a consumer that wraps an nfft_plan in a matrix-vector
  plan of its own and runs the solver on it, the pattern of the reference's
  kernel/mri/mri.c:159-262, in C99 complex arithmetic (<complex.h> included
  before nfft3.h, so fftw_complex is double _Complex as with FFTW's header).
  Compiled here with -Wall -Werror; run on the GPU (fast vs direct CGNR
  iterates agree to 1e-8, the residual falls).
* The reference's Julia shim julia/nfft/libnfftjulia.c, compiled from where it
  lies against include/nfft3.h and linked against the library (CPU: every
  symbol it binds resolves at load time).  Its one autotools dependency,
  `#include "config.h"`, is left out of the compile and the one macro the
  shim reads from it, HAVE_COMPLEX_H, is set on the command line.  Skipped
  where the reference checkout is absent.
"""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "nfft_amd")
SHIM = "/root/reference/julia/nfft/libnfftjulia.c"


def _link_args():
    return ["-L", LIBDIR, "-lnfft3_hip", "-lm", "-Wl,-rpath," + LIBDIR]


def _build_mv_caller(tmp_path):
    exe = tmp_path / "mv_caller"
    subprocess.run(["gcc", "-std=gnu99", "-O1", "-Wall", "-Werror", "-I", INC,
                    os.path.join(ROOT, "tests", "c", "mv_caller.c"), "-o", str(exe)]
                   + _link_args(), check=True)
    return exe


def test_mv_caller_compiles_and_links(tmp_path):
    assert os.path.exists(os.path.join(LIBDIR, "libnfft3_hip.so"))
    assert _build_mv_caller(tmp_path).exists()


def test_header_complex_types_follow_fftw(tmp_path):
    """With <complex.h> first the plan's f / f_hat are double _Complex *, else
    double (*)[2]; both layouts are 16 bytes per value and the struct offsets
    are the same."""
    src = tmp_path / "t.c"
    src.write_text(
        "#include <complex.h>\n#include <stddef.h>\n#include \"nfft3.h\"\n"
        "_Static_assert(sizeof(fftw_complex) == 16, \"c16\");\n"
        "_Static_assert(sizeof(fftwf_complex) == 8, \"c8\");\n"
        "int main(void){ nfft_plan p; double _Complex z = 1.0 + 2.0 * I; p.f = &z;\n"
        "  p.f[0] *= 2.0; return (int)(offsetof(nfft_plan, x) != 0 && creal(p.f[0]) == 2.0 ? 0 : 1); }\n")
    src2 = tmp_path / "u.c"
    src2.write_text(
        "#include <stddef.h>\n#include <stdio.h>\n#include \"nfft3.h\"\n"
        "int main(void){ printf(\"%zu %zu\\n\", offsetof(nfft_plan, x), offsetof(solver_plan_complex, dot_r_iter)); return 0; }\n")
    src3 = tmp_path / "v.c"
    src3.write_text(
        "#include <complex.h>\n#include <stddef.h>\n#include <stdio.h>\n#include \"nfft3.h\"\n"
        "int main(void){ printf(\"%zu %zu\\n\", offsetof(nfft_plan, x), offsetof(solver_plan_complex, dot_r_iter)); return 0; }\n")
    outs = []
    for s in (src, src2, src3):
        exe = tmp_path / (s.stem + ".x")
        subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror", "-I", INC, str(s), "-o", str(exe)],
                       check=True)
        r = subprocess.run([str(exe)], capture_output=True, text=True)
        assert r.returncode == 0
        outs.append(r.stdout)
    assert outs[1] == outs[2]


@pytest.mark.skipif(not os.path.exists(SHIM), reason="reference checkout absent")
def test_julia_shim_links_against_library(tmp_path):
    so = tmp_path / "libnfftjulia.so"
    with open(SHIM) as fh:
        text = fh.read()
    assert '#include "config.h"' in text
    code = text.replace('#include "config.h"', "")
    subprocess.run(["gcc", "-std=gnu99", "-shared", "-fPIC", "-DHAVE_COMPLEX_H", "-I", INC,
                    "-x", "c", "-", "-o", str(so)] + _link_args(),
                   input=code, text=True, check=True)
    lib = ctypes.CDLL(str(so), mode=ctypes.RTLD_GLOBAL | os.RTLD_NOW)
    for sym in ("jnfft_alloc", "jnfft_init", "jnfft_set_x", "jnfft_set_fhat", "jnfft_set_f",
                "jnfft_trafo", "jnfft_adjoint", "jnfft_trafo_direct", "jnfft_adjoint_direct",
                "jnfft_finalize"):
        assert hasattr(lib, sym), sym
    # every nfft_ symbol the shim leaves undefined is exported by the library,
    # and so are the ones NFFT.jl ccalls through the shim's library path
    und = subprocess.run(["nm", "-D", "--undefined-only", str(so)], capture_output=True,
                         text=True, check=True).stdout.split()
    need = {s for s in und if s.startswith("nfft_")}
    need |= {"nfft_get_default_window_cut_off", "nfft_get_num_threads", "nfft_set_num_threads"}
    have = set(subprocess.run(["nm", "-D", "--defined-only",
                               os.path.join(LIBDIR, "libnfft3_hip.so")],
                              capture_output=True, text=True, check=True).stdout.split())
    assert need and need <= have, sorted(need - have)


@pytest.mark.gpu
def test_mv_caller_runs(tmp_path):
    exe = _build_mv_caller(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "mv_caller: OK" in r.stdout, r.stdout + r.stderr
