"""CPU checks of the C ABI boundary (no kernel runs here).

* the library loads and exports every symbol that include/*.h declares;
* the plan and solver structs seen by C (gcc, from include/nfft3.h) have
  exactly the field offsets the ctypes mirror uses;
* a C caller written against the reference API compiles and links;
* without a HIP device, creating a plan fails loudly through nfft_die.
"""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import nfft_amd as nf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


@pytest.fixture(scope="module")
def libpath():
    if not os.path.exists(nf.LIB_PATH):
        from nfft_amd import build
        build.build(verbose=False)
    return nf.LIB_PATH


def _declared_symbols(tmp_path):
    src = tmp_path / "decl.c"
    src.write_text('#include "nfft3.h"\n#include "nfft3_hip.h"\n')
    out = subprocess.run(["gcc", "-E", "-P", "-I", INC, str(src)], capture_output=True, text=True,
                         check=True).stdout
    funcs = set(re.findall(r"\b((?:nfftf?|solverf?)_\w+)\s*\(", out))
    hooks = set(re.findall(r"extern\s+\w+\s+((?:nfftf?)_\w+_hook)\s*;", out))
    types = set(re.findall(r"typedef[^;]*?\b((?:nfftf?|solverf?)_\w+)\s*;", out))
    types |= set(re.findall(r"\(\*((?:nfftf?)_\w+_function)\)", out))
    return (funcs - types), hooks


def test_exports_every_declared_symbol(tmp_path, libpath):
    funcs, hooks = _declared_symbols(tmp_path)
    assert len(funcs) > 100
    L = ctypes.CDLL(libpath)
    missing = [s for s in sorted(funcs | hooks) if not hasattr(L, s)]
    assert not missing, missing


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "nfft3.h"
#define F(T, f) printf("%s %s %zu\n", #T, #f, offsetof(T, f));
#define PLAN(T) F(T,N_total) F(T,M_total) F(T,f_hat) F(T,f) F(T,mv_trafo) F(T,mv_adjoint) \
  F(T,d) F(T,N) F(T,sigma) F(T,n) F(T,n_total) F(T,m) F(T,b) F(T,K) F(T,flags) F(T,fftw_flags) \
  F(T,x) F(T,MEASURE_TIME_t) F(T,my_fftw_plan1) F(T,my_fftw_plan2) F(T,c_phi_inv) F(T,psi) \
  F(T,psi_index_g) F(T,psi_index_f) F(T,g) F(T,g_hat) F(T,g1) F(T,g2) F(T,spline_coeffs) \
  F(T,index_x) printf("%s sizeof %zu\n", #T, sizeof(T));
#define SOLV(T) F(T,mv) F(T,flags) F(T,w) F(T,w_hat) F(T,y) F(T,f_hat_iter) F(T,r_iter) \
  F(T,z_hat_iter) F(T,p_hat_iter) F(T,v_iter) F(T,alpha_iter) F(T,beta_iter) F(T,dot_r_iter) \
  F(T,dot_r_iter_old) F(T,dot_z_hat_iter) F(T,dot_z_hat_iter_old) F(T,dot_p_hat_iter) \
  F(T,dot_v_iter) printf("%s sizeof %zu\n", #T, sizeof(T));
#define MV(T) F(T,N_total) F(T,M_total) F(T,f_hat) F(T,f) F(T,mv_trafo) F(T,mv_adjoint) \
  printf("%s sizeof %zu\n", #T, sizeof(T));
int main(void) { PLAN(nfftf_plan) PLAN(nfft_plan) SOLV(solverf_plan_complex) SOLV(solver_plan_complex)
  SOLV(solverf_plan_double) SOLV(solver_plan_double) MV(nfftf_mv_plan_double) MV(nfft_mv_plan_double)
  return 0; }
"""


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", INC, str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    types = {"nfftf_plan": nf.nfftf_plan, "nfft_plan": nf.nfft_plan,
             "solverf_plan_complex": nf.solverf_plan_complex, "solver_plan_complex": nf.solver_plan_complex,
             "solverf_plan_double": nf.solverf_plan_double, "solver_plan_double": nf.solver_plan_double,
             "nfftf_mv_plan_double": nf.nfftf_mv_plan_double, "nfft_mv_plan_double": nf.nfft_mv_plan_double}
    n = 0
    for line in lines:
        if not line.strip():
            continue
        T, field, val = line.split()
        if field == "sizeof":
            assert ctypes.sizeof(types[T]) == int(val), (T, val)
        else:
            assert getattr(types[T], field).offset == int(val), (T, field, val)
        n += 1
    assert n == 2 * 31 + 4 * 19 + 2 * 7


def test_c_caller_compiles_and_links(tmp_path, libpath):
    src = os.path.join(ROOT, "tests", "c", "simple_test.c")
    exe = tmp_path / "simple_test"
    subprocess.run(["gcc", "-std=c99", "-O1", "-I", INC, src, "-o", str(exe),
                    "-L", os.path.dirname(libpath), "-lnfft3_hip", "-lm",
                    "-Wl,-rpath," + os.path.dirname(libpath)], check=True)
    assert exe.exists()


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_no_device_fails_loudly(libpath):
    code = ("import ctypes, nfft_amd as nf\n"
            "L = nf.lib()\n"
            "p = nf.nfftf_plan()\n"
            "L.nfftf_init_1d(ctypes.byref(p), 16, 10)\n"
            "print('not reached')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode != 0
    assert "not reached" not in r.stdout
    assert "no HIP device" in (r.stderr + r.stdout)


@pytest.mark.parametrize("cnt", [1, 2, 7, 255, 256, 257, 4095, 4096, 4097, 76825, 1 << 20])
@pytest.mark.parametrize("max_per_sub,cap", [(64, 0), (256, 0), (256, 300), (256, 3), (1024, 0),
                                             (1024, 1), (4096, 2), (512, 100000)])
def test_work_split_covers_bin(libpath, cnt, max_per_sub, cap):
    """The B-stage work items of one bin cover its nodes exactly once, each
    item non-empty and within the size limit unless the item cap forces longer
    items (the node binning's split, host copy of the device function)."""
    L = ctypes.CDLL(libpath)
    k_max = (cnt + max_per_sub - 1) // max_per_sub
    first = (ctypes.c_int * max(1, k_max))()
    count = (ctypes.c_int * max(1, k_max))()
    k = L.nfft_hip_work_split(cnt, max_per_sub, cap, first, count)
    assert k == (min(k_max, cap) if cap > 0 else k_max)
    f = list(first[:k])
    c = list(count[:k])
    assert all(x >= 1 for x in c)
    assert sum(c) == cnt
    assert f[0] == 0 and all(f[i] + c[i] == f[i + 1] for i in range(k - 1))
    assert max(c) - min(c) <= 1
    if cap == 0 or k_max <= cap:
        assert max(c) <= max_per_sub
