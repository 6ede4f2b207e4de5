"""Drop-in check of the reference's driver call sites.

tests/compat/driver.cpp holds the reference's own call lines (3B/ht_init.cu:45,
3O/ht_init.cu:47, 3T/ht_init.cu:35, 4B/ht_init.cu:29, 4O/ht_init.cu:53 of the reference's
third_order_tensor_decompsition/CUDA and fourth_order_tensor_decomposition/single_GPU/CUDA
trees). On the CPU it must compile unchanged against include/ht_amd.h and link with
libht_amd.so; on the GPU each variant runs on a seeded tensor and its printed error is
checked against the FP64 oracle.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import gen as G
from oracle import ht_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ht_tensor_learning_using_gpu_tensor_cores_amd")
SRC = os.path.join(ROOT, "tests", "compat", "driver.cpp")


def build_driver(outdir):
    exe = os.path.join(str(outdir), "ht_compat_driver")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable", SRC, "-o",
           exe, f"-L{PKG}", "-lht_amd", f"-Wl,-rpath,{PKG}", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_reference_call_lines_compile_and_link(tmp_path):
    exe = build_driver(tmp_path)
    und = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    syms = set(re.findall(r"\bU (\w+)", und))
    # the 7-argument htd and the 6-argument htd4 bind to the C exports directly; the
    # optimized drivers' 5- and 8-argument calls go through the inline overloads to
    # htd_tc / htd4_tc
    assert {"htd", "htd_tc", "htd4", "htd4_tc"} <= syms, syms
    src = open(SRC).read()
    for line in ("error = htd(X,a,b,c,k,rel_eps,max_rank);", "error = htd(X,a,b,c,k);", "htd(X,a,b,c,k);",
                 "htd4(X,a,b,c,d,k);", "htd4(X,a,b,c,d,k,handle,cusolverH);"):
        assert line in src


def _run(exe, variant, x, dims, k, tmp_path):
    f = os.path.join(str(tmp_path), f"x_{variant}.f32")
    np.asarray(x, dtype=np.float32).tofile(f)
    args = [exe, variant, f] + [str(n) for n in dims] + [str(v) for v in k]
    print("driver", variant, dims, k, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    m = re.findall(r"error rate ([0-9.eE+-]+)", r.stdout)
    assert m, r.stdout
    return float(m[-1]), r.stdout


@pytest.mark.gpu
def test_reference_drivers_run_against_the_oracle(tmp_path):
    exe = build_driver(tmp_path)
    # third-order drivers: the reference's tree {1,2},{3}
    dims3 = [40, 32, 24]
    x3 = G.tucker(dims3, [5, 4, 6], seed=0, noise_rel=1e-3)
    k3 = [1, 12, 6, 5, 4]
    e3 = O.rel_error(x3, O.plain_ht(x3, dims3, O.tree_reference_3d(), k3).reconstruct())
    for variant, tol in (("3B", 1e-5), ("3O", 1e-4), ("3T", 1e-4)):
        e, out = _run(exe, variant, x3, dims3, k3, tmp_path)
        assert abs(e - e3) <= tol + 1e-6 * e3, (variant, e, e3, out)
        if variant in ("3B", "3O"):
            d = float(re.findall(r"driver error ([0-9.eE+-]+)", out)[-1])
            assert abs(d - e) <= 1e-5 * max(e, 1e-30) + 1e-7
    # fourth-order drivers: the balanced tree {1,2},{3,4}
    dims4 = [24, 20, 16, 12]
    x4 = G.tucker(dims4, [4, 5, 3, 4], seed=1, noise_rel=1e-3)
    k4 = [1, 16, 12, 4, 5, 3, 4]
    e4 = O.rel_error(x4, O.plain_ht(x4, dims4, O.tree_balanced_4d(), k4).reconstruct())
    for variant, tol in (("4B", 1e-5), ("4O", 1e-4)):
        e, out = _run(exe, variant, x4, dims4, k4, tmp_path)
        assert abs(e - e4) <= tol + 1e-6 * e4, (variant, e, e4, out)
