"""The CPU baseline of bench.py (TEST INFRASTRUCTURE ONLY): the reference's
CPU algorithms restated, timed on the host's cores, in a process of its own.

It runs in a child process without torch: MKL Pardiso and the OpenMP runtime
torch loads do not coexist in one process (Pardiso's symbolic phase then
fails).  It prints one JSON object:

* ``host``: the CPU model and the threads used;
* ``assembly``: the C/OpenMP restatement of the reference's CPU stiffness
  assembly (oracle/assemble_ref.c, FEM_model_csr.py:150-165, 246-257: per
  element intermediate, linear slot search, atomic +=) on the 60x60x55 slab,
  compiled here with -march=native (or the prebuilt x86-64-v3 build), the
  median of as many runs as fill about 10 s after a warm-up;
* ``static``: the reference's static solve (SOLVE_static.py:31-38: Pardiso
  mtype 2 on the penalised K, relative residual) on the 20x20x50 cantilever,
  MKL Pardiso when libmkl_rt loads, else SciPy SuperLU (named in ``solver``);
* ``modal``: the reference's shift-invert block Lanczos (SOLVE_eigen.py:45-89,
  the NumPy restatement fem_oracle.lanczos) with that factorisation, 20 modes
  with block 8, on the same cantilever.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import fem_oracle as O  # noqa: E402
from oracle import c_oracle  # noqa: E402
from oracle.pardiso import DirectSolver  # noqa: E402


def _mesh_module():
    """fem_taichi_amd/mesh.py on its own (the package imports torch)."""
    spec = importlib.util.spec_from_file_location("fem_mesh_gen", os.path.join(ROOT, "fem_taichi_amd", "mesh.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def native_oracle():
    """Build assemble_ref.c with -march=native for this host into a temporary
    file and load it in place of the prebuilt one; returns the build flag used."""
    src = os.path.join(ROOT, "oracle", "assemble_ref.c")
    out = os.path.join(tempfile.gettempdir(), f"libfem_oracle_native_{os.getpid()}.so")
    r = subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-o", out, src,
                        "-lm"], capture_output=True, text=True)
    if r.returncode != 0:
        return "x86-64-v3 (prebuilt; -march=native build failed)"
    c_oracle._PATH = out
    c_oracle._lib = None
    c_oracle.lib()
    return "-march=native"


def assembly_leg(mesh, seconds):
    g = mesh.slab_60x60x55()
    pp = O.pre_process(g.Node, g.Elem)
    xyz = np.ascontiguousarray(g.Node[1:].T)
    en = O.element_nodes(g.Elem, pp["Nind"]).astype(np.int32)
    ex = np.asarray(g.EX)
    args = (xyz, en, ex if ex.ndim else None, float(ex) if ex.ndim == 0 else 0.0, g.PO,
            pp["SmNode"], pp["CoNode"], pp["ReNode"])
    c_oracle.assemble_K_csr(*args, with_cols=False)          # warm-up (like excluding Taichi JIT)
    times = []
    while len(times) < 5 or sum(times) < seconds:
        t = time.perf_counter()
        c_oracle.assemble_K_csr(*args, with_cols=False)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    return {"value": g.Num_Elem / med, "unit": "hex8 elems assembled/s", "seconds_median": med, "runs": len(times),
            "workload": "slab_60x60x55 (198000 elements), FP64 CSR K"}


def penalised_K(g, pp):
    """Penalised FP64 K in the reference's CSR layout (C port values, Q1 penalty)."""
    import scipy.sparse as sp
    xyz = np.ascontiguousarray(g.Node[1:].T)
    en = O.element_nodes(g.Elem, pp["Nind"]).astype(np.int32)
    ex = np.asarray(g.EX)
    val, col = c_oracle.assemble_K_csr(xyz, en, ex if ex.ndim else None, float(ex) if ex.ndim == 0 else 0.0, g.PO,
                                       pp["SmNode"], pp["CoNode"], pp["ReNode"])
    val = O.modify_K_csr(val, O.stage_bounds(g.Bounds), pp["Nind"], pp["SmNode"], pp["ReNode"], pp["CoNode"])
    row = O.csr_index_K(pp["SmNode"], pp["CoNode"])
    n = 3 * pp["Num_Node"]
    return sp.csr_matrix((val, col, row), shape=(n, n))


def solve_legs(mesh):
    g = mesh.cantilever_20x20x50()
    pp = O.pre_process(g.Node, g.Elem)
    K = penalised_K(g, pp)
    F = O.model_F(g.Node, g.Elem, pp, g.DE, (0.0, 0.0, 0.0), g.Forces)
    t0 = time.perf_counter()
    ds = DirectSolver(K)
    t1 = time.perf_counter()
    U = ds.solve(F)
    t2 = time.perf_counter()
    res = float(np.linalg.norm(K @ U - F) / np.linalg.norm(F))
    static = {"workload": "cantilever_20x20x50 (67473 DOF), penalised FP64 K", "solver": ds.kind,
              "factor_seconds": t1 - t0, "solve_seconds": t2 - t1, "seconds": t2 - t0, "rel_residual": res}
    M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    t3 = time.perf_counter()
    lam, X, _ = O.lanczos(K, M, ds.solve, num_eigen=20, num_block=8, rng=np.random.default_rng(0))
    t4 = time.perf_counter()
    modal = {"workload": "cantilever_20x20x50, 20 modes, block 8 (NumPy restatement of SOLVE_eigen.py:45-89)",
             "solver": ds.kind, "seconds": t4 - t3, "freq_hz_first": [float(v) for v in np.sqrt(lam[:4]) / 2 / np.pi]}
    ds.release()
    return static, modal


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seconds", type=float, default=10.0)
    args = ap.parse_args()
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS"):
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("MKL_NUM_THREADS", str(threads))
    mesh = _mesh_module()
    build = native_oracle()
    out = {"host": {"cpu": host_cpu(), "threads": threads, "affinity_cpus": len(os.sched_getaffinity(0))},
           "assembly": dict(assembly_leg(mesh, args.seconds), build=build)}
    out["static"], out["modal"] = solve_legs(mesh)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
