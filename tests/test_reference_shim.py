"""The reference-side binding (include/aprl_reference_shim.hpp) compiled
against the reference's own headers (eb_field.h, eb_cpml.h, eb_particles.h,
boostframe.h, moving_window_fields.h, eb_bunch.h) with the MPICH headers,
linked with the library and the reference build, and run for what needs no
GPU (tests/shim/shim_check.cpp): status-to-exception mapping, the
geometry / CPML / frame conversions, and the window planes' component shapes
against the reference's own EBFieldGrid."""
import os
import subprocess

import pytest

import harness as H

ROOT = H.ROOT
SHIM_DIR = os.path.join(ROOT, "tests", "shim")
LIB_DIR = os.path.join(ROOT, "aprl_amd")
REF_SRC = os.path.join(H.REFERENCE_TREE, "src")
MPI_INC = "/opt/conda/include"


@pytest.mark.skipif(not (os.path.isdir(REF_SRC) and H.have_ref() and
                         os.path.exists(os.path.join(LIB_DIR, "libaprl_gpu.so"))),
                    reason="reference tree, reference build or library unavailable")
def test_reference_shim_compiles_and_maps_types(tmp_path):
    exe = str(tmp_path / "shim_check")
    ref_dir = os.path.dirname(H.REF_LIB)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wno-unused-function", "-DMPICH_SKIP_MPICXX", "-DOMPI_SKIP_MPICXX",
           "-I", os.path.join(ROOT, "include"), "-I", REF_SRC, "-I", MPI_INC,
           os.path.join(SHIM_DIR, "shim_check.cpp"), "-o", exe,
           "-L", LIB_DIR, "-laprl_gpu", f"-Wl,-rpath,{LIB_DIR}",
           H.REF_LIB, f"-Wl,-rpath,{ref_dir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and "shim check passed" in run.stdout, run.stdout + run.stderr[-2000:]
