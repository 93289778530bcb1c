"""The C-ABI library loads and exports every entry point include/faunus_amd.h
declares; without a GPU the engine refuses to run (no CPU fallback)."""
import os
import re
import subprocess

import pytest

import faunus_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "faunus_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fmc_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ["fmc_create", "fmc_destroy", "fmc_delta_u", "fmc_energy", "fmc_sync", "fmc_stage_trial",
              "fmc_total_energy", "fmc_init", "fmc_replay_translate"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = F.lib()
    syms = declared_symbols()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fmc_[a-z_]+)", out))
    assert set(syms) <= exported
    # the Python wrapper binds exactly the declared entry points
    assert set(F.SIGNATURES) == set(syms)


def test_library_is_built_for_gfx950():
    data = open(F.LIB_PATH, "rb").read()
    assert b".hip_fatbin" in data
    assert b"amdgcn-amd-amdhsa--gfx950" in data


@pytest.mark.skipif(F.device_count() > 0, reason="a GPU is visible")
def test_no_cpu_fallback_without_gpu():
    from tests import systems
    with pytest.raises(RuntimeError, match="device"):
        F.Engine(systems.primitive_model(n_pairs=5, box=40.0)[0])


def test_cpp_adapters_compile_standalone():
    """The EnergyTerm-shaped C++ adapters build against the C header alone."""
    src = os.path.join(ROOT, "faunus_amd", "host", "gpu_energy.hpp")
    out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-x", "c++", src],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
