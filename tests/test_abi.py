"""The C-ABI library loads and exports every entry point declared in include/pfem.h
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pfem.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pfem_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("pfem_create", "pfem_assemble", "pfem_linear_solve", "pfem_newton", "pfem_impose_dirichlet"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from plasticity_amd import _lib
    L = ctypes.CDLL(_lib.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding knows every declared symbol
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_defaults_and_status_strings_need_no_gpu():
    from plasticity_amd import _lib, default_options
    o = default_options()
    assert o.cg_itmax == 200 and o.eta_min == 1e-8 and o.eta_max == 0.1  # Solver.jl:103-106
    assert o.ew_gamma == 0.9 and o.ew_alpha == 1.5
    assert _lib.lib().pfem_status_string(1).decode().startswith("inverted")
