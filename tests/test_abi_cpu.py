"""CPU-side checks of the C ABI: the library loads and exports every declared
symbol; host-side logic that needs no GPU."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "issampler.h")
LIB = os.path.join(ROOT, "importancesamplers.jl_amd", "libissampler.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|const char\*)\s+(is_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("is_ctx_create", "is_job_create", "is_job_run", "is_job_get_logweights", "is_job_resample",
                 "is_rng_raw", "is_cdf_device", "is_select_device"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libissampler.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_matches_header(pkg):
    from importlib import import_module

    sigs = pkg._lib.SIGNATURES
    assert sorted(sigs) == declared_symbols()


def test_descriptor_layout_matches_header(pkg):
    # is_job_desc field order / types as declared in include/issampler.h
    text = open(HEADER).read()
    body = text[text.index("typedef struct is_job_desc {"):text.index("} is_job_desc;")]
    names = re.findall(r"\b(\w+);", body)
    assert names == [f[0] for f in pkg._lib.JobDesc._fields_]


def _header_desc_fields():
    """(name, C type) of is_job_desc, in order, from include/issampler.h."""
    text = open(HEADER).read()
    body = text[text.index("typedef struct is_job_desc {") + len("typedef struct is_job_desc {"):
                text.index("} is_job_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        m = re.match(r"(.*?)\s*(\w+)$", decl)
        out.append((m.group(2), m.group(1).replace(" *", "*").replace("const ", "")))
    return out


# Julia field type for each C type of the descriptor
_JULIA_OF_C = {"int32_t": "Int32", "int64_t": "Int64", "uint64_t": "UInt64", "double": "Float64",
               "int64_t*": "Ptr{Int64}", "double*": "Ptr{Float64}", "int32_t*": "Ptr{Int32}",
               "is_batch_target_fn": "Ptr{Cvoid}", "void*": "Ptr{Cvoid}",
               "is_transform_block*": "Ptr{TransformBlock}"}


def test_integration_julia_descriptor_matches_header():
    """The Julia JobDesc that INTEGRATION.md documents mirrors every is_job_desc field,
    in order and with the C type's Julia counterpart (a short struct would make
    is_job_create read past the Julia object)."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = text[text.index("struct JobDesc"):]
    body = body[body.index("\n") + 1: body.index("\nend")]
    julia = [(m.group(1), m.group(2)) for m in re.finditer(r"^\s*(\w+)::([\w{}]+)", body, re.M)]
    header = _header_desc_fields()
    assert [n for n, _ in julia] == [n for n, _ in header]
    assert [t for _, t in julia] == [_JULIA_OF_C[c] for _, c in header]


def test_proposal_constructors_and_constants(pkg, oracle_mod):
    import math

    for nu in (0.5, 1.0, 3.0, 5.0, 40.0):
        for d in (1, 2, 3, 10, 51):
            assert pkg.student_t_logconstant(nu, d) == oracle_mod.student_t_logconstant(nu, d)
    p = pkg.FactorGaussian([0.0, 0.0], [[2.0, 0.0], [0.5, 1.5]])
    assert p.lognormalizer == pytest.approx(-math.log(2 * math.pi) - math.log(3.0), rel=1e-15)
    with pytest.raises(ValueError):
        pkg.FactorGaussian([0.0, 0.0], [[2.0, 1.0], [0.5, 1.5]])
    with pytest.raises(ValueError):
        pkg.SphericalGaussian([0.0, np.inf], 1.0)
    with pytest.raises(ValueError):
        pkg.AMIS(pkg.SphericalStudentT(2.0, [0.0], 1.0), rounds=2, round_size=4)
    with pytest.raises(ValueError):
        pkg.AMIS(pkg.SphericalGaussian([0.0], 1.0), rounds=2, round_size=[4])


def test_graft_entry_loads_package():
    import __graft_entry__

    pkg = __graft_entry__.load_package()
    assert hasattr(pkg, "importance_sample")
