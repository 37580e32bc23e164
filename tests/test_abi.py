"""CPU tests of the C-ABI boundary: the library loads, exports every function
declared in include/*.h, and rejects invalid parameters before touching a
device (reference error behaviour: std::invalid_argument in
path_integral.C:212-213 / temperature.h:86-100 -> LG_ERR_ARG)."""

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(lg_[A-Za-z0-9_]+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lg_chain_create" in names and "lg_chain_sweep" in names and "lg_batch_run" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    import loopgpu
    lib = ctypes.CDLL(loopgpu.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(loopgpu.EXPORTED) <= set(declared_functions())


@pytest.mark.parametrize("kw,frag", [
    (dict(lattice=2, L=8, beta=1.0), "lattice"),
    (dict(lattice=0, L=7, beta=1.0), "even"),
    (dict(lattice=0, L=0, beta=1.0), "even"),
    (dict(lattice=1, L=8, beta=-1.0), "negative temperature"),
    (dict(lattice=1, L=8, beta=float("inf")), "beta"),
])
def test_invalid_parameters_rejected_without_device(kw, frag):
    import loopgpu
    p = loopgpu.params(kw["lattice"], kw["L"], kw["beta"])
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Chain(p)
    assert e.value.code == -1
    assert frag in str(e.value)


def test_record_layout_matches_header():
    import loopgpu
    txt = open(os.path.join(ROOT, "include", "loopgpu.h")).read()
    body = txt[txt.index("typedef struct lg_sweep_record"):txt.index("} lg_sweep_record;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in re.findall(r"double\s+([^;]+);", body):
        fields += [f.strip() for f in decl.split(",")]
    assert fields == loopgpu.REC_FIELDS
