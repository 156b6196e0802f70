"""Host logic of the product library (no GPU calls): the C ABI loads and
exports every declared entry point, the .properties parser mirrors
Params::initFromPropertiesFile, and the host generators / MTX reader equal the
oracle's."""
import ctypes as C
import os

import numpy as np
import pytest

import bootcmatchgx_amd as bx
import oracle_py as O
from bootcmatchgx_amd import abi


def test_library_exports_every_declared_symbol():
    L = bx.lib()
    names = bx.declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n


def test_defaults_match_reference_struct():
    p = bx.Params()
    ref = abi.reference_defaults()
    for f, _ in abi.Params._fields_:
        assert getattr(p.s, f) == getattr(ref, f), f
    # Params.h:250-275
    assert p.rtol == 1e-6 and p.itnlim == 2000 and p.aggrsweeps == 2 and p.max_levels == 39
    assert p.relaxnumber_coarse == 20 and p.prerelax_sweeps == 4 and p.postrelax_sweeps == 4
    assert p.solver == abi.SOLVER["CGHS"] and p.preconditioner == abi.PREC["NONE"]


@pytest.mark.parametrize("fname,solver,prec", [
    ("FCG_BCMG.properties", "FCG", "BCMG"),
    ("FCG_AFSAI.properties", "FCG", "AFSAI"),
    ("FCG_l1Jacobi.properties", "FCG", "L1_JACOBI"),
    ("FCG_noPreconditioner.properties", "FCG", "NONE"),
])
def test_reference_settings_files(golden, fname, solver, prec):
    p = bx.Params.init_from_properties_file(os.path.join(golden, fname))
    assert p.solver == abi.SOLVER[solver] and p.preconditioner == abi.PREC[prec]
    assert p.rtol == 1e-6


def test_sample_properties(golden):
    p = bx.Params.init_from_properties_file(os.path.join(golden, "sample.properties"))
    assert p.solver == abi.SOLVER["CGHS"] and p.preconditioner == abi.PREC["AFSAI"]
    assert p.itnlim == 2000 and p.relaxnumber_coarse == 20 and p.aggrsweeps == 2 and p.l1jacsweeps == 4
    # `coarsesolver_type` is not a properties key of the reference: ignored
    q = bx.Params.init_from_properties_file(os.path.join(golden, "FCG_BCMG.prova.properties"))
    assert q.itnlim == 1000 and q.preconditioner == abi.PREC["BCMG"] and q.prerelax_sweeps == 4


def test_properties_quirks(tmp_path):
    f = tmp_path / "a.properties"
    f.write_text("solver = FCG\n  \n# comment\nrtol = 1.e-8 # trailing\nunknown_key = 3\n"
                 "preconditioner=BCMG\nmax_coarse_size = 99\nprecision = FP32\n")
    p = bx.Params.init_from_properties_file(str(f))
    assert p.solver == abi.SOLVER["FCG"] and p.rtol == 1e-8 and p.preconditioner == abi.PREC["BCMG"]
    assert p.max_coarse_size == 99 and p.precision == abi.PRECISION["FP32"]
    # an enum followed by a comment keeps the comment in the value (Params.cu:311-316): INVALID
    g = tmp_path / "b.properties"
    g.write_text("solver = FCG # x\n")
    assert bx.Params.init_from_properties_file(str(g)).solver == abi.SOLVER["INVALID"]
    # a non-empty line without '=' is an error (Params.cu:306-309), reported as a status
    h = tmp_path / "c.properties"
    h.write_text("solver FCG\n")
    with pytest.raises(bx.BcmgxError) as e:
        bx.Params.init_from_properties_file(str(h))
    assert e.value.status == abi.ERR_IO
    with pytest.raises(bx.BcmgxError):
        bx.Params.init_from_properties_file(str(tmp_path / "missing.properties"))


def test_enum_strings():
    for kind, tab in [("solver", abi.SOLVER), ("preconditioner", abi.PREC), ("cycle_type", abi.CYCLE)]:
        for name, v in tab.items():
            assert bx.enum_string(kind, v) == name
            assert bx.lib().bcmgx_enum_from_string(kind.encode(), name.encode()) == v


@pytest.mark.parametrize("args", [(7, 9, 8, 7, 1, 1, 1), (27, 6, 5, 4, 1, 1, 1), (7, 2, 2, 5, 3, 4, 2), (27, 5, 5, 5, 1, 2, 2),
                                  (27, 5, 5, 5, 4, 2, 1), (7, 4, 14, 4, 2, 2, 1)])
def test_host_generator_equals_oracle(args):
    n, row, col, val = bx.laplacian3d(*args)
    r2, c2, v2 = O.laplacian3d(*args).arrays()
    assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(val, v2)


@pytest.mark.parametrize("kind,p1,p2,seed", [(0, 2.0, 0.0, 2), (1, 0.5, 2.0, 4)])
def test_host_diffusion_equals_oracle(kind, p1, p2, seed):
    n, row, col, val = bx.diffusion7p(7, 6, 5, kind, p1, p2, seed)
    r2, c2, v2 = O.diffusion7p(7, 6, 5, kind, p1, p2, seed).arrays()
    assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(val, v2)


def test_host_mtx_reader_equals_oracle(golden):
    n, m, row, col, val = bx.read_mtx(os.path.join(golden, "bcsstk09.mtx"))
    r2, c2, v2 = O.read_mtx(os.path.join(golden, "bcsstk09.mtx")).arrays()
    assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(val, v2)
    with pytest.raises(bx.BcmgxError):
        bx.read_mtx(os.path.join(golden, "nonexistent.mtx"))
