"""ctypes mirror of ``include/bcmgx.h`` (data layouts and enum values only).

The ``bcmgx_params`` structure mirrors ``struct params``
(BCMGX/src/config/Params.h:238-291) followed by this build's extension keys.
"""
import ctypes as C

# status codes
OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_ALLOC = -3
ERR_IO = -4
ERR_PARAM = -5
ERR_UNSUPPORTED = -6
ERR_NUMERIC = -7
ERR_STATE = -8

MEM_HOST = 0
MEM_DEVICE = 1

# enumerations, reference order (Params.h:21-212)
SOLVER = {"CGHS": 0, "FCG": 1, "CGS": 2, "PIPELINED_CGS": 3, "CGS_CUBLAS": 4, "INVALID": 5}
BOOTSTRAP = {"MULTIPLICATIVE": 0, "SYMMETRIZED_MULTIPLICATIVE": 1, "ADDITIVE": 2, "INVALID": 3}
MATCH = {"SUITOR": 0, "INVALID": 1}
CYCLE = {"V_CYCLE": 0, "H_CYCLE": 1, "W_CYCLE": 2, "VARIABLE_V_CYCLE": 3, "INVALID": 4}
COARSE = {"L1_JACOBI": 0, "INVALID": 1}
RELAX = {"L1_JACOBI": 0, "INVALID": 1}
PREC = {"NONE": 0, "L1_JACOBI": 1, "BCMG": 2, "AFSAI": 3, "INVALID": 4}
PRECISION = {"FP64": 0, "FP32": 1}


class Params(C.Structure):
    _fields_ = [
        ("mem_alloc_size", C.c_int),
        ("error", C.c_int),
        ("solver", C.c_int),
        ("itnlim", C.c_int),
        ("rtol", C.c_double),
        ("dispnorm", C.c_int),
        ("preconditioner", C.c_int),
        ("bootstrap_type", C.c_int),
        ("max_hrc", C.c_int),
        ("conv_ratio", C.c_double),
        ("matchtype", C.c_int),
        ("aggrsweeps", C.c_int),
        ("aggr_type", C.c_int),
        ("max_levels", C.c_int),
        ("cycle_type", C.c_int),
        ("coarse_solver", C.c_int),
        ("relax_type", C.c_int),
        ("relaxnumber_coarse", C.c_int),
        ("prerelax_sweeps", C.c_int),
        ("postrelax_sweeps", C.c_int),
        ("l1jacsweeps", C.c_int),
        ("sstep", C.c_int),
        ("stop_criterion", C.c_int),
        ("ru_res", C.c_int),
        ("rec_res_int", C.c_int),
        ("max_coarse_size", C.c_int),
        ("precision", C.c_int),
        ("use_graph", C.c_int),
        ("autotune", C.c_int),
        ("check_every", C.c_int),
        ("fcg_z32", C.c_int),
        ("post_galerkin", C.c_int),
    ]


class PrecInfo(C.Structure):
    _fields_ = [
        ("ptype", C.c_int),
        ("num_levels", C.c_int),
        ("op_cmplx", C.c_double),
        ("op_wcmplx", C.c_double),
        ("avg_cratio", C.c_double),
        ("max_coarse_size", C.c_int),
        ("setup_seconds", C.c_double),
        ("g_nnz", C.c_longlong),
        ("num_hrc", C.c_int),
        ("estimated_ratio", C.c_double),
    ]


class SolverOut(C.Structure):
    _fields_ = [
        ("retv", C.c_int),
        ("niter", C.c_int),
        ("exit_res", C.c_double),
        ("rel_res", C.c_double),
        ("del0", C.c_double),
        ("solve_seconds", C.c_double),
        ("full_n", C.c_int),
        ("local_n", C.c_int),
    ]


def reference_defaults():
    """The defaults of ``struct params`` (Params.h:238-291) plus extension defaults."""
    p = Params()
    p.mem_alloc_size = 2000000
    p.error = 0
    p.solver = SOLVER["CGHS"]
    p.itnlim = 2000
    p.rtol = 1.0e-6
    p.dispnorm = 1
    p.preconditioner = PREC["NONE"]
    p.bootstrap_type = BOOTSTRAP["MULTIPLICATIVE"]
    p.max_hrc = 1
    p.conv_ratio = 0.8
    p.matchtype = MATCH["SUITOR"]
    p.aggrsweeps = 2
    p.aggr_type = 0
    p.max_levels = 39
    p.cycle_type = CYCLE["V_CYCLE"]
    p.coarse_solver = COARSE["L1_JACOBI"]
    p.relax_type = RELAX["L1_JACOBI"]
    p.relaxnumber_coarse = 20
    p.prerelax_sweeps = 4
    p.postrelax_sweeps = 4
    p.l1jacsweeps = 4
    p.sstep = 1
    p.stop_criterion = 1
    p.ru_res = 1
    p.rec_res_int = 0
    p.max_coarse_size = 0
    p.precision = PRECISION["FP64"]
    p.use_graph = 1
    p.autotune = 1
    p.check_every = 8
    p.fcg_z32 = 1
    p.post_galerkin = 1
    return p
