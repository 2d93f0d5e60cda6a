"""Loads libchfem.so (the gfx950 HIP library) and declares its C ABI.

The library must have been built (``python -m chfem_amd.build``); there is no
CPU fallback: every entry point here runs the HIP implementation or fails.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchfem.so")

_lib = None

vp = C.c_void_p
u8 = C.c_uint8
uint = C.c_uint
dbl = C.c_double
cp = C.c_char_p


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint), ("sweeps", C.c_uint), ("solve_seconds", C.c_double),
                ("residual", C.c_double), ("x0_residual", C.c_double), ("x0_used", C.c_uint)]


# (name, restype, argtypes) of every entry point declared in include/*.h
SIGNATURES = [
    # include/cudapcg.h
    ("cudapcgInit", u8, [u8, u8]),
    ("cudapcgEnd", u8, []),
    ("cudapcgSetModelConstructorFcn", u8, [vp]),
    ("cudapcgBuildModel", u8, [vp]),
    ("cudapcgNewModel", vp, []),
    ("cudapcgSetModel", u8, [vp]),
    ("cudapcgSetNumTol", u8, [dbl]),
    ("cudapcgSetMaxIterations", u8, [uint]),
    ("cudapcgSetXReduceMode", u8, [u8]),
    ("cudapcgSetXReduceShift", u8, [uint]),
    ("cudapcgSetXReduceScale", u8, [dbl]),
    ("cudapcgSetReduceStabFactor", u8, [dbl]),
    ("cudapcgSetSolver", u8, [u8]),
    ("cudapcgSetPreconditioner", u8, [u8]),
    ("cudapcgSetResNorm", u8, [u8]),
    ("cudapcgAllocateArrays", u8, []),
    ("cudapcgFreeArrays", u8, []),
    ("cudapcgSetRHS", u8, [vp]),
    ("cudapcgSetX0", u8, [vp, u8]),
    ("cudapcgSetImage", u8, [vp]),
    ("cudapcgSetParametricDensityField", u8, [vp, dbl, dbl]),
    ("cudapcgSetPoreMap", u8, [vp]),
    ("cudapcgSetPeriodic2DOFMap", u8, [vp]),
    ("cudapcgSetDOF2PeriodicMap", u8, [vp]),
    ("cudapcgSetLclMtxs", u8, [vp]),
    ("cudapcgSolve", u8, [vp]),
    ("cudapcgSetHeaderString", u8, [cp]),
    ("cudapcgPrintSolverReport", u8, []),
    ("cudapcgPrintSolverReport2", u8, [cp]),
    ("cudapcgPrintSolverMetrics", u8, []),
    ("cudapcgPrintSolverMetrics2", u8, [cp]),
    ("cudapcgGetNumIterations", uint, []),
    ("cudapcgGetMaxNumIterations", uint, []),
    ("cudapcgGetResidual", dbl, []),
    # include/femhmg.h
    ("hmgInit", u8, [cp, cp, cp, vp]),
    ("hmgEnd", u8, []),
    ("hmgSetXReduceFlag", None, [u8]),
    ("hmgSetParallelStrategyFlag", None, [u8]),
    ("hmgSetSolverFlag", None, [u8]),
    ("hmgSetPreConditionerFlag", None, [u8]),
    ("hmgSetStoppingCriteria", None, [u8]),
    ("hmgSetPoreMappingStrategy", u8, [u8]),
    ("hmgFindInitialGuesses", None, [uint]),
    ("hmgExportX", None, [u8]),
    ("hmgImportX", None, [cp]),
    ("hmgSetHomogenizationFlag", None, [C.c_uint16]),
    ("hmgSolveHomogenization", None, []),
    ("hmgPrintModelData", None, []),
    ("hmgPrintConstitutiveMtx", None, []),
    ("hmgGetConstitutiveMtx", C.POINTER(dbl), []),
    ("hmgGetConstitutiveMtxDim", uint, []),
    ("hmgGetThermalExpansion", C.POINTER(dbl), []),
    ("hmgGetThermalExpansionDim", uint, []),
    ("hmgSaveConstitutiveMtx", u8, [cp]),
    ("hmgSaveFields", None, [u8, u8]),
    ("hmgKeepTrackOfAnalysisReport", None, [u8]),
    ("hmgPrintReport", u8, [cp]),
    ("initDefaultInput", None, [vp]),
    ("runAnalysis", C.c_int, [vp]),
    # include/chfem_amd.h
    ("chfem_run", C.c_int, [vp, C.c_size_t, cp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                            C.POINTER(uint), C.POINTER(uint)]),
    ("cudapcgSetPrecision", u8, [u8]),
    ("hmgSetPrecisionFlag", None, [u8]),
    ("cudapcgGetSolveStats", u8, [C.POINTER(SolveStats)]),
    ("hmgGetSolveStats", u8, [C.POINTER(SolveStats), C.POINTER(uint)]),
    ("cudapcgApplyOperator", u8, [vp, vp]),
    ("cudapcgApplyOperatorF32", u8, [vp, vp]),
    ("cudapcgApplyPreconditioner", u8, [vp, vp]),
    ("hmgAssembleRHS", u8, [uint, vp]),
    ("hmgUpdateC", u8, [uint, vp, vp]),
    ("hmgUpdateCNoAlloc", u8, [uint, vp, vp]),
    ("hmgGetMaps", u8, [vp, vp, vp, vp, vp]),
    ("hmgGetLocalMatrices", u8, [vp, vp]),
    ("hmgGetSizes", u8, [C.POINTER(uint), C.POINTER(uint), C.POINTER(uint), C.POINTER(uint)]),
    ("hmgSetSolverTolerance", None, [dbl]),
    ("hmgSetMaxIterationsOverride", None, [uint]),
    ("hmgPrepare", u8, []),
    ("hmgSolveResident", u8, [vp]),
    ("hmgRelease", None, []),
    ("cudapcgSetKernelTiming", u8, [u8]),
    ("cudapcgGetKernelTiming", u8, [C.POINTER(dbl), C.POINTER(uint)]),
    ("hmgGetKernelTiming", u8, [C.POINTER(dbl), C.POINTER(uint)]),
    ("chfemCopyKernelBench", dbl, [C.c_size_t, uint]),
    ("cudapcgOperatorBench", dbl, [u8, uint, u8]),
    ("hmgSolveLoadCase", u8, [uint, vp, vp, vp]),
    ("chfemGuessCreate", vp, [C.c_int, uint]),
    ("chfemGuessDestroy", None, [vp]),
    ("chfemGuessNumParams", C.c_size_t, [vp]),
    ("chfemGuessNumOutputs", uint, [vp]),
    ("chfemGuessGetParams", C.c_int, [vp, vp]),
    ("chfemGuessSetParams", C.c_int, [vp, vp]),
    ("chfemGuessGetGradients", C.c_int, [vp, vp]),
    ("chfemGuessZeroGradients", C.c_int, [vp]),
    ("chfemGuessPredict", C.c_int, [vp, vp, vp, uint, uint, uint, uint, vp]),
    ("chfemGuessAccumulate", C.c_int, [vp, vp, vp, uint, uint, uint, uint, vp, vp]),
    ("chfemGuessStep", C.c_int, [vp, C.c_float, C.c_float]),
    ("hmgSetInitialGuessNetwork", C.c_int, [vp]),
    ("chfemGuessMfmaCheck", C.c_int, [vp, vp, vp]),
    ("chfemMfmaCheckF", C.c_int, [vp, vp, vp, vp, vp, vp]),
    ("chfemMfmaCheck4", C.c_int, [vp, vp, vp]),
    ("cudapcgGetOperatorTile", u8, [u8, vp]),
    ("cudapcgGetLoopForm", C.c_int, [u8]),
    ("chfemBatchCreate", vp, []),
    ("chfemBatchDestroy", None, [vp]),
    ("chfemBatchAddSample", C.c_int, [vp, cp, vp]),
    ("chfemBatchSolve", C.c_int, [vp, dbl, uint]),
    ("chfemBatchGetC", C.c_int, [vp, uint, vp]),
    ("chfemBatchGetStats", C.c_int, [vp, uint, vp, C.POINTER(C.c_uint)]),
    ("chfemBatchSeconds", dbl, [vp]),
    ("chfemBatchSetInitialGuessNetwork", C.c_int, [vp, vp]),
]


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m chfem_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class CudapcgModel(C.Structure):
    """cudapcgModel_t (include/cudapcg.h; chfem/src/femhmg/cudapcg/includes.h:93-129)."""
    _fields_ = [("name", C.c_void_p), ("freeAllowed_flag", C.c_uint8), ("parStrategy_flag", C.c_uint8),
                ("poremap_flag", C.c_uint8), ("parametric_density_field_flag", C.c_uint8),
                ("SBS_flag", C.c_uint8), ("nrows", C.c_uint), ("ncols", C.c_uint), ("nlayers", C.c_uint),
                ("nelem", C.c_uint), ("nvars", C.c_uint), ("nvarspernode", C.c_uint), ("nkeys", C.c_uint),
                ("localmtxdim", C.c_uint), ("nporenodes", C.c_uint), ("nbordernodes", C.c_uint),
                ("nhmgvars", C.c_uint), ("limits_density_field", C.c_double * 2), ("image", C.c_void_p),
                ("parametric_density_field", C.c_void_p), ("pore_map", C.c_void_p),
                ("border_pore_map", C.c_void_p), ("periodic2DOF_map", C.c_void_p),
                ("DOF2periodic_map", C.c_void_p)]
