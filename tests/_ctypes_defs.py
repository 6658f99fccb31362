"""ctypes mirrors of the C structs in include/stt.h (shared by the tests)."""
import ctypes as C


class SttGeometry(C.Structure):
    _fields_ = [("vol", C.c_int * 3), ("spacing", C.c_float * 3), ("det", C.c_int * 2),
                ("ds", C.c_double), ("sid", C.c_double), ("sdd", C.c_double),
                ("oxyz", C.c_double * 3)]


class SttParams(C.Structure):
    _fields_ = [("sigma", C.c_double), ("tau", C.c_double), ("lambda_tv", C.c_double),
                ("k_tp", C.c_double), ("e_bc", C.c_double), ("huber", C.c_double),
                ("alpha", C.c_double), ("g_max", C.c_float), ("prior", C.c_int),
                ("bp", C.c_int), ("interp", C.c_int), ("roi_cut", C.c_int * 8),
                ("roi_jtop", C.c_int), ("fp64_reductions", C.c_int)]


def geometry(vol=(32, 32, 32), spacing=(1.0, 1.0, 1.0), det=(48, 48), ds=1.0,
             sid=600.0, sdd=1000.0, oxyz=(0.0, 0.0, 0.0)):
    g = SttGeometry()
    g.vol[:] = list(vol)
    g.spacing[:] = list(spacing)
    g.det[:] = list(det)
    g.ds, g.sid, g.sdd = ds, sid, sdd
    g.oxyz[:] = list(oxyz)
    return g
