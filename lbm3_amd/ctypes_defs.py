"""ctypes mirrors of include/lbm3mi_types.h (the plain C data of the C-ABI)."""
import ctypes as C

MAX_LEVELS = 8
MAX_RULES = 16

FACE_LOW, FACE_MID, FACE_HIGH, FACE_ANY = 1, 2, 4, 7
SET_VELOCITY, SET_ROTATION, SET_FORCE, SET_LIMITER, SET_NU_MULTIPLIER, SET_USE_IBB = 1, 2, 4, 8, 16, 32
INIT_UNIFORM, INIT_TAYLOR_GREEN = 0, 1
REFINE_SET, REFINE_ADD, REFINE_REMOVE = 0, 1, 2
STORE_F32, STORE_F16, STORE_BF16, STORE_F64, STORE_F16C = 0, 1, 2, 3, 4


class Bounds(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("xMin", "yMin", "zMin", "xMax", "yMax", "zMax", "rxMax", "ryMax", "rzMax")]


class Info(C.Structure):
    _fields_ = [("gridID", C.c_float), ("gridMemoryBytes", C.c_ulonglong),
                ("iterationsFinished", C.c_int), ("iterationsSinceTrackerReset", C.c_int),
                ("res", C.c_float), ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
                ("Bounds", Bounds), ("nu", C.c_float), ("dtPhys", C.c_float),
                ("cellCountX", C.c_int), ("cellCountY", C.c_int), ("cellCountZ", C.c_int),
                ("cellCount", C.c_int), ("esotwistFlipper", C.c_int), ("useRotors", C.c_int),
                ("iRegulatorInlet", C.c_float), ("iRegulatorOutlet", C.c_float)]


class BC(C.Structure):
    _fields_ = [("dRho", C.c_float), ("ux", C.c_float), ("uy", C.c_float), ("uz", C.c_float),
                ("gx", C.c_float), ("gy", C.c_float), ("gz", C.c_float),
                ("wallID", C.c_int), ("openBCID", C.c_int), ("useIBB", C.c_int),
                ("dirichletU", C.c_int), ("dirichletRho", C.c_int), ("periodic", C.c_int),
                ("nonReflective", C.c_int), ("nonReflectiveBeta", C.c_float),
                ("collisionLimiter", C.c_float), ("nuMultiplier", C.c_float)]


def default_bc(**kw):
    """BCStruct defaults (reference include/types.h:131-136) with overrides."""
    b = BC(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1, 0, 1, 0, 0, 0, 1, 7.0, 0.05, 1.0)
    for k, v in kw.items():
        setattr(b, k, v)
    return b


class Profile(C.Structure):
    """max(0, 1 - r^2 / radius^2) about the line through origin along axis; radius <= 0: none."""
    _fields_ = [("axis", C.c_int), ("origin", C.c_float * 3), ("radius", C.c_float)]


class OpenRule(C.Structure):
    _fields_ = [("levelMask", C.c_int), ("faceMask", C.c_int * 3), ("bc", BC), ("profile", Profile)]


class LocalRule(C.Structure):
    _fields_ = [("levelMask", C.c_int), ("wallID", C.c_int), ("set", C.c_int),
                ("ux", C.c_float), ("uy", C.c_float), ("uz", C.c_float),
                ("omega", C.c_float * 3), ("origin", C.c_float * 3),
                ("gx", C.c_float), ("gy", C.c_float), ("gz", C.c_float),
                ("collisionLimiter", C.c_float), ("nuMultiplier", C.c_float), ("useIBB", C.c_int)]


class Init(C.Structure):
    _fields_ = [("kind", C.c_int), ("dRho", C.c_float), ("ux", C.c_float), ("uy", C.c_float), ("uz", C.c_float),
                ("u0", C.c_float), ("length", C.c_float), ("pressure", C.c_int), ("profile", Profile)]


class RefineRule(C.Structure):
    _fields_ = [("level", C.c_int), ("mode", C.c_int)] + \
               [(n, C.c_float) for n in ("xMin", "yMin", "zMin", "xMax", "yMax", "zMax")] + [("cylinder", Profile)]


OPEN_CHARACTERISTIC, OPEN_IMPEDANCE = 0, 1
INTERFACE_QUADRATIC, INTERFACE_LINEAR = 0, 1
MAX_ROTORS = 8
AXIS_NONE, AXIS_X, AXIS_Y, AXIS_Z = -1, 0, 1, 2


class Case(C.Structure):
    _fields_ = [("gridLevelCount", C.c_int), ("wallRefinementCount", C.c_int),
                ("resGlobal", C.c_float), ("dtPhysGlobal", C.c_float), ("nuPhys", C.c_float), ("rhoPhys", C.c_float),
                ("trackWallForce", C.c_int), ("trackRotorForce", C.c_int), ("trackOpenBoundaries", C.c_int),
                ("init", Init),
                ("nOpenRules", C.c_int), ("openRules", OpenRule * MAX_RULES), ("openDefault", BC),
                ("nLocalRules", C.c_int), ("localRules", LocalRule * MAX_RULES),
                ("nRefineRules", C.c_int), ("refineRules", RefineRule * MAX_RULES),
                ("storeRef", C.c_float * 3), ("openScheme", C.c_int), ("rhoReflectionTolerance", C.c_float),
                ("interfaceScheme", C.c_int)]
