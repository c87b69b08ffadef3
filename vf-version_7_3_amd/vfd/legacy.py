"""ctypes mirror of the reference-compatible interface (include/vf_diffusion.h)."""
import ctypes

from . import lib, _ptr

N_CHEM = 8
_f, _i, _p = ctypes.c_float, ctypes.c_int, ctypes.c_void_p


class CCtx(ctypes.Structure):
    """c_ctx (convolutionFFT_common.h:305-346), identical field order."""
    _fields_ = [("dt", _f), ("dx2", _f), ("kernelRadius", _i), ("niter", _i),
                ("KD", _i), ("KH", _i), ("KW", _i),
                ("KZ", _i), ("KX", _i), ("KY", _i),
                ("DD", _i), ("DH", _i), ("DW", _i),
                ("FFTD", _i), ("FFTH", _i), ("FFTW", _i),
                ("windowD", _i), ("windowH", _i), ("windowW", _i),
                ("h_ibuffs", _p * N_CHEM), ("h_obuffs", _p * N_CHEM),
                ("gpu_id", _i * N_CHEM), ("dev_id", _i * N_CHEM),
                ("d_data", _p * N_CHEM), ("h_data", _p * N_CHEM),
                ("h_kernelspectrum", _p * N_CHEM), ("d_kernelspectrum_h", _p * N_CHEM)]


lib.snapTransformSize.argtypes = [_i]
lib.snapTransformSize.restype = _i
lib.fftDiffuse3D.argtypes = [_p, _p, _i, _i, CCtx, _i, _f, _i, _i]
lib.fftDiffuse3D.restype = ctypes.c_bool
lib.computeKernel3D.argtypes = [_i, _f, _f, _f, CCtx, _i, _i, ctypes.c_short]
lib.computeKernel3D.restype = ctypes.c_bool
lib.computeKernelSpectrum3D.argtypes = [CCtx, CCtx, ctypes.c_short]
lib.computeKernelSpectrum3D.restype = ctypes.c_bool
lib.computeKernel3DBatch.argtypes = [_i, ctypes.POINTER(_f), ctypes.POINTER(_f), _f, CCtx, CCtx]
lib.computeKernel3DBatch.restype = ctypes.c_bool
# the 2D operator (the reference's !MODEL_3D build)
lib.computeKernel.argtypes = [_p, _i, _f, _f, _f, CCtx]
lib.computeKernel.restype = ctypes.c_bool
lib.computeKernelSpectrum.argtypes = [_p, _p, CCtx, CCtx]
lib.computeKernelSpectrum.restype = ctypes.c_bool
lib.fftDiffuse2D.argtypes = [_p, _p, CCtx, _i, _f]
lib.fftDiffuse2D.restype = ctypes.c_bool
lib.vfdPlan3d.argtypes = [ctypes.POINTER(_i), _i, _i, _i, _i]
lib.vfdDestroy.argtypes = [_i]
lib.H2D.argtypes = [_i, ctypes.POINTER(CCtx), _i]
lib.H2D.restype = None
lib.D2H.argtypes = [_i, ctypes.POINTER(CCtx), _i]
lib.D2H.restype = None
lib.vfdDiffuseChems.argtypes = [CCtx, _i, _i, ctypes.POINTER(_f)]
lib.vfdDiffuseChems.restype = ctypes.c_bool
lib.vfdTickChems.argtypes = [CCtx, _i, ctypes.POINTER(_f), ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(_f),
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                             ctypes.POINTER(ctypes.c_double)]
lib.vfdTickChems.restype = ctypes.c_bool
lib.vfdTickChemsBegin.argtypes = [CCtx, _i, ctypes.POINTER(_f)]
lib.vfdTickChemsBegin.restype = ctypes.c_bool
lib.vfdTickChemsEnd.argtypes = [CCtx, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(_f), ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_double)]
lib.vfdTickChemsEnd.restype = ctypes.c_bool
lib.vfdSetKernelPowerQuirk.argtypes = [_i]
lib.vfdSetKernelPowerQuirk.restype = None
for name, n in (("extractCenter3D", 8), ("unpadResult3D", 8), ("padKernel3D", 11),
                ("padDataClampToBorder3D", 14)):
    fn = getattr(lib, name)
    fn.argtypes = [_p, _p] + [_i] * (n - 2)
    fn.restype = None
lib.padDataConstantVF3D.argtypes = [_p, _p] + [_i] * 13 + [_f]
lib.padDataConstantVF3D.restype = None
for name in ("modulateAndNormalize3D", "modulate3D"):
    getattr(lib, name).argtypes = [_p, _p, _i, _i, _i, _i]
    getattr(lib, name).restype = None
lib.complexPower.argtypes = [_p, _i, _i, _i, _i, _i]
lib.complexPower.restype = None
for name in ("padDataMirror", "padDataRightWall", "padDataClampToBorder"):
    getattr(lib, name).argtypes = [_p, _p] + [_i] * 8
    getattr(lib, name).restype = None
lib.padDataConstantVF.argtypes = [_p, _p] + [_i] * 9 + [_f]
lib.padDataConstantVF.restype = None
lib.padKernel.argtypes = [_p, _p] + [_i] * 6
lib.padKernel.restype = None
lib.extractCenter.argtypes = [_p, _p, _i, _i, _i, _i]
lib.extractCenter.restype = None
lib.unpadResult.argtypes = [_p, _p, _i, _i, _i, _i]
lib.unpadResult.restype = None
lib.modulateAndNormalize.argtypes = [_p, _p, _i, _i, _i]
lib.modulateAndNormalize.restype = None
lib.convolutionClampToBorderCPU.argtypes = [_p, _p, _p] + [_i] * 6
lib.convolutionClampToBorderCPU.restype = None


def ptr(t):
    return _ptr(t)
