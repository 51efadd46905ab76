"""Algorithmic work per unit: the fixed figures roofline fractions are computed from.

These are frozen conventions (they do not follow the implementation's actual
traffic): bytes and flops a kernel cannot avoid for one unit of work.

Beam_1 assembly, per element and Newton iteration (FP64 tangent):
    node displacements  3 x 6 x 8     =   144 B
    Gauss-point state   2 x 15 x 8    =   240 B   (Q_i, kappa_i, z'_i)
    element constants   21 x 8        =   168 B   (T3, D, e3r, L/2, rho A)
    connectivity + DOF  3x4 + 18x4    =    84 B
    tangent scatter     324 x 8 x 2   =  5184 B   (read-modify-write of the global entries)
    residual scatter    18 x 8 x 2    =   288 B
                                         ------
                                         6108 B
    flops: 1.5e4 (structure-aware count of the Gauss-point kinematics, the 9x9
    operator and the nine 6x6 node blocks; the dense products as the reference
    writes them are 7.4e4).

Shell_1 assembly, per element and Newton iteration (FP64 tangent):
    node displacements  6 x 6 x 8         =   288 B
    Gauss-point state   3 x 24 x 8        =   576 B
    element constants                     ~   300 B
    tangent scatter     729 x 8 x 2       = 11664 B
    residual scatter    27 x 8            =   216 B
                                            -------
                                            13044 B

Band factorisation (half-bandwidth b, n unknowns, element size s):
    flops 2 n b^2, bytes 2 n (2b+1) s  (read and write the band once)
Triangular solves per right-hand side:
    flops 4 n b,   bytes n (2b+1) s
"""

BEAM_BYTES_FP64 = 6108
BEAM_BYTES_FP32_TANGENT = 144 + 240 + 168 + 84 + 324 * 4 * 2 + 288
BEAM_FLOPS = 1.5e4
SHELL_BYTES_FP64 = 288 + 576 + 300 + 729 * 8 * 2 + 27 * 8

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6      # dense FP64 (vector and MFMA)
FP32_PEAK_TFLOPS = 157.3     # dense FP32


def band_lu(n, b, elem=8):
    return {"flops": 2.0 * n * b * b, "bytes": 2.0 * n * (2 * b + 1) * elem}


def band_trsv(n, b, elem=8):
    return {"flops": 4.0 * n * b, "bytes": 1.0 * n * (2 * b + 1) * elem}
