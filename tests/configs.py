"""The five benchmark workloads at their full sizes, shared by the golden-digest script
(tests/golden/make_config_digests.py, which runs the oracle on them in this container) and the
config-size GPU parity tests (tests/test_gpu_configs.py, which hash the device outputs).

  C1  1024x1024 uint8 (seed 1): Dilate / Erode by SESquare2D, iter 1
  C2  8192x8192 uint8 (seed 2): dilation, erosion, opening, closing, gradient by squares of
      side 3, 7, 15, 31, 63; 512^3 uint8 (seed 3): SECross3D and SESquare3D, iter 1..3
  C3  4096x4096 blurred-noise mask (seed 4): reconstruction by dilation of max(mask - 32, 0)
      (= h_maxima(mask, 32)), SESquare2D; leveling(mask, opening_7(mask)) (15x15 opening)
  C4  512 uint8 images (seed 5): dilation and opening by squares of side 3..31; 256
      blurred-noise masks (seed 6): reconstruction by dilation of max(mask - 16, 0)
  C5  [64, 32, 256, 256] float32 (seed 7): flat 3x3 / 7x7 dilation and erosion, 5x5 max-plus
      dilation (w ~ N(0, 0.1), seed 8), forward and backward with an integer-valued output
      gradient (seed 9), so that every gradient sum is exact in any order
"""
import hashlib

import numpy as np

SIDES = [3, 7, 15, 31, 63]
OPS_2D = ["dilate", "erode", "opening", "closing", "gradient"]
SES_3D = ["SECross3D", "SESquare3D"]
ITERS_3D = [1, 2, 3]
C5_SHAPE = (64, 32, 256, 256)
# (name, k, with w, dilate)
C5_OPS = [("dilate3", 3, False, True), ("erode3", 3, False, False), ("dilate7", 7, False, True),
          ("erode7", 7, False, False), ("maxplus5", 5, True, True)]


def digest(a):
    """sha256 of the bytes of a C-contiguous array (floats: -0.0 folded into +0.0)"""
    a = np.ascontiguousarray(a)
    if a.dtype.kind == "f":
        a = a + a.dtype.type(0)
    return hashlib.sha256(a.data).hexdigest()


def digest_of_digests(ds):
    return hashlib.sha256("".join(ds).encode()).hexdigest()


def c5_input():
    from poutre2_amd import synth
    return synth.normal_f32(C5_SHAPE, 7)


def c5_w():
    from poutre2_amd import synth
    return (0.1 * synth.rng(8).standard_normal((5, 5))).astype(np.float32)


def c5_gy():
    """output gradient with values in {-3..3}.  gx[p] sums at most k*k of them, so it is an
    exact float32 integer in any summation order.  gw[t] sums millions of them; with signs
    drawn at random every partial sum over a block of outputs stays a random walk of a few
    thousand, far below 2^24, so the device's blockwise float32 sums are exact too."""
    from poutre2_amd import synth
    return synth.rng(9).integers(-3, 4, size=C5_SHAPE, dtype=np.int8)
