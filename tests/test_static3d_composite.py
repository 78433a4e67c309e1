"""The reference's static SESquare3D / SECross3D dilation and erosion
(include/poutre/low_level_morpho/details/ero_dil_static_se_t.hpp:761-915), restated LITERALLY
on the CPU to show what it returns on volumes whose slices differ.

Those two dispatchers build the 3D operator from 2D slice operators and keep five 2D
temporaries.  Each temporary is an ``image_t`` that owns a ``std::vector``: ``swap`` exchanges
the buffers of two temporaries (image_t.hpp:261-270) and ``a = b`` copies b's values into a's
own buffer (image_t.hpp:164, the defaulted copy assignment).  The slice copies, however, go
through raw pointers ``ptmp1/2/3/L`` taken from the temporaries BEFORE the loop
(:791-794, :873-876): after the first ``swap`` those pointers no longer point at the buffer
the temporary of the same name owns.  The restatement below models exactly that: a set of
buffers, five names bound to buffers, four raw pointers bound to buffers, and the statements
of the two dispatchers in their order.  Slices are taken the way t_3DCopyOneSliceTo2D
(:86-101) addresses memory, with xsize = bound[0], ysize = bound[1], zsize = bound[2]
(:777-779).

Findings (checked below):
* on the 3 x 3 x 3 z-constant volumes of the reference's 3D KATs
  (test/low_level_morpho/ero_dil_static_se_t.cpp:502-645) the literal code returns the
  mathematical 3D dilation / erosion, so the reference's own tests pass;
* on z-varying volumes it does not.  SESquare3D: slice 0 is right, slice 1 combines the 2D
  operator on slices 0 and 1 only (slice 2 is missing), later slices read stale or
  twice-dilated temporaries.  SECross3D: slice 0 is the 2D cross of slice 0 alone (slice 1
  is missing);
* z-constant cubes larger than the KATs' diverge too: SESquare3D at every side from 4 up
  (stale temporaries from slice 2 on), SECross3D at side 4.

The library and the oracle compute the mathematical definition (the maximum / minimum over
the 27- or 7-voxel neighbourhood), which is also what the reference's runtime neighbour-list
path returns (ero_dil_runtime_nl_se_t.hpp:48-97) for the same SE.
"""
import numpy as np
import pytest

import oracle
from tests import kats


class _Machine:
    """Buffers, image_t names bound to buffers, and raw pointers bound to buffers."""

    def __init__(self, shape2d, dtype):
        self.buf = {k: np.zeros(shape2d, dtype) for k in ("B1", "B2", "B3", "BL", "BT")}
        self.obj = {"imtmp1": "B1", "imtmp2": "B2", "imtmp3": "B3", "imtmpL": "BL", "imtmp": "BT"}
        self.ptr = {"ptmp1": "B1", "ptmp2": "B2", "ptmp3": "B3", "ptmpL": "BL"}

    def img(self, name):
        return self.buf[self.obj[name]]

    def swap(self, a, b):  # image_t::swap exchanges the storage vectors
        self.obj[a], self.obj[b] = self.obj[b], self.obj[a]

    def assign(self, a, b):  # image_t copy assignment: values into a's own buffer
        self.img(a)[...] = self.img(b)

    def op2d(self, se, dil, src, dst):  # HelperOp::EroDilOp2D(view(src), view(dst))
        self.img(dst)[...] = oracle.static(self.img(src).copy(), se, dil)

    def arith(self, dil, a, b, dst):  # t_ArithSup / t_ArithInf, elementwise (dst may be a)
        f = np.maximum if dil else np.minimum
        self.img(dst)[...] = f(self.img(a), self.img(b))

    def rotate_ptrs(self):
        p = self.ptr
        p["ptmp1"], p["ptmp2"], p["ptmp3"] = p["ptmp2"], p["ptmp3"], p["ptmp1"]


def reference_static3d(vol, se3d, dil):
    """Literal restatement of t_ErodeDilateDispatcher<SESquare3D|SECross3D, ..., 3>
    (ero_dil_static_se_t.hpp:761-839 and :841-915)."""
    b0, b1, b2 = vol.shape
    xsize, ysize, zsize = b0, b1, b2
    flat_in = np.ascontiguousarray(vol).reshape(-1)
    flat_out = np.zeros_like(flat_in)
    se2d = {"SESquare3D": "SESquare2D", "SECross3D": "SECross2D"}[se3d]
    m = _Machine((xsize, ysize), vol.dtype)

    def copy_to(slice_no, ptr):  # t_3DCopyOneSliceTo2D(psrc, x, y, x, y, i, ptr)
        n = xsize * ysize
        m.buf[m.ptr[ptr]].reshape(-1)[...] = flat_in[slice_no * n:(slice_no + 1) * n]

    def copy_from(ptr, slice_no):  # t_3DCopyOneSliceFrom2D(ptr, x, y, x, y, i, pdest)
        n = xsize * ysize
        flat_out[slice_no * n:(slice_no + 1) * n] = m.buf[m.ptr[ptr]].reshape(-1)

    if se3d == "SESquare3D":  # :800-838
        copy_to(0, "ptmp1")
        copy_to(1, "ptmp2")
        m.op2d(se2d, dil, "imtmp1", "imtmp")
        m.swap("imtmp", "imtmp1")
        m.op2d(se2d, dil, "imtmp2", "imtmp")
        m.swap("imtmp", "imtmp2")
        m.arith(dil, "imtmp1", "imtmp2", "imtmpL")
        copy_from("ptmpL", 0)
        for i in range(2, zsize):
            copy_to(i, "ptmp3")
            m.op2d(se2d, dil, "imtmp3", "imtmp")
            m.swap("imtmp", "imtmp3")
            m.arith(dil, "imtmp1", "imtmp2", "imtmpL")
            m.arith(dil, "imtmpL", "imtmp3", "imtmp")
            m.swap("imtmp", "imtmpL")
            copy_from("ptmpL", i - 1)
            m.rotate_ptrs()
            m.assign("imtmp", "imtmp1")
            m.assign("imtmp1", "imtmp2")
            m.assign("imtmp2", "imtmp3")
            m.assign("imtmp3", "imtmp")
        m.arith(dil, "imtmp1", "imtmp2", "imtmpL")
        copy_from("ptmpL", zsize - 1)
    else:  # SECross3D, :882-914
        copy_to(0, "ptmp1")
        copy_to(1, "ptmp2")
        m.op2d(se2d, dil, "imtmp1", "imtmpL")
        m.arith(dil, "imtmpL", "imtmp2", "imtmp")
        m.swap("imtmp", "imtmpL")
        copy_from("ptmpL", 0)
        for i in range(2, zsize):
            copy_to(i, "ptmp3")
            m.op2d(se2d, dil, "imtmp2", "imtmpL")
            m.arith(dil, "imtmpL", "imtmp1", "imtmpL")
            m.arith(dil, "imtmpL", "imtmp3", "imtmp")
            m.swap("imtmp", "imtmpL")
            copy_from("ptmpL", i - 1)
            m.rotate_ptrs()
            m.assign("imtmp", "imtmp1")
            m.assign("imtmp1", "imtmp2")
            m.assign("imtmp2", "imtmp3")
            m.assign("imtmp3", "imtmp")
        m.op2d(se2d, dil, "imtmp2", "imtmpL")
        m.arith(dil, "imtmpL", "imtmp1", "imtmpL")
        copy_from("ptmpL", zsize - 1)
    return flat_out.reshape(vol.shape)


def _kat_cases():
    return kats.cases(lambda c: c.get("se") in ("SESquare3D", "SECross3D") and not c["runtime_se"])


@pytest.mark.parametrize("case", _kat_cases(), ids=kats.case_id)
def test_literal_restatement_reproduces_reference_kats(case):
    got = reference_static3d(kats.image(case["inputs"][0]), case["se"], case["op"] == "t_Dilate")
    np.testing.assert_array_equal(got, kats.image(case["expected"]))


def _zconst(n, seed):
    sl = np.random.default_rng(seed).integers(0, 256, size=(n, n)).astype(np.uint8)
    return np.repeat(sl[None], n, axis=0)


@pytest.mark.parametrize("se", ["SESquare3D", "SECross3D"])
@pytest.mark.parametrize("dil", [True, False])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_z_constant_3cube_agrees_with_definition(se, dil, seed):
    vol = _zconst(3, seed)
    np.testing.assert_array_equal(reference_static3d(vol, se, dil), oracle.static(vol, se, dil))


@pytest.mark.parametrize("se", ["SESquare3D", "SECross3D"])
@pytest.mark.parametrize("dil", [True, False])
def test_z_constant_4cube_diverges(se, dil):
    vol = _zconst(4, 4)
    assert not np.array_equal(reference_static3d(vol, se, dil), oracle.static(vol, se, dil))


@pytest.mark.parametrize("se", ["SESquare3D", "SECross3D"])
@pytest.mark.parametrize("dil", [True, False])
def test_z_varying_cube_diverges_from_definition(se, dil):
    """On a 6^3 cube with distinct slices the literal code disagrees with the definition the
    library computes.  SESquare3D: slice 0 is right and slice 1 is op(D2 s0, D2 s1), slice 2
    missing.  SECross3D: slice 0 is the 2D cross of slice 0 alone, slice 1 missing."""
    n = 6
    rng = np.random.default_rng(11)
    vol = rng.integers(0, 256, size=(n, n, n)).astype(np.uint8)
    lit = reference_static3d(vol, se, dil)
    want = oracle.static(vol, se, dil)
    f = np.maximum if dil else np.minimum
    se2d = "SESquare2D" if se == "SESquare3D" else "SECross2D"
    assert not np.array_equal(lit, want)
    d0 = oracle.static(vol[0], se2d, dil)
    if se == "SESquare3D":
        np.testing.assert_array_equal(lit[0], want[0])
        np.testing.assert_array_equal(lit[1], f(d0, oracle.static(vol[1], se2d, dil)))
        assert not np.array_equal(lit[1], want[1])
    else:
        np.testing.assert_array_equal(lit[0], d0)
        np.testing.assert_array_equal(want[0], f(d0, vol[1]))
        assert not np.array_equal(lit[0], want[0])


def test_non_cubic_volume_reads_memory_as_w_h_d():
    """xsize = bound[0] (the slowest axis) is used as the row length of a slice
    (:777-779, :86-101): on a D x H x W volume with D != W the slices are D*H-element runs of
    memory, i.e. the volume is read as if its shape were (W, H, D) -- even a z-constant
    volume of that shape does not give the 3D dilation."""
    rng = np.random.default_rng(5)
    D, H, W = 3, 4, 5
    sl = rng.integers(0, 256, size=(H, W)).astype(np.uint8)
    vol = np.repeat(sl[None], D, axis=0)
    lit = reference_static3d(vol, "SESquare3D", True)
    assert not np.array_equal(lit, oracle.static(vol, "SESquare3D", True))
