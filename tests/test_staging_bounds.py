"""CPU checks of the LDS staging loads' addresses and of the Python layer's
argument checks (no GPU needed).

The LDS backprojection stages a box of padded detector pixels per view with
raw buffer loads whose row offset is a scalar operand; the buffer's range
check does not cover that operand, so the filtered buffer itself must hold
every element a load touches.  The replay below follows stage_block's
address arithmetic (bp_device.hpp) over every box tile_boxes can produce
and checks that the highest element read stays below the end of the buffer
the plan allocates: rows_pv * pitch elements per view, then
filt_spare_elems(pitch) = 2 * pitch + 64 spare elements (plan_internal.hpp).
"""
import numpy as np
import pytest

import gct_amd as g

# (z unroll, views per block, staged rows TR, pairs per row TP) of kLdsShapes
LDS_SHAPES = [(16, 2, 40, 40), (16, 2, 32, 40), (16, 2, 40, 48), (16, 3, 28, 34), (16, 4, 24, 28), (32, 3, 48, 34),
              (32, 4, 36, 28)]


def spare_elems(pitch):
    return 2 * pitch + 64


def staged_max_index(nu, nv, pitch, TR, TP):
    """Highest element index (from the view's first element) that any staging
    load of any box reads, by enumerating the boxes: clo, rlo in [-1, nu] x
    [-1, nv], widths nc - 1 <= TP, heights nr <= TR, cut at the padded
    detector (chi <= nu + 1, rhi <= nv + 1).  Loads of box row r (r < nr;
    rows past the box re-read row 0) start at padded pixel (rlo + 1 + r,
    clo + 1); lane l adds l elements (boxes of 32 pairs or more) or
    (l // 32) rows and l % 32 elements (half-row form)."""
    worst = -1
    for clo in range(-1, nu + 1):
        for nc in range(2, TP + 2):
            chi = clo + nc - 1
            if chi > nu + 1:
                break
            half = nc - 1 < 32
            # lane 63 reads the highest address of a row
            lane = 63
            voff = (lane >> 5) * pitch + (lane & 31) if half else lane
            for rlo in range(-1, nv + 1):
                nr_max = min(TR, nv + 1 - rlo + 1)
                if nr_max < 2:
                    continue
                # the highest row a load reads is the box's last (rows past it re-read row 0)
                top = (rlo + 1 + nr_max - 1) * pitch + (clo + 1) + voff
                worst = max(worst, top)
    return worst


@pytest.mark.parametrize("nu,nv", [(29, 4), (45, 6), (61, 9), (77, 3), (100, 20), (253, 12), (512, 40)])
@pytest.mark.parametrize("shape", LDS_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_staging_loads_stay_inside_the_filtered_buffer(nu, nv, shape):
    _, _, TR, TP = shape
    pitch = (nu + 3 + 15) & ~15
    rows_pv = nv + 3
    top = staged_max_index(nu, nv, pitch, TR, TP)
    # the last view's loads: inside the view plus the spare elements after it
    assert top < rows_pv * pitch + spare_elems(pitch), (nu, nv, pitch, top)


def test_single_spare_row_is_exceeded_somewhere():
    """The bound is tight enough to matter: for some detector widths the
    half-row loads of the last box row run past one spare row."""
    over = []
    for nu in range(29, 140):
        pitch = (nu + 3 + 15) & ~15
        for nv in (4, 9):
            top = staged_max_index(nu, nv, pitch, 40, 40)
            if top >= (nv + 3) * pitch + pitch:
                over.append((nu, nv))
    assert over, "no width exceeds a single spare row"


# ---- argument checks of the Python layer (before any library call) ----------------------

def test_device_arguments_are_checked_before_any_call():
    torch = pytest.importorskip("torch")
    with pytest.raises(TypeError):
        g._dev(np.zeros(6, np.float32), (2, 3), "vol")
    with pytest.raises(TypeError):
        g._dev(12345, (2, 3), "vol")
    with pytest.raises(ValueError, match="GPU"):
        g._dev(torch.zeros(2, 3), (2, 3), "vol")


def test_host_arguments_are_checked_before_any_call():
    torch = pytest.importorskip("torch")
    assert g._host(np.zeros((2, 3), np.float32), (2, 3), "vol") != 0
    assert g._host(torch.zeros(2, 3), (2, 3), "vol") != 0
    with pytest.raises(ValueError, match="float32"):
        g._host(np.zeros((2, 3), np.float64), (2, 3), "vol")
    with pytest.raises(ValueError, match="float32"):
        g._host(np.zeros((3, 2), np.float32).T, (2, 3), "vol")
    with pytest.raises(ValueError, match="needs 6"):
        g._host(np.zeros(5, np.float32), (2, 3), "vol")
    with pytest.raises(TypeError):
        g._host(12345, (2, 3), "vol")


def test_batch_argument_lengths_are_checked():
    class FakePlan:
        def __init__(self):
            self.scan = type("S", (), {"proj_shape": (2, 3, 4)})()
            self.img = type("I", (), {"shape": (4, 4)})()

    plans = [FakePlan(), FakePlan()]
    with pytest.raises(ValueError, match="2 plans"):
        g._batch_args(plans, [object()], [object(), object()])
    with pytest.raises(ValueError, match="2 plans"):
        g._batch_args(plans, [object(), object()], [object()])
    with pytest.raises(TypeError):
        g._batch_args(plans, [np.zeros((2, 3, 4), np.float32)] * 2, [np.zeros((4, 4), np.float32)] * 2)
