"""Pins of the oracle's open-space domain (simulation_type 0,
FDTDSolver.cu:479-550): frame and positions as the reference computes them
(integer halves), causality of the stencil, the constant power-of-two depth
0.5 making 2.5D bit-identical to 2D, and exact mirror symmetry of the
x-directed point source about its row before the wave meets the boundary."""
import numpy as np
import pytest

import oracle as orc
from vocalsim import tubes


@pytest.mark.parametrize("pml,H,W,ey,ex", [(1, 57, 355, 27, 176), (0, 45, 343, 21, 170)])
def test_frame_and_positions(pml, H, W, ey, ex):
    t = orc.OracleTube(tubes.open_space(num_steps=5, pml=pml), 64)
    d = t.dims
    assert (d["frameH"], d["frameW"], d["domainH"], d["domainW"]) == (H, W, 43, 341)
    assert (d["excitationY"], d["excitationX"]) == (ey, ex)
    assert (d["listenerY"], d["listenerX"]) == (ey + 5, ex + 5)
    ct = t.field("cell_type")
    assert ct[ey, ex] == 2                                   # the one excitation cell
    assert (ct == 2).sum() == 1
    assert (t.field("boundary_seg") == 0).all() and (t.field("z_inv") == 0).all()
    inner = ct[1 + 6 * pml:H - 1 - 6 * pml, 1 + 6 * pml:W - 1 - 6 * pml]
    assert ((inner == 1) | (inner == 2)).all()
    for name in ("depth_x", "depth_y", "depth_p"):
        assert (t.field(name) == 0.5).all()


def test_causality_of_the_stencil():
    p, _ = orc.OracleTube(tubes.open_space(num_steps=30), 64).run()
    # the listener is 5 + 5 cells from the source: nothing arrives before step 10
    assert (p[:10] == 0).all()
    assert (p[10:20] != 0).all()


def test_constant_half_depth_equals_2d_bitwise():
    sp = tubes.open_space(num_steps=300)
    a, _ = orc.OracleTube(sp, 64).run()
    b, _ = orc.OracleTube(sp.replace(solver_type=0), 64).run()
    assert np.array_equal(a, b)
    a32, _ = orc.OracleTube(sp, 32).run()
    b32, _ = orc.OracleTube(sp.replace(solver_type=0), 32).run()
    assert np.array_equal(a32, b32)


def test_mirror_symmetry_about_the_source_row():
    t = orc.OracleTube(tubes.open_space(num_steps=18), 64)
    t.run()
    P, Vx, Vy = t.state()
    e = t.dims["excitationY"]
    k = np.arange(1, 19)
    assert np.abs(P).max() > 0
    assert np.array_equal(P[e + k], P[e - k])
    assert np.array_equal(Vx[e + k], Vx[e - k])
