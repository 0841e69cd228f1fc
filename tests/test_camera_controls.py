"""Camera controls (src/Camera.cpp:36-166) against rotation geometry.

glm itself is absent here, so these are checked against what the
operations mean: moves of move_speed along the camera's axes, a yaw of
yaw * rotation_speed radians about world +Y, a pitch about the camera's
side axis that is refused when it would tip `up` below the horizon, an
orthonormal right-handed basis throughout, and setBuffer's packing.
"""
import math

import numpy as np
import pytest

import yune_amd as Y


def test_identity_view_packs_like_yn_camera():
    c = Y.CameraControl(45.0, 0.25, 0.1)
    assert bytes(c.cam()) == bytes(Y.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0, 0, 0), 45.0))


@pytest.mark.parametrize("direction,axis,sign", [((0, 0, 1), 2, 1), ((0, 0, -1), 2, -1), ((1, 0, 0), 0, 1),
                                                 ((-1, 0, 0), 0, -1), ((0, 1, 0), 1, 1), ((0, -1, 0), 1, -1)])
def test_moves_along_camera_axes(direction, axis, sign):
    c = Y.CameraControl(60.0, 0.25, 0.5)
    c.set_view((0, 0, 1), (0, 1, 0), (1, 0, 0), (1, 2, 3))       # side +z, look +x
    side, up, look, eye0 = c.basis()
    c.orient(direction)
    eye = c.basis()[3]
    along = [side, up, look][axis]
    assert np.allclose(eye[:3] - eye0[:3], sign * 0.5 * along[:3], atol=1e-6)
    cam = c.cam()
    assert np.allclose([cam.r1[3], cam.r2[3], cam.r3[3]], eye[:3])


def test_first_nonzero_component_wins_z_then_x_then_y():
    c = Y.CameraControl(60.0, 0.25, 1.0)
    c.orient((1, 1, 1))
    assert np.allclose(c.basis()[3][:3], (0, 0, -1))              # moved along look only


def test_yaw_is_rotation_about_world_y():
    c = Y.CameraControl(60.0, 0.25, 0.1)
    c.orient((0, 0, 0), 0.0, 2.0)                                 # 0.5 rad
    side, up, look, _ = c.basis()
    th = 0.5
    assert np.allclose(look[:3], (-math.sin(th), 0, -math.cos(th)), atol=1e-6)
    assert np.allclose(side[:3], (math.cos(th), 0, -math.sin(th)), atol=1e-6)
    assert np.allclose(up[:3], (0, 1, 0), atol=1e-6)


def test_pitch_rotates_about_side_and_stops_at_vertical():
    c = Y.CameraControl(60.0, 0.25, 0.1)
    c.orient((0, 0, 0), 1.0, 0.0)                                 # +0.25 rad about side (+x): look tilts up
    side, up, look, _ = c.basis()
    assert np.allclose(look[:3], (0, math.sin(0.25), -math.cos(0.25)), atol=1e-6)
    for _ in range(40):                                           # far past vertical
        c.orient((0, 0, 0), 1.0, 0.0)
        side, up, look, _ = c.basis()
        assert up[1] >= -1e-6
    assert look[1] > 0.9                                          # stopped near straight up


def test_basis_stays_orthonormal_and_right_handed():
    rng = np.random.default_rng(5)
    c = Y.CameraControl(60.0, 0.3, 0.2)
    for _ in range(200):
        d = [(0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 0, 0)][int(rng.integers(4))]
        c.orient(d, float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)))
        side, up, look, _ = c.basis()
        B = np.stack([side[:3], up[:3], look[:3]])
        assert np.allclose(B @ B.T, np.eye(3), atol=1e-4)
        assert np.allclose(np.cross(side[:3], up[:3]), -look[:3], atol=1e-4)
    cam = c.cam()
    assert tuple(cam.r4) == (0, 0, 0, 1)


def test_rejects_bad_arguments():
    with pytest.raises(Y.YuneError):
        Y.CameraControl(0.0)
    with pytest.raises(Y.YuneError):
        Y.CameraControl(200.0)
