"""Host-side logic that needs no GPU: the defect-ring locator."""
import numpy as np

from qmin import defects


def _ring_field(shape, center, radius, normal, core=1.6):
    """A largest-eigenvalue field of 0.53 with a ring of reduced order (0.2)
    around `center`: sites within `core` of the circle."""
    lx, ly, lz = shape
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(float) - np.asarray(center)
    p -= np.asarray(shape) * np.round(p / np.asarray(shape))
    n = np.asarray(normal, float) / np.linalg.norm(normal)
    h = p @ n
    r = np.linalg.norm(p - np.outer(h, n), axis=1)
    dist = np.sqrt((r - radius) ** 2 + h ** 2)
    smax = np.full(len(p), 0.53)
    smax[dist < core] = 0.2
    return smax


def test_ring_fit_recovers_plane_radius_and_centre():
    shape = (64, 60, 56)
    center = (30.0, 28.0, 27.0)
    normal = (0.6, 0.0, 0.8)
    smax = _ring_field(shape, center, 14.0, normal)
    types = np.zeros(smax.size, np.int32)
    pts, bulk = defects.defect_sites(smax, types, shape)
    assert bulk == 0.53
    fit = defects.fit_ring(pts, center, shape, 30.0)
    assert fit is not None and fit["sites"] > 100
    assert abs(abs(np.dot(fit["normal"], np.array(normal))) - 1.0) < 1e-2
    assert abs(fit["radius"] - 14.0) < 0.5
    assert np.abs(fit["center_offset"]).max() < 0.3
    assert fit["rms_distance"] < 1.5


def test_ring_across_the_periodic_boundary():
    shape = (40, 40, 40)
    center = (1.0, 38.5, 20.0)
    smax = _ring_field(shape, center, 9.0, (0.0, 0.0, 1.0))
    pts, _ = defects.defect_sites(smax, np.zeros(smax.size, np.int32), shape)
    fit = defects.fit_ring(pts, center, shape, 20.0)
    assert abs(fit["normal"][2] - 1.0) < 1e-3
    assert abs(fit["radius"] - 9.0) < 0.5


def test_object_sites_are_not_defects():
    shape = (20, 20, 20)
    smax = np.full(8000, 0.53)
    types = np.zeros(8000, np.int32)
    smax[:100] = 0.0
    types[:100] = 1
    pts, _ = defects.defect_sites(smax, types, shape)
    assert len(pts) == 0
    assert defects.fit_ring(pts, (10, 10, 10), shape, 10.0) is None
