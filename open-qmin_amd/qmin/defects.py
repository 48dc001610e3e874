"""Defect-ring locator on the largest-eigenvalue field.

The reference writes the largest eigenvalue of Q per site with its states
(qTensorLatticeModel::computeDefectMeasures, qTensorLatticeModel.cpp:95-133;
multirankSimulation.cpp:644-661) and leaves locating defects to the user.  A
homeotropic colloid in a uniform nematic carries a Saturn ring: a loop of
reduced order around it.  ring_locus() finds the liquid-crystal sites whose
largest eigenvalue is below a fraction of the bulk value and fits a circle to
those near a given centre: the plane's normal (the direction of least spread),
the ring's radius and the offset of its centre.  The field comes from the
device (Lattice.defects(0)); the fit runs on the few defect sites.
"""
import numpy as np


def _min_image(d, L):
    return d - L * np.round(d / L)


def defect_sites(smax, types, shape, fraction=0.8):
    """Coordinates (n, 3) of liquid-crystal sites (type <= 0) whose largest
    eigenvalue is below `fraction` of the median over bulk sites (type 0)."""
    lx, ly, lz = shape
    lc = types <= 0
    bulk = np.median(smax[types == 0])
    idx = np.flatnonzero(lc & (smax < fraction * bulk))
    x = idx % lx
    y = (idx // lx) % ly
    z = idx // (lx * ly)
    return np.stack([x, y, z], axis=1).astype(np.float64), float(bulk)


def fit_ring(points, center, shape, rmax):
    """Circle through `points` (min-image offsets from `center` on the periodic
    lattice, kept within rmax of it): normal, radius, centre offset and the
    rms distance of the points from the fitted circle."""
    d = np.stack([_min_image(points[:, k] - center[k], shape[k]) for k in range(3)], axis=1)
    d = d[np.linalg.norm(d, axis=1) <= rmax]
    if len(d) < 6:
        return None
    mean = d.mean(axis=0)
    cov = np.cov((d - mean).T)
    w, v = np.linalg.eigh(cov)
    normal = v[:, 0]
    if normal[np.argmax(np.abs(normal))] < 0:
        normal = -normal
    # in-plane distances from the centre of the point cloud
    rel = d - mean
    inplane = rel - np.outer(rel @ normal, normal)
    r = np.linalg.norm(inplane, axis=1)
    radius = float(r.mean())
    height = rel @ normal
    rms = float(np.sqrt(np.mean((r - radius) ** 2 + height ** 2)))
    return {"sites": int(len(d)), "normal": normal.tolist(), "radius": radius,
            "center_offset": mean.tolist(), "rms_distance": rms}


def ring_locus(lattice, center, sphere_radius, fraction=0.8, reach=2.5):
    """The Saturn ring of the sphere at `center`: defect sites within
    reach * sphere_radius of the centre, fitted by fit_ring.  Returns None if
    fewer than six defect sites are found there."""
    smax = lattice.defects(0)
    pts, bulk = defect_sites(smax, lattice.types(), lattice.shape, fraction)
    out = fit_ring(pts, center, lattice.shape, reach * sphere_radius)
    if out is not None:
        out.update(bulk_largest_eigenvalue=bulk, threshold_fraction=fraction)
    return out
