"""Benchmark-job helpers: the pre-cracked mode-I plate and its crack-tip tracker."""
import numpy as np


def crack_tip(position, damage, dmg_min=0.35, dmg_max=0.9, xmin=0.0, cluster_radius=None):
    """Leading crack tip as (x, |y|), or (nan, nan) without a crack: crack_tip of
    jobs/jobs.jl (the tip is the point of largest x among the partially damaged points
    ahead of xmin, averaged over the cluster of such points within cluster_radius of it;
    |y| is the largest in that cluster).  position is 3 x n (Julia layout) or n x 3."""
    pos = np.asarray(position, np.float64)
    if pos.shape[0] != 3:
        pos = pos.T
    dmg = np.asarray(damage, np.float64)
    ids = np.nonzero((pos[0] >= xmin) & (dmg >= dmg_min) & (dmg <= dmg_max))[0]
    if ids.size == 0:
        return (float("nan"), float("nan"))
    lead = ids[np.argmax(pos[0, ids])]
    d = np.sqrt(np.sum((pos[:, ids] - pos[:, [lead]]) ** 2, axis=0))
    cl = ids[d < cluster_radius]
    return (float(pos[0, cl].mean()), float(np.max(np.abs(pos[1, cl]))))


def mode_i_plate(npxy=500, nz=4, dx=2e-4, a_frac=0.25, m=3.015, E=70e9, rho=2700.0, Gc=100.0,
                 v=2.0, filter_bonds=True):
    """Pre-cracked plate pulled apart in mode I (the mode_i job of jobs/mode_i.jl, with the
    plate thickness given in points): an npxy x npxy x nz box of spacing dx, a precrack of
    length a_frac * l from the left edge at mid-height between the point sets within 2 delta
    above and below it, and velocities +-v in y on the top and bottom point rows.
    filter_bonds=True removes the crossing bonds (precrack!(...; update_dmg=false))."""
    from .body import Body, material, point_set, precrack, velocity_bc
    from .materials import BBMaterial
    from .point_generators import uniform_box
    l = npxy * dx
    delta = m * dx
    pos, vol = uniform_box(l, l, nz * dx, dx)
    body = Body(BBMaterial(), pos, vol)
    material(body, horizon=delta, E=E, rho=rho, Gc=Gc)
    x, y = pos[0], pos[1]
    a = a_frac * l
    point_set(body, "set_a", (x <= -l / 2 + a) & (y >= 0) & (y <= 2 * delta))
    point_set(body, "set_b", (x <= -l / 2 + a) & (y >= -2 * delta) & (y < 0))
    precrack(body, "set_a", "set_b", update_dmg=not filter_bonds)
    point_set(body, "set_top", y > l / 2 - dx)
    point_set(body, "set_bottom", y < -l / 2 + dx)
    velocity_bc(lambda t: -v, body, "set_bottom", "y")
    velocity_bc(lambda t: v, body, "set_top", "y")
    return body, {"l": l, "a": a, "delta": delta, "xmin": -l / 2 + a}
