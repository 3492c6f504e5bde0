"""Coil-port excitations: wire loops as current segments and their incident
fields on the voxel grid (CCL_Excitation.m, eval_DGF.m) through the device
kernel `marie_vie_dgf_fields`.

Segments follow the reference's CCL ascii format (CCL_Excitation.m:41-56):
one row per straight segment, [P(3) N(3)] its positive and negative end
points; the segment carries current Jc from N to P and acts as the current
element J = (P - N) Jc at its midpoint (P + N)/2.
"""
import numpy as np

from . import lib, vie


def rect_loop(center, tangent, axis, width, height, seg_per_side=8):
    """Segments (n x 6, CCL rows) of a rectangular loop: `width` along the unit
    vector `tangent`, `height` along `axis`, centred at `center`; current
    circulates counter-clockwise about tangent x axis."""
    c = np.asarray(center, float)
    t = np.asarray(tangent, float)
    a = np.asarray(axis, float)
    corners = [c - t * width / 2 - a * height / 2, c + t * width / 2 - a * height / 2,
               c + t * width / 2 + a * height / 2, c - t * width / 2 + a * height / 2]
    rows = []
    for k in range(4):
        p0, p1 = corners[k], corners[(k + 1) % 4]
        for s in range(seg_per_side):
            n_end = p0 + (p1 - p0) * s / seg_per_side
            p_end = p0 + (p1 - p0) * (s + 1) / seg_per_side
            rows.append(np.concatenate([p_end, n_end]))
    return np.array(rows)


def cylinder_loop_array(n_ports=16, radius=0.15, width=0.1, height=0.05, z_center=0.0, seg_per_side=8, rings=1,
                        length=None):
    """Rectangular loops on a cylinder of `radius` about the z axis: `rings`
    rings of n_ports / rings loops each, at angles 2 pi p / (n_ports / rings)
    and, for several rings, z centres spread evenly over `length` (default
    rings * width) about z_center; `width` along z, `height` along the
    circumference (the chord of the cylinder). Returns a list of CCL arrays,
    ring by ring."""
    if n_ports % rings:
        raise ValueError('n_ports must be a multiple of rings')
    per = n_ports // rings
    length = rings * width if length is None else length
    ports = []
    for q in range(rings):
        zc = z_center + (q - (rings - 1) / 2) * (length / rings)
        for p in range(per):
            phi = 2 * np.pi * p / per
            center = [radius * np.cos(phi), radius * np.sin(phi), zc]
            tangent = [-np.sin(phi), np.cos(phi), 0.0]
            ports.append(rect_loop(center, [0.0, 0.0, 1.0], tangent, width, height, seg_per_side))
    return ports


def eval_DGF(J, Icoord, Ocoord, ko, device=0):
    """E, H (No x 3, or No x 3 x Ne for J of shape Ni x 3 x Ne) of current
    elements J at Icoord, observed at Ocoord (eval_DGF.m:1)."""
    J = np.asarray(J, complex)
    if J.ndim == 2:
        J = J[..., None]
    Ne = J.shape[2]
    em = vie.EMconstants(vie.C0 * ko / (2 * np.pi))
    Ic = np.asarray(Icoord, float).reshape(-1, 3)
    sc = np.tile(Ic, (Ne, 1))
    sj = np.concatenate([J[:, :, e] for e in range(Ne)])
    ps = np.arange(Ne + 1) * Ic.shape[0]
    E, H = lib.dgf_fields(sc, sj, ps, Ocoord, ko, em['omega_mu'], device)
    E, H = np.moveaxis(E, 0, -1), np.moveaxis(H, 0, -1)
    return (E[..., 0], H[..., 0]) if Ne == 1 else (E, H)


def CCL_Excitation(r, ko, Jc, ccl, device=0):
    """[E, H] (L x M x N x 3) of a segmented coil (CCL_Excitation.m:1); ccl is
    an n x 6 array of [P N] rows (or a list of them, one per port, returning
    lists) and Jc the current per segment (scalar or per row)."""
    ports = ccl if isinstance(ccl, (list, tuple)) else [ccl]
    r = np.asarray(r, float)
    L, M, N, _ = r.shape
    pts = r.reshape(L * M * N, 3, order='F')
    centers, elems, starts = [], [], [0]
    for rows in ports:
        rows = np.asarray(rows, float)
        P, Nn = rows[:, 0:3], rows[:, 3:6]
        centers.append((P + Nn) / 2)
        elems.append((P - Nn) * np.broadcast_to(np.asarray(Jc, complex), (rows.shape[0],))[:, None])
        starts.append(starts[-1] + rows.shape[0])
    em = vie.EMconstants(vie.C0 * ko / (2 * np.pi))
    E, H = lib.dgf_fields(np.concatenate(centers), np.concatenate(elems), starts, pts, ko, em['omega_mu'], device)
    Es = [E[p].reshape(L, M, N, 3, order='F') for p in range(len(ports))]
    Hs = [H[p].reshape(L, M, N, 3, order='F') for p in range(len(ports))]
    if isinstance(ccl, (list, tuple)):
        return Es, Hs
    return Es[0], Hs[0]
