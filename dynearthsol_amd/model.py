"""Host-side model set-up: parameters, box meshes, boundary artefacts, initial
conditions.

This mirrors what the reference does on the host before its step loop
(dynearthsol.cxx:31-101): parameter defaults (input.cxx:33-459), a box mesh
with the reference's renumbering (mesh.cxx:1431-1530), boundary flags, facets
and normals (mesh.cxx:1546-1704, bc.cxx:87-176), and the initial temperature,
stress and weak zone (ic.cxx:96-231).  Meshes come from a structured box
tetrahedralization instead of Triangle/TetGen (those meshers are not on the
device path); any mesh with the same arrays can be used instead.

All arrays are numpy, in the reference's AoS layout.
"""
import configparser
import math

import numpy as np

from . import abi

YEAR2SEC = 365.2422 * 86400

# Boundary flag bits (constants.hpp:38-51)
BOUNDX0, BOUNDX1, BOUNDY0, BOUNDY1, BOUNDZ0, BOUNDZ1 = 1, 2, 4, 8, 16, 32
IBOUNDX0, IBOUNDX1, IBOUNDY0, IBOUNDY1, IBOUNDZ0, IBOUNDZ1 = range(6)

# local node numbers of each facet (constants.hpp:61-71)
NODE_OF_FACET = {
    3: np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]], dtype=np.int32),
    2: np.array([[1, 2], [2, 0], [0, 1]], dtype=np.int32),
}

RHEOLOGY = {
    "elastic": abi.RH_ELASTIC,
    "viscous": abi.RH_VISCOUS,
    "maxwell": abi.RH_MAXWELL,
    "elasto-plastic": abi.RH_EP,
    "elasto-visco-plastic": abi.RH_EVP,
}

# Defaults of the keys the explicit step reads (input.cxx:33-459).
DEFAULTS = {
    "mesh.xlength": None, "mesh.ylength": None, "mesh.zlength": None,
    "mesh.resolution": None,
    "mesh.smallest_size": 0.01, "mesh.min_quality": 0.4,
    "mesh.quality_check_step_interval": 100,
    "markers.markers_per_element": 4,
    "control.gravity": 10.0, "control.characteristic_speed": 0.0,
    "control.is_quasi_static": True, "control.dt_fraction": 1.0,
    "control.fixed_dt": 0.0, "control.inertial_scaling": 1e5,
    "control.damping_option": 1, "control.damping_factor": 0.8,
    "control.ref_pressure_option": 0, "control.surface_process_option": 0,
    "control.has_thermal_diffusion": True,
    "bc.surface_temperature": 273.0, "bc.mantle_temperature": 1600.0,
    "bc.has_fixed_bottom_temperature": True,
    "bc.has_winkler_foundation": True, "bc.winkler_delta_rho": 0.0,
    "bc.has_elastic_foundation": False, "bc.elastic_foundation_constant": 1e11,
    "bc.has_water_loading": True,
    "bc.vbc_x0": 1, "bc.vbc_x1": 1, "bc.vbc_y0": 0, "bc.vbc_y1": 0,
    "bc.vbc_z0": 0, "bc.vbc_z1": 0,
    "bc.vbc_n0": 1, "bc.vbc_n1": 1, "bc.vbc_n2": 1, "bc.vbc_n3": 1,
    "bc.vbc_val_x0": -1e-9, "bc.vbc_val_x1": 1e-9, "bc.vbc_val_y0": 0.0,
    "bc.vbc_val_y1": 0.0, "bc.vbc_val_z0": 0.0, "bc.vbc_val_z1": 0.0,
    "bc.vbc_val_n0": 0.0, "bc.vbc_val_n1": 0.0, "bc.vbc_val_n2": 0.0,
    "bc.vbc_val_n3": 0.0,
    "ic.weakzone_option": 1, "ic.weakzone_plstrain": 0.1,
    "ic.weakzone_azimuth": 0.0, "ic.weakzone_inclination": 90.0,
    "ic.weakzone_halfwidth": 1.5, "ic.weakzone_y_min": 0.0,
    "ic.weakzone_y_max": 1.0, "ic.weakzone_depth_min": 0.0,
    "ic.weakzone_depth_max": 1.0, "ic.weakzone_xcenter": 0.5,
    "ic.weakzone_ycenter": 0.5, "ic.weakzone_zcenter": 0.0,
    "ic.weakzone_xsemi_axis": 1e3, "ic.weakzone_ysemi_axis": 1e3,
    "ic.weakzone_zsemi_axis": 1e3,
    "ic.temperature_option": 0, "ic.oceanic_plate_age_in_yr": 60e6,
    "mat.rheology_type": None, "mat.is_plane_strain": False,
    "mat.num_materials": 1,
    "mat.max_viscosity": 1e24, "mat.min_viscosity": 1e18,
    "mat.max_tension": 1e9, "mat.max_thermal_diffusivity": 5e-6,
    "mat.rho0": [3210.0], "mat.alpha": [3e-5],
    "mat.bulk_modulus": [128.2e9], "mat.shear_modulus": [80.5e9],
    "mat.visc_exponent": [3.05], "mat.visc_coefficient": [1.25e-1],
    "mat.visc_activation_energy": [3.76e5],
    "mat.heat_capacity": [1e3], "mat.therm_cond": [3.0],
    "mat.radiogenic_heating": [0.0],
    "mat.pls0": [0.0], "mat.pls1": [0.1],
    "mat.cohesion0": [4e7], "mat.cohesion1": [4e6],
    "mat.friction_angle0": [30.0], "mat.friction_angle1": [5.0],
    "mat.dilation_angle0": [0.0], "mat.dilation_angle1": [0.0],
}

MAT_VECTORS = (
    "rho0", "alpha", "bulk_modulus", "shear_modulus", "visc_exponent",
    "visc_coefficient", "visc_activation_energy", "heat_capacity",
    "therm_cond", "radiogenic_heating", "pls0", "pls1", "cohesion0",
    "cohesion1", "friction_angle0", "friction_angle1", "dilation_angle0",
    "dilation_angle1",
)


def _parse_value(text):
    t = text.strip()
    if t.startswith("["):
        return [float(x) for x in t.strip("[]").split(",") if x.strip()]
    low = t.lower()
    if low in ("yes", "true", "on"):
        return True
    if low in ("no", "false", "off"):
        return False
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(t)
    except ValueError:
        return t


def load_cfg(text):
    """Parse a reference-style INI config (sections sim/mesh/control/bc/ic/mat)
    into a {"section.key": value} dict."""
    cp = configparser.ConfigParser(inline_comment_prefixes=("#",))
    cp.optionxform = str
    cp.read_string(text)
    out = {}
    for sec in cp.sections():
        for k, v in cp.items(sec):
            out[f"{sec}.{k}"] = _parse_value(v)
    return out


def resolve_cfg(overrides):
    """Defaults + overrides with the silent overrides of validate_parameters
    (input.cxx:650-670)."""
    cfg = dict(DEFAULTS)
    for k, v in overrides.items():
        cfg[k] = v
    if cfg["bc.has_winkler_foundation"] and cfg["control.gravity"] == 0:
        cfg["bc.has_winkler_foundation"] = False
    if cfg["bc.has_winkler_foundation"] and cfg["bc.vbc_z0"] != 0:
        cfg["bc.vbc_z0"] = 0
    if cfg["bc.has_water_loading"] and cfg["bc.vbc_z1"] != 0:
        cfg["bc.vbc_z1"] = 0
    if int(cfg["mat.num_materials"]) == 1 and cfg["control.ref_pressure_option"] != 0:
        cfg["control.ref_pressure_option"] = 0       # input.cxx:774-777
    return cfg


def find_max_vbc(cfg):
    """bc.cxx:59-84"""
    m = 0.0
    for s in ("x0", "x1", "y0", "y1", "z0", "z1", "n0", "n1", "n2", "n3"):
        if cfg[f"bc.vbc_{s}"] % 2 == 1:
            m = max(m, abs(cfg[f"bc.vbc_val_{s}"]))
    return m


def make_params(cfg, ndims):
    """Fill the C-ABI parameter struct from a resolved config dict."""
    p = abi.DesParams()
    p.ndims = ndims
    rt = cfg["mat.rheology_type"]
    p.rheol_type = RHEOLOGY[rt] if isinstance(rt, str) else int(rt)
    nmat = int(cfg["mat.num_materials"])
    if not 1 <= nmat <= abi.DES_MAX_MAT:
        raise ValueError("mat.num_materials out of range")
    p.nmat = nmat
    p.is_plane_strain = int(bool(cfg["mat.is_plane_strain"]) and ndims == 2)
    p.visc_min = cfg["mat.min_viscosity"]
    p.visc_max = cfg["mat.max_viscosity"]
    p.tension_max = cfg["mat.max_tension"]
    p.therm_diff_max = cfg["mat.max_thermal_diffusivity"]
    for name in MAT_VECTORS:
        vals = cfg[f"mat.{name}"]
        if not isinstance(vals, (list, tuple)):
            vals = [vals]
        if len(vals) == 1:          # Vector1 broadcast (matprops.cxx:233-251)
            vals = list(vals) * nmat
        if len(vals) != nmat:
            raise ValueError(f"mat.{name}: expected {nmat} values")
        arr = getattr(p, name)
        for i, v in enumerate(vals):
            arr[i] = float(v)
    p.gravity = cfg["control.gravity"]
    p.characteristic_speed = cfg["control.characteristic_speed"]
    p.inertial_scaling = cfg["control.inertial_scaling"]
    p.dt_fraction = cfg["control.dt_fraction"]
    p.fixed_dt = cfg["control.fixed_dt"]
    p.damping_factor = cfg["control.damping_factor"]
    p.damping_option = int(cfg["control.damping_option"])
    p.ref_pressure_option = int(cfg["control.ref_pressure_option"])
    p.is_quasi_static = int(bool(cfg["control.is_quasi_static"]))
    p.has_thermal_diffusion = int(bool(cfg["control.has_thermal_diffusion"]))
    p.surface_process_option = int(cfg["control.surface_process_option"])
    p.surface_temperature = cfg["bc.surface_temperature"]
    p.mantle_temperature = cfg["bc.mantle_temperature"]
    p.has_fixed_bottom_temperature = int(bool(cfg["bc.has_fixed_bottom_temperature"]))
    p.has_winkler_foundation = int(bool(cfg["bc.has_winkler_foundation"]))
    p.winkler_delta_rho = cfg["bc.winkler_delta_rho"]
    p.has_elastic_foundation = int(bool(cfg["bc.has_elastic_foundation"]))
    p.elastic_foundation_constant = cfg["bc.elastic_foundation_constant"]
    p.has_water_loading = int(bool(cfg["bc.has_water_loading"]))
    for i, s in enumerate(("x0", "x1", "y0", "y1", "z0", "z1", "n0", "n1", "n2", "n3")):
        p.vbc_types[i] = int(cfg[f"bc.vbc_{s}"])
        p.vbc_values[i] = float(cfg[f"bc.vbc_val_{s}"])
    p.zlength = cfg["mesh.zlength"]
    p.resolution = cfg["mesh.resolution"]
    p.smallest_size = cfg["mesh.smallest_size"]
    p.min_quality = cfg["mesh.min_quality"]
    # init_var (dynearthsol.cxx:36-39)
    cs = cfg["control.characteristic_speed"]
    p.max_vbc_val = find_max_vbc(cfg) if cs == 0 else cs
    return p


# ----------------------------------------------------------------- meshing

def _renumber(coord, conn, lengths):
    """renumbering_mesh (mesh.cxx:1431-1530): sort nodes by a weighted
    coordinate along the longest, middle and shortest axes, elements by the
    sum of their nodes' weights."""
    nd = coord.shape[1]
    order = np.argsort(np.asarray(lengths), kind="stable")
    dmin, dmid, dmax = order[0], order[nd // 2] if nd == 3 else None, order[-1]
    f = 1e-3
    wn = coord[:, dmax].copy()
    if nd == 3:
        wn += f * coord[:, dmid]
    wn += f * f * coord[:, dmin]
    we = wn[conn].sum(axis=1)
    nd_idx = np.argsort(wn, kind="stable")
    el_idx = np.argsort(we, kind="stable")
    nd_inv = np.empty_like(nd_idx)
    nd_inv[nd_idx] = np.arange(len(nd_idx))
    coord2 = np.ascontiguousarray(coord[nd_idx])
    conn2 = np.ascontiguousarray(nd_inv[conn[el_idx]].astype(np.int32))
    return coord2, conn2


def _tet_volume(x, conn):
    d0, d1, d2, d3 = (x[conn[:, i]] for i in range(4))
    x01, x12, x23 = d0[:, 0] - d1[:, 0], d1[:, 0] - d2[:, 0], d2[:, 0] - d3[:, 0]
    y01, y12, y23 = d0[:, 1] - d1[:, 1], d1[:, 1] - d2[:, 1], d2[:, 1] - d3[:, 1]
    z01, z12, z23 = d0[:, 2] - d1[:, 2], d1[:, 2] - d2[:, 2], d2[:, 2] - d3[:, 2]
    return (x01 * (y23 * z12 - y12 * z23) + x12 * (y01 * z23 - y23 * z01)
            + x23 * (y12 * z01 - y01 * z12)) / 6


def box_mesh(ndims, lengths, cells, jitter=0.0, seed=0):
    """Structured box mesh: x in [0,Lx], (y in [0,Ly],) z in [-Lz,0].

    3D cells are split into the 6 Kuhn tetrahedra (conforming across cells),
    2D cells into 2 triangles with alternating diagonals.  Interior node
    coordinates are jittered by up to `jitter` of a cell along each free
    axis.  Returns (coord, connectivity) renumbered as the reference does.
    """
    rng = np.random.default_rng(seed)
    if ndims == 3:
        Lx, Ly, Lz = lengths
        nx, ny, nz = cells
        xs = np.linspace(0.0, Lx, nx + 1)
        ys = np.linspace(0.0, Ly, ny + 1)
        zs = np.linspace(-Lz, 0.0, nz + 1)
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        coord = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        h = np.array([Lx / nx, Ly / ny, Lz / nz])
        if jitter > 0:
            ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1),
                                     np.arange(nz + 1), indexing="ij")
            free = [(ix.ravel() > 0) & (ix.ravel() < nx),
                    (iy.ravel() > 0) & (iy.ravel() < ny),
                    (iz.ravel() > 0) & (iz.ravel() < nz)]
            for d in range(3):
                coord[:, d] += np.where(free[d], rng.uniform(-jitter, jitter, len(coord)) * h[d], 0.0)

        def nid(i, j, k):
            return (i * (ny + 1) + j) * (nz + 1) + k

        I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        corner = {}
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    corner[(a, b, c)] = nid(I + a, J + b, K + c)
        tets = []
        for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
            p = [0, 0, 0]
            path = [corner[tuple(p)]]
            for ax in perm:
                p[ax] = 1
                path.append(corner[tuple(p)])
            tets.append(np.stack(path, axis=1))
        conn = np.concatenate(tets, axis=0).astype(np.int64)
        vol = _tet_volume(coord, conn)
        neg = vol < 0
        conn[neg, 1], conn[neg, 2] = conn[neg, 2].copy(), conn[neg, 1].copy()
    else:
        Lx, Lz = lengths
        nx, nz = cells
        xs = np.linspace(0.0, Lx, nx + 1)
        zs = np.linspace(-Lz, 0.0, nz + 1)
        X, Z = np.meshgrid(xs, zs, indexing="ij")
        coord = np.stack([X.ravel(), Z.ravel()], axis=1)
        h = np.array([Lx / nx, Lz / nz])
        if jitter > 0:
            ix, iz = np.meshgrid(np.arange(nx + 1), np.arange(nz + 1), indexing="ij")
            free = [(ix.ravel() > 0) & (ix.ravel() < nx), (iz.ravel() > 0) & (iz.ravel() < nz)]
            for d in range(2):
                coord[:, d] += np.where(free[d], rng.uniform(-jitter, jitter, len(coord)) * h[d], 0.0)

        def nid(i, k):
            return i * (nz + 1) + k

        I, K = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
        I, K = I.ravel(), K.ravel()
        a, b, c, d = nid(I, K), nid(I + 1, K), nid(I + 1, K + 1), nid(I, K + 1)
        alt = ((I + K) % 2) == 0
        t1 = np.where(alt[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
        t2 = np.where(alt[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
        conn = np.concatenate([t1, t2], axis=0).astype(np.int64)
        # counter-clockwise, as Triangle emits (the 2D shape functions assume it)
        x0, x1, x2 = (coord[conn[:, i]] for i in range(3))
        sa = (x1[:, 0] - x0[:, 0]) * (x2[:, 1] - x0[:, 1]) - (x2[:, 0] - x0[:, 0]) * (x1[:, 1] - x0[:, 1])
        neg = sa < 0
        conn[neg, 1], conn[neg, 2] = conn[neg, 2].copy(), conn[neg, 1].copy()
    return _renumber(coord, conn, lengths)


def box_delaunay_mesh(ndims, lengths, cells, jitter=0.35, seed=0, joggle=False):
    """Unstructured box mesh: the Delaunay triangulation (scipy.spatial,
    Qhull) of a lattice of points jittered by up to `jitter` of a cell along
    every axis in which the point is not on the boundary, so boundary points
    stay on their faces.  Node degrees vary as in a mesher's output (a
    structured Kuhn mesh has the same few degrees everywhere).  Degenerate
    simplices Qhull leaves on the planar faces are dropped; elements are
    oriented as the structured mesher orients them, then renumbered as the
    reference does.  joggle: Qhull's QJ option (triangulate slightly joggled
    input instead of merging the lattice's nearly co-spherical facets), about
    20x faster at 10^5 points and more, with more sliver-like elements.
    Returns (coord, connectivity)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    axes = [np.linspace(0.0, L, n + 1) for L, n in zip(lengths, cells)]
    axes[-1] = axes[-1] - lengths[-1]                       # z in [-Lz, 0]
    grids = np.meshgrid(*axes, indexing="ij")
    coord = np.stack([g.ravel() for g in grids], axis=1)
    idx = np.meshgrid(*[np.arange(n + 1) for n in cells], indexing="ij")
    h = np.array([L / n for L, n in zip(lengths, cells)])
    for d in range(ndims):
        free = (idx[d].ravel() > 0) & (idx[d].ravel() < cells[d])
        coord[:, d] += np.where(free, rng.uniform(-jitter, jitter, len(coord)) * h[d], 0.0)
    conn = Delaunay(coord, qhull_options="QJ" if joggle else None).simplices.astype(np.int64)
    if ndims == 3:
        vol = _tet_volume(coord, conn)
        keep = np.abs(vol) > 1e-9 * float(np.prod(h))
        conn, vol = conn[keep], vol[keep]
        neg = vol < 0
    else:
        x0, x1, x2 = (coord[conn[:, i]] for i in range(3))
        sa = (x1[:, 0] - x0[:, 0]) * (x2[:, 1] - x0[:, 1]) - (x2[:, 0] - x0[:, 0]) * (x1[:, 1] - x0[:, 1])
        keep = np.abs(sa) > 1e-9 * float(np.prod(h))
        conn, sa = conn[keep], sa[keep]
        neg = sa < 0
    conn[neg, 1], conn[neg, 2] = conn[neg, 2].copy(), conn[neg, 1].copy()
    return _renumber(coord, conn, lengths)


def box_cells_for(ndims, lengths, nelem_target):
    """Cell counts of a box mesh with about nelem_target elements and
    near-cubic cells."""
    per_cell = 6 if ndims == 3 else 2
    ncell = nelem_target / per_cell
    vol = float(np.prod(lengths))
    h = (vol / ncell) ** (1.0 / ndims)
    return tuple(max(1, int(round(L / h))) for L in lengths)


def box_bcflag(coord, lengths):
    """Boundary flags of the box's nodes (create_boundary_flags, mesh.cxx:1534-1554)."""
    nd = coord.shape[1]
    flag = np.zeros(len(coord), dtype=np.uint32)
    tol = 1e-9 * max(lengths)
    flag[np.abs(coord[:, 0]) <= tol] |= BOUNDX0
    flag[np.abs(coord[:, 0] - lengths[0]) <= tol] |= BOUNDX1
    if nd == 3:
        flag[np.abs(coord[:, 1]) <= tol] |= BOUNDY0
        flag[np.abs(coord[:, 1] - lengths[1]) <= tol] |= BOUNDY1
    flag[np.abs(coord[:, nd - 1] + lengths[nd - 1]) <= tol] |= BOUNDZ0
    flag[np.abs(coord[:, nd - 1]) <= tol] |= BOUNDZ1
    return flag


def boundary_facets(conn, bcflag, ndims):
    """var.bfacets (mesh.cxx:1637-1704): (boundary, element, facet) of every
    element facet whose nodes share one boundary bit, ordered by boundary,
    then element, then facet."""
    nof = NODE_OF_FACET[ndims]
    bdry, elem, face = [], [], []
    E = conn.shape[0]
    for f in range(ndims + 1):
        common = np.full(E, 0x3FF, dtype=np.uint32)
        for j in range(ndims):
            common &= bcflag[conn[:, nof[f, j]]]
        for b in range(abi.DES_NBDRY):
            hit = np.nonzero(common & (1 << b))[0]
            bdry.append(np.full(len(hit), b, dtype=np.int32))
            elem.append(hit.astype(np.int32))
            face.append(np.full(len(hit), f, dtype=np.int32))
    bdry, elem, face = (np.concatenate(a) for a in (bdry, elem, face))
    order = np.lexsort((face, elem, bdry))
    return (np.ascontiguousarray(bdry[order]), np.ascontiguousarray(elem[order]),
            np.ascontiguousarray(face[order]))


def facet_normal(coord, conn, e, f, ndims):
    """normal_vector_of_facet (bc.cxx:12-47): area-weighted outward normal."""
    nof = NODE_OF_FACET[ndims]
    n0, n1 = coord[conn[e, nof[f, 0]]], coord[conn[e, nof[f, 1]]]
    if ndims == 3:
        n2 = coord[conn[e, nof[f, 2]]]
        v01, v02 = n1 - n0, n2 - n0
        return np.array([(v01[1] * v02[2] - v01[2] * v02[1]) / 2,
                         (v01[2] * v02[0] - v01[0] * v02[2]) / 2,
                         (v01[0] * v02[1] - v01[1] * v02[0]) / 2])
    v01 = n1 - n0
    return np.array([v01[1], -v01[0]])


def boundary_normals(coord, conn, bfacets, ndims):
    """create_boundary_normals (bc.cxx:87-176): unit normal of each boundary
    (first facet) and the edge vector of each pair of boundaries."""
    bdry, elem, face = bfacets
    bn = np.zeros((abi.DES_NBDRY, ndims))
    has = np.zeros(abi.DES_NBDRY, dtype=bool)
    for i in range(abi.DES_NBDRY):
        idx = np.nonzero(bdry == i)[0]
        if len(idx) == 0:
            continue
        has[i] = True
        for k, j in enumerate(idx):
            n = facet_normal(coord, conn, elem[j], face[j], ndims)
            n = n / math.sqrt(float(np.sum(n * n)))
            if k == 0:
                bn[i] = n
                if i < 6:
                    break
            elif float(np.sum((bn[i] - n) ** 2)) > 1e-12:
                raise ValueError(f"boundary {i} is curved")
    ev = np.zeros((abi.DES_NBDRY, abi.DES_NBDRY, ndims))
    valid = np.zeros((abi.DES_NBDRY, abi.DES_NBDRY), dtype=np.int32)
    for i in range(abi.DES_NBDRY):
        if not has[i]:
            continue
        for j in range(i + 1, abi.DES_NBDRY):
            if not has[j]:
                continue
            if ndims == 3:
                if abs(bn[i, 2]) < 1e-15 and abs(bn[j, 2]) < 1e-15:
                    s = np.array([0.0, 0.0, 1.0])
                else:
                    s = np.array([bn[i, 1] * bn[j, 2] - bn[i, 2] * bn[j, 1],
                                  bn[i, 2] * bn[j, 0] - bn[i, 0] * bn[j, 2],
                                  bn[i, 0] * bn[j, 1] - bn[i, 1] * bn[j, 0]])
            else:
                s = np.array([0.0, 1.0])
            ev[i, j] = s
            valid[i, j] = 1
    return np.ascontiguousarray(bn), np.ascontiguousarray(ev), np.ascontiguousarray(valid)


PREM_DEPTH = np.array([
    0e3, 3e3, 15e3, 24.4e3, 40e3, 60e3, 80e3, 115e3, 150e3, 185e3, 220e3, 265e3, 310e3, 355e3,
    400e3, 450e3, 500e3, 550e3, 600e3, 635e3, 670e3, 721e3, 771e3, 871e3, 971e3, 1071e3,
    1171e3, 1271e3, 1371e3, 1471e3, 1571e3, 1671e3, 1771e3, 1871e3, 1971e3, 2071e3, 2171e3,
    2271e3, 2371e3, 2471e3, 2571e3, 2671e3, 2741e3, 2771e3, 2871e3, 2891e3])
_PREM_TAIL = [11.2e8, 17.8e8, 24.5e8, 36.1e8, 47.8e8, 59.4e8, 71.1e8, 86.4e8,
              102.0e8, 117.7e8, 133.5e8, 152.2e8, 171.3e8, 190.7e8, 210.4e8, 224.3e8, 238.3e8,
              260.7e8, 282.9e8, 327.6e8, 372.8e8, 418.6e8, 464.8e8, 511.6e8, 558.9e8, 606.8e8,
              655.2e8, 704.1e8, 753.5e8, 803.6e8, 854.3e8, 905.6e8, 957.6e8, 1010.3e8, 1063.8e8,
              1118.2e8, 1173.4e8, 1229.7e8, 1269.7e8, 1287.0e8, 1345.6e8, 1357.5e8]
PREM_P = {1: np.array([0e8, 0.3e8, 3.3e8, 6.0e8] + _PREM_TAIL),
          2: np.array([0e8, 0.82e8, 4.1e8, 6.7e8] + _PREM_TAIL)}


def prem_pressure(depth, option):
    """get_prem_pressure[_modified] (matprops.cxx:11-99), vectorised."""
    tab = PREM_P[option]
    depth = np.asarray(depth, dtype=np.float64)
    out = np.zeros_like(depth)
    for k, d in enumerate(depth):
        if d <= 0:
            continue
        n = 1
        while n < 46 and not d <= PREM_DEPTH[n]:
            n += 1
        n = min(n, 45) if n == 46 else n
        out[k] = tab[n - 1] + (tab[n] - tab[n - 1]) * (d - PREM_DEPTH[n - 1]) / (PREM_DEPTH[n] - PREM_DEPTH[n - 1])
    return out


# ------------------------------------------------------- seeded random fields

def splitmix_uniform(seed, n, stream=0):
    """Counter-based uniform [0,1) per index: SplitMix64 of (seed, stream, i).
    Identical on every platform and thread count."""
    with np.errstate(over="ignore"):
        z = (np.arange(n, dtype=np.uint64)
             + np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF))
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def splitmix_normal(seed, n, stream=0):
    u1 = splitmix_uniform(seed, n, 2 * stream + 1)
    u2 = splitmix_uniform(seed, n, 2 * stream + 2)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2 * math.pi * u2)


# -------------------------------------------------------------- the model

class Model:
    """A model ready for the step loop: parameters, mesh, set-up artefacts and
    the initial fields, all host arrays in the reference's layout."""

    def __init__(self, cfg_overrides, ndims, cells=None, nelem_target=None,
                 jitter=0.15, mesh_seed=0, plstrain_noise=None, elem_visc=None,
                 elem_cohesion=None, temperature="erf", elemmarkers=None, mesh="structured"):
        cfg = resolve_cfg(cfg_overrides)
        self.cfg = cfg
        self.ndims = nd = ndims
        self.npe = nd + 1
        self.nstr = 6 if nd == 3 else 3
        self.params = make_params(cfg, nd)
        if nd == 3:
            lengths = (cfg["mesh.xlength"], cfg["mesh.ylength"], cfg["mesh.zlength"])
        else:
            lengths = (cfg["mesh.xlength"], cfg["mesh.zlength"])
        self.lengths = lengths
        if cells is None:
            cells = box_cells_for(nd, lengths, nelem_target)
        self.cells = cells
        if mesh in ("delaunay", "delaunay-joggled"):
            self.coord, self.conn = box_delaunay_mesh(nd, lengths, cells, seed=mesh_seed,
                                                      joggle=mesh == "delaunay-joggled")
        elif mesh == "structured":
            self.coord, self.conn = box_mesh(nd, lengths, cells, jitter=jitter, seed=mesh_seed)
        else:
            raise ValueError(mesh)
        self.nnode, self.nelem = len(self.coord), len(self.conn)
        self.bcflag = box_bcflag(self.coord, lengths)
        self.bfacets = boundary_facets(self.conn, self.bcflag, nd)
        self.bnormals, self.edge_vectors, self.edge_valid = boundary_normals(
            self.coord, self.conn, self.bfacets, nd)
        self.markers_per_element = int(cfg["markers.markers_per_element"])
        # marker counts per element and material (markerset.cxx:145-173); None
        # with nmat == 1 means every element holds markers_per_element of mat 0
        self.elemmarkers = None
        if elemmarkers is not None:
            em = elemmarkers(self) if callable(elemmarkers) else elemmarkers
            self.elemmarkers = np.ascontiguousarray(em, dtype=np.int32)
            assert self.elemmarkers.shape == (self.nelem, self.params.nmat)
        elif self.params.nmat != 1:
            raise ValueError("nmat > 1 needs elemmarkers")
        self._initial_conditions(temperature, plstrain_noise, elem_visc, elem_cohesion)

    # ic.cxx
    def _elem_center(self):
        return self.coord[self.conn].mean(axis=1)

    def _initial_conditions(self, temperature, plstrain_noise, elem_visc, elem_cohesion):
        cfg, nd, p = self.cfg, self.ndims, self.params
        N, E = self.nnode, self.nelem
        z = self.coord[:, nd - 1]
        Ts, Tm = cfg["bc.surface_temperature"], cfg["bc.mantle_temperature"]
        if temperature == "erf":
            # initial_temperature, option 0 (ic.cxx:207-231); rho(0) sees T = 0 here
            age = cfg["ic.oceanic_plate_age_in_yr"] * YEAR2SEC
            rho_T0 = p.rho0[0] * (1 - p.alpha[0] * (0.0 - 273))
            diff = p.therm_cond[0] / rho_T0 / p.heat_capacity[0]
            w = -z / math.sqrt(4 * diff * age)
            from scipy.special import erf
            T = Ts + (Tm - Ts) * erf(w)
        elif temperature == "linear":
            # extension: linear profile from the top to the bottom boundary
            T = Ts + (Tm - Ts) * (-z / cfg["mesh.zlength"])
        else:
            raise ValueError(temperature)
        self.temperature0 = np.ascontiguousarray(T)

        # initial_stress_state (ic.cxx:96-132): lithostatic, option 0 only
        stress = np.zeros((E, self.nstr))
        strain = np.zeros((E, self.nstr))
        stressyy = np.zeros(E)
        if p.gravity != 0:
            zc = np.zeros(E)
            for i in range(self.npe):
                zc += self.coord[self.conn[:, i], nd - 1]
            zc /= self.npe
            if p.ref_pressure_option == 0:
                ks = p.bulk_modulus[0]      # bulkm(0) with one material
                pr = p.rho0[0] * p.gravity * (-zc)
            else:
                ks = self.bulkm_elem()
                pr = prem_pressure(-zc, p.ref_pressure_option)
            for i in range(nd):
                stress[:, i] = -pr
                strain[:, i] = -pr / ks / nd
            if p.is_plane_strain:
                stressyy[:] = -pr
            if p.ref_pressure_option == 0:
                self.compensation_pressure = p.rho0[0] * p.gravity * cfg["mesh.zlength"]
            else:
                self.compensation_pressure = float(
                    prem_pressure(np.array([cfg["mesh.zlength"]]), p.ref_pressure_option)[0])
        else:
            self.compensation_pressure = 0.0
        self.stress0, self.strain0, self.stressyy0 = stress, strain, stressyy

        # initial_weak_zone (ic.cxx:135-204)
        pl = np.zeros(E)
        c = self._elem_center()
        opt = cfg["ic.weakzone_option"]
        L = self.lengths
        if opt == 1:
            x0 = [cfg["ic.weakzone_xcenter"] * L[0]]
            if nd == 3:
                x0.append(cfg["ic.weakzone_ycenter"] * L[1])
            x0.append(-cfg["ic.weakzone_zcenter"] * L[nd - 1])
            az = math.tan(cfg["ic.weakzone_azimuth"] * math.pi / 180)
            incl = 1 / math.tan(cfg["ic.weakzone_inclination"] * math.pi / 180)
            hw = cfg["ic.weakzone_halfwidth"] * cfg["mesh.resolution"]
            zmin = -cfg["ic.weakzone_depth_max"] * L[nd - 1]
            zmax = -cfg["ic.weakzone_depth_min"] * L[nd - 1]
            inside = (c[:, nd - 1] > zmin) & (c[:, nd - 1] < zmax)
            dist = c[:, 0] - x0[0]
            if nd == 3:
                inside &= (c[:, 1] > cfg["ic.weakzone_y_min"] * L[1]) & (c[:, 1] < cfg["ic.weakzone_y_max"] * L[1])
                dist = dist - az * (c[:, 1] - x0[1])
            dist = dist + incl * (c[:, nd - 1] - x0[nd - 1])
            inside &= np.abs(dist) < hw
            pl[inside] = cfg["ic.weakzone_plstrain"]
        elif opt == 2:
            x0 = [cfg["ic.weakzone_xcenter"] * L[0]]
            ax = [cfg["ic.weakzone_xsemi_axis"]]
            if nd == 3:
                x0.append(cfg["ic.weakzone_ycenter"] * L[1])
                ax.append(cfg["ic.weakzone_ysemi_axis"])
            x0.append(-cfg["ic.weakzone_zcenter"] * L[nd - 1])
            ax.append(cfg["ic.weakzone_zsemi_axis"])
            r = sum((c[:, d] - x0[d]) ** 2 / (ax[d] * ax[d]) for d in range(nd))
            pl[r < 1] = cfg["ic.weakzone_plstrain"]
        if plstrain_noise is not None:
            # extension: seeded uniform[0, amp] initial plastic strain per element
            amp, seed = plstrain_noise
            pl = pl + amp * splitmix_uniform(seed, E)
        self.plstrain0 = np.ascontiguousarray(pl)
        self.elem_visc = None
        if elem_visc is not None:
            lo, hi, seed = elem_visc       # log10 viscosity uniform [lo, hi]
            self.elem_visc = 10.0 ** (lo + (hi - lo) * splitmix_uniform(seed, E, 7))
        self.elem_cohesion = None
        if elem_cohesion is not None:
            kind, a, b, seed = elem_cohesion
            if kind == "normal":
                v = a + b * splitmix_normal(seed, E, 3)
                self.elem_cohesion = np.maximum(v, 1.0)
            else:
                self.elem_cohesion = a + (b - a) * splitmix_uniform(seed, E, 5)

    def bulkm_elem(self):
        """MatProps::bulkm for every element (matprops.cxx:263-266, 117-128)."""
        p = self.params
        if p.nmat == 1:
            return np.full(self.nelem, p.bulk_modulus[0])
        em = self.elemmarkers
        result = np.zeros(self.nelem)
        for i in range(p.nmat):
            result += em[:, i] / p.bulk_modulus[i]
        return em.sum(axis=1) / result

    def initial_dt(self):
        """compute_dt (geometry.cxx:239-310) of the initial mesh, single
        material, in numpy: the step size a fixed-dt run can be given."""
        p = self.params
        x = self.coord[self.conn]                      # E x NPE x ND
        if self.ndims == 3:
            e1, e2, e3 = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]
            vol = np.einsum("ij,ij->i", e1, np.cross(e2, e3)) / 6
            faces = [(0, 1, 2), (0, 1, 3), (2, 3, 0), (2, 3, 1)]
            area = np.max([np.linalg.norm(np.cross(x[:, b] - x[:, a], x[:, c] - x[:, a]), axis=1) / 2
                           for a, b, c in faces], axis=0)
            minh = 3 * vol / area
        else:
            ab, ac = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
            vol = np.abs(ab[:, 0] * ac[:, 1] - ab[:, 1] * ac[:, 0]) / 2
            maxl = np.sqrt(np.max([np.sum((x[:, j] - x[:, i]) ** 2, axis=1)
                                   for i, j in [(0, 1), (1, 2), (0, 2)]], axis=0))
            minh = 2 * vol / maxl
        minl = float(np.min(minh))
        dt_maxwell = 0.5 * p.visc_min / (1e-40 + p.shear_modulus[0])
        dt_diffusion = 0.5 * minl * minl / p.therm_diff_max if p.has_thermal_diffusion else math.inf
        dt_advection = 0.5 * minl / p.max_vbc_val
        if p.is_quasi_static:
            dt_elastic = 0.5 * minl / (p.max_vbc_val * p.inertial_scaling)
        else:
            dt_elastic = 0.5 * minl / math.sqrt(p.bulk_modulus[0] / p.rho0[0])
        return min(dt_elastic, dt_maxwell, dt_advection, dt_diffusion) * p.dt_fraction

    def mesh_struct(self):
        """C-ABI mesh struct; keeps references so the arrays outlive it."""
        m = abi.DesMesh()
        m.nnode, m.nelem = self.nnode, self.nelem
        self._keep = [np.ascontiguousarray(self.conn, dtype=np.int32),
                      np.ascontiguousarray(self.coord, dtype=np.float64)]
        m.connectivity = abi.iptr(self._keep[0])
        m.coord = abi.dptr(self._keep[1])
        m.bcflag = abi.uptr(self.bcflag)
        b, e, f = self.bfacets
        m.nbfacet = len(b)
        m.bfacet_bdry, m.bfacet_elem, m.bfacet_face = abi.iptr(b), abi.iptr(e), abi.iptr(f)
        m.bnormals = abi.dptr(self.bnormals)
        m.edge_vectors = abi.dptr(self.edge_vectors)
        m.edge_valid = abi.iptr(self.edge_valid)
        m.elemmarkers = abi.iptr(self.elemmarkers)
        m.markers_per_element = self.markers_per_element
        return m


def config_ensemble(n2d=128, n3d=128, seed=6, size_scale=1.0):
    """The batched ensemble: n2d 2D models with log-uniform 1k-50k triangles and
    n3d 3D models with log-uniform 10k-200k tetrahedra, each with its own
    extension rate U[0.5, 5] cm/yr, friction angle U[20, 35] deg and weak-seed
    x-centre U[0.3, 0.7], drawn from a counter-based generator keyed by seed."""
    n = n2d + n3d
    u = [splitmix_uniform(seed, n, k) for k in range(4)]
    models = []
    for i in range(n):
        nd = 2 if i < n2d else 3
        lo, hi = (1e3, 5e4) if nd == 2 else (1e4, 2e5)
        ne = int(round(size_scale * math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u[0][i])))
        rate = (0.5 + 4.5 * u[1][i]) * 1e-2 / YEAR2SEC
        phi = 20.0 + 15.0 * u[2][i]
        xc = 0.3 + 0.4 * u[3][i]
        cfg = {
            "mat.rheology_type": "elasto-visco-plastic",
            "mat.rho0": [2800.0], "mat.bulk_modulus": [50e9], "mat.shear_modulus": [30e9],
            "mat.cohesion0": [44e6], "mat.cohesion1": [4e6],
            "mat.friction_angle0": [phi], "mat.friction_angle1": [5.0],
            "mat.pls0": [0.0], "mat.pls1": [0.1],
            "mat.min_viscosity": 1e24, "mat.max_viscosity": 1e24,
            "mat.heat_capacity": [1000.0], "mat.therm_cond": [2.8],
            "bc.surface_temperature": 273.0, "bc.mantle_temperature": 1573.0,
            "bc.has_fixed_bottom_temperature": False,
            "bc.vbc_x0": 1, "bc.vbc_val_x0": -rate, "bc.vbc_x1": 1, "bc.vbc_val_x1": rate,
            "bc.vbc_z0": 1, "bc.vbc_val_z0": 0.0, "bc.vbc_z1": 0,
            "bc.has_winkler_foundation": False, "bc.has_water_loading": False,
            "mesh.xlength": 100e3, "mesh.zlength": 10e3, "mesh.resolution": 1e3,
            "ic.weakzone_xcenter": xc,
        }
        if nd == 2:
            cfg.update({"ic.weakzone_option": 2, "ic.weakzone_xsemi_axis": 2.5e3,
                        "ic.weakzone_zsemi_axis": 0.5e3, "ic.weakzone_zcenter": 0.5})
        else:
            cfg.update({"mesh.ylength": 100e3, "bc.vbc_y0": 1, "bc.vbc_val_y0": 0.0,
                        "bc.vbc_y1": 1, "bc.vbc_val_y1": 0.0,
                        "ic.weakzone_option": 1, "ic.weakzone_halfwidth": 2.5,
                        "ic.weakzone_depth_min": 0.45, "ic.weakzone_depth_max": 0.55})
        models.append(Model(cfg, nd, nelem_target=ne, temperature="linear", mesh_seed=i))
    return models


def config_adjoint(case="maxwell", nelem_target=200_000, log_visc=(22.0, 24.0)):
    """The differentiable configuration: 50 x 50 x 10 km box, per-element
    log10 viscosity U[22, 24] and cohesion U[20, 60] MPa (seed 3), visco-elastic
    ("maxwell") or elasto-visco-plastic ("plastic"), constant density
    (alpha = 0), 1 cm/yr extension, and a fixed dt equal to the initial
    compute_dt."""
    v = 1.0e-2 / YEAR2SEC
    cfg = {
        "mat.rheology_type": "maxwell" if case == "maxwell" else "elasto-visco-plastic",
        "mat.rho0": [2800.0], "mat.alpha": [0.0], "mat.bulk_modulus": [50e9], "mat.shear_modulus": [30e9],
        "mat.cohesion0": [44e6], "mat.cohesion1": [4e6],
        "mat.friction_angle0": [30.0], "mat.friction_angle1": [5.0],
        "mat.pls0": [0.0], "mat.pls1": [0.1],
        "mat.min_viscosity": 1e18, "mat.max_viscosity": 1e26,
        "mat.heat_capacity": [1000.0], "mat.therm_cond": [2.8],
        "bc.surface_temperature": 273.0, "bc.mantle_temperature": 1573.0,
        "bc.vbc_x0": 1, "bc.vbc_val_x0": -v, "bc.vbc_x1": 1, "bc.vbc_val_x1": v,
        "bc.vbc_y0": 1, "bc.vbc_val_y0": 0.0, "bc.vbc_y1": 1, "bc.vbc_val_y1": 0.0,
        "bc.vbc_z0": 1, "bc.vbc_val_z0": 0.0, "bc.vbc_z1": 0,
        "bc.has_winkler_foundation": False, "bc.has_water_loading": False,
        "control.is_quasi_static": True, "control.inertial_scaling": 1e5,
        "control.damping_option": 1, "control.damping_factor": 0.8, "control.gravity": 10.0,
        "mesh.xlength": 50e3, "mesh.ylength": 50e3, "mesh.zlength": 10e3, "mesh.resolution": 1e3,
        "mesh.quality_check_step_interval": 1 << 30,
        "ic.weakzone_option": 1, "ic.weakzone_halfwidth": 2.5,
        "ic.weakzone_depth_min": 0.45, "ic.weakzone_depth_max": 0.55,
    }
    m = Model(cfg, 3, nelem_target=nelem_target, temperature="linear",
              elem_visc=(log_visc[0], log_visc[1], 3), elem_cohesion=("uniform", 20e6, 60e6, 3))
    m.params.fixed_dt = m.initial_dt()
    return m


def config_benchmark(which, nelem_target=None, mesh="structured"):
    """The workloads the benchmark and the tests are quoted on.

    which = "2d-extension": 100x10 km box, elasto-visco-plastic Mohr-Coulomb,
            linear temperature, +-1 cm/yr side walls (the reference's CPU case);
    which = "3d-fp64":      100x100x10 km box, same material, weak slab plus
            seeded plastic-strain noise;
    which = "3d-mixed":     200x200x20 km box, per-element viscosity and cohesion.
    mesh = "structured" (jittered Kuhn tetrahedra / triangles), "delaunay"
    (the Delaunay mesh of a jittered lattice, node degrees as a mesher's) or
    "delaunay-joggled" (the same built with Qhull's joggle, for large meshes).
    """
    v = 1.0e-2 / YEAR2SEC          # 1 cm/yr
    base = {
        "mat.rheology_type": "elasto-visco-plastic",
        "mat.rho0": [2800.0], "mat.bulk_modulus": [50e9], "mat.shear_modulus": [30e9],
        "mat.cohesion0": [44e6], "mat.cohesion1": [4e6],
        "mat.friction_angle0": [30.0], "mat.friction_angle1": [5.0],
        "mat.pls0": [0.0], "mat.pls1": [0.1],
        "mat.dilation_angle0": [0.0], "mat.dilation_angle1": [0.0],
        "mat.min_viscosity": 1e24, "mat.max_viscosity": 1e24,
        "mat.heat_capacity": [1000.0], "mat.therm_cond": [2.8],
        "bc.surface_temperature": 273.0, "bc.mantle_temperature": 1573.0,
        "bc.has_fixed_bottom_temperature": False,
        "bc.vbc_x0": 1, "bc.vbc_val_x0": -v, "bc.vbc_x1": 1, "bc.vbc_val_x1": v,
        "bc.vbc_z0": 1, "bc.vbc_val_z0": 0.0, "bc.vbc_z1": 0,
        "bc.has_winkler_foundation": False, "bc.has_water_loading": False,
        "control.is_quasi_static": True, "control.inertial_scaling": 1e5,
        "control.damping_option": 1, "control.damping_factor": 0.8,
        "control.gravity": 10.0,
        "mesh.quality_check_step_interval": 1 << 30,
    }
    if which == "2d-extension":
        cfg = dict(base)
        cfg.update({"mesh.xlength": 100e3, "mesh.zlength": 10e3, "mesh.resolution": 1e3,
                    "ic.weakzone_option": 2, "ic.weakzone_xsemi_axis": 2.5e3,
                    "ic.weakzone_zsemi_axis": 0.5e3, "ic.weakzone_zcenter": 0.5})
        return Model(cfg, 2, nelem_target=nelem_target or 4000, temperature="linear", mesh=mesh)
    if which == "3d-fp64":
        cfg = dict(base)
        cfg.update({"mesh.xlength": 100e3, "mesh.ylength": 100e3, "mesh.zlength": 10e3,
                    "mesh.resolution": 1e3,
                    "bc.vbc_y0": 1, "bc.vbc_val_y0": 0.0, "bc.vbc_y1": 1, "bc.vbc_val_y1": 0.0,
                    "ic.weakzone_option": 1, "ic.weakzone_halfwidth": 2.5,
                    "ic.weakzone_y_min": 0.0, "ic.weakzone_y_max": 1.0,
                    "ic.weakzone_depth_min": 0.45, "ic.weakzone_depth_max": 0.55})
        return Model(cfg, 3, nelem_target=nelem_target or 1_000_000, temperature="linear",
                     plstrain_noise=(0.02, 1), mesh=mesh)
    if which == "3d-mixed":
        cfg = dict(base)
        cfg.update({"mesh.xlength": 200e3, "mesh.ylength": 200e3, "mesh.zlength": 20e3,
                    "mesh.resolution": 1e3,
                    "mat.min_viscosity": 1e18, "mat.max_viscosity": 1e26,
                    "bc.vbc_y0": 1, "bc.vbc_val_y0": 0.0, "bc.vbc_y1": 1, "bc.vbc_val_y1": 0.0,
                    "ic.weakzone_option": 1, "ic.weakzone_halfwidth": 2.5,
                    "ic.weakzone_depth_min": 0.45, "ic.weakzone_depth_max": 0.55})
        return Model(cfg, 3, nelem_target=nelem_target or 8_000_000, temperature="linear",
                     elem_visc=(21.0, 25.0, 2), elem_cohesion=("normal", 44e6, 4.4e6, 2), mesh=mesh)
    raise ValueError(which)
