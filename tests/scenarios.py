"""Small models that drive every branch of the step: plastic yielding (shear
and tensile), power-law viscosity, each rheology, stress boundary conditions,
damping options, plane strain, several materials, per-element fields."""
import numpy as np

from dynearthsol_amd import model as M

V1 = 1.0e-2 / M.YEAR2SEC    # 1 cm/yr

BASE = {
    "mat.rheology_type": "elasto-visco-plastic",
    "mat.rho0": [2800.0], "mat.bulk_modulus": [50e9], "mat.shear_modulus": [30e9],
    "mat.cohesion0": [10e6], "mat.cohesion1": [1e6],
    "mat.friction_angle0": [30.0], "mat.friction_angle1": [5.0],
    "mat.dilation_angle0": [5.0], "mat.dilation_angle1": [0.0],
    "mat.pls0": [0.0], "mat.pls1": [0.1],
    "mat.min_viscosity": 1e18, "mat.max_viscosity": 1e26,
    "mat.heat_capacity": [1000.0], "mat.therm_cond": [2.8],
    "mat.radiogenic_heating": [1e-9],
    "bc.surface_temperature": 273.0, "bc.mantle_temperature": 1573.0,
    "bc.has_fixed_bottom_temperature": False,
    "bc.vbc_x0": 1, "bc.vbc_val_x0": -V1, "bc.vbc_x1": 1, "bc.vbc_val_x1": V1,
    "bc.vbc_z0": 1, "bc.vbc_val_z0": 0.0, "bc.vbc_z1": 0,
    "bc.has_winkler_foundation": False, "bc.has_water_loading": False,
    "control.inertial_scaling": 1e3,
    "ic.oceanic_plate_age_in_yr": 20e6,
}


def _perturb_stress(m, amp, seed):
    """Seeded deviatoric perturbation of the initial stress, so that many
    elements start near or beyond yield."""
    E, nstr, nd = m.nelem, m.nstr, m.ndims
    r = np.stack([M.splitmix_uniform(seed, E, 100 + k) * 2 - 1 for k in range(nstr)], axis=1)
    r[:, :nd] -= r[:, :nd].mean(axis=1)[:, None]      # deviatoric
    m.stress0 += amp * r


def _layers(m):
    """Two materials: upper half material 0, lower half material 1, with a
    mixed band of elements in between."""
    c = m.coord[m.conn].mean(axis=1)[:, m.ndims - 1]
    zl = m.lengths[-1]
    em = np.zeros((m.nelem, m.params.nmat), dtype=np.int32)
    mpe = m.markers_per_element
    upper = c > -0.45 * zl
    lower = c < -0.55 * zl
    mixed = ~(upper | lower)
    em[upper, 0] = mpe
    em[lower, 1] = mpe
    k = (M.splitmix_uniform(9, m.nelem) * (mpe + 1)).astype(np.int32)
    em[mixed, 0] = k[mixed]
    em[mixed, 1] = mpe - k[mixed]
    return em


def _slant(m, shear):
    """Shear the box (x -> x - shear * z, z <= 0) so that its x walls become
    slanted planes, and flag them as the slant boundaries n0 (the former x1
    wall) and n1 (the former x0 wall) with their normals and the edge vectors
    of their intersections (bc.cxx:87-176); the velocity BCs then take the
    slant branch of apply_vbcs (bc.cxx:350-436)."""
    nd = m.ndims
    m.coord = m.coord.copy()
    m.coord[:, 0] -= shear * m.coord[:, nd - 1]
    f = m.bcflag.copy()
    x0, x1 = (f & M.BOUNDX0) != 0, (f & M.BOUNDX1) != 0
    f &= ~np.uint32(M.BOUNDX0 | M.BOUNDX1)
    f[x1] |= np.uint32(1 << 6)
    f[x0] |= np.uint32(1 << 7)
    m.bcflag = f
    m.bfacets = M.boundary_facets(m.conn, m.bcflag, nd)
    m.bnormals, m.edge_vectors, m.edge_valid = M.boundary_normals(m.coord, m.conn, m.bfacets, nd)


def build(name, nelem=None, **mesh_kw):
    cfg = dict(BASE)
    kw = {}
    nd = 3 if name.endswith("3") else 2
    if nd == 3:
        cfg.update({"mesh.xlength": 40e3, "mesh.ylength": 20e3, "mesh.zlength": 10e3,
                    "mesh.resolution": 2e3,
                    "bc.vbc_y0": 1, "bc.vbc_val_y0": 0.0, "bc.vbc_y1": 1, "bc.vbc_val_y1": 0.0,
                    "ic.weakzone_option": 1, "ic.weakzone_halfwidth": 1.5,
                    "ic.weakzone_depth_min": 0.3, "ic.weakzone_depth_max": 0.7})
        ne = nelem or 4000
    else:
        cfg.update({"mesh.xlength": 40e3, "mesh.zlength": 10e3, "mesh.resolution": 1e3,
                    "ic.weakzone_option": 2, "ic.weakzone_xsemi_axis": 3e3,
                    "ic.weakzone_zsemi_axis": 1e3, "ic.weakzone_zcenter": 0.5})
        ne = nelem or 3000
    perturb = 30e6
    if name.startswith("evp"):
        pass
    elif name.startswith("ep"):
        cfg["mat.rheology_type"] = "elasto-plastic"
    elif name.startswith("psevp"):
        cfg["mat.is_plane_strain"] = True
    elif name.startswith("ps"):        # plane-strain elasto-plastic (2D only)
        cfg["mat.rheology_type"] = "elasto-plastic"
        cfg["mat.is_plane_strain"] = True
    elif name.startswith("maxwell"):
        cfg["mat.rheology_type"] = "maxwell"
        cfg["mat.min_viscosity"] = 1e19
        cfg["mat.max_viscosity"] = 1e25
    elif name.startswith("elastic"):
        cfg["mat.rheology_type"] = "elastic"
        cfg["control.is_quasi_static"] = False
        cfg["control.damping_option"] = 0
        cfg["mat.alpha"] = [0.0]
        perturb = 1e6
    elif name.startswith("viscous"):
        cfg["mat.rheology_type"] = "viscous"
        cfg["mat.min_viscosity"] = 1e20
        cfg["mat.max_viscosity"] = 1e22
        # explicit viscous stepping is stable for eta below ~2 h K / (v s)
        # (h element size, K bulk modulus, v the boundary speed, s the
        # inertial scaling): 6e20 Pa s for the 3D mesh, 3e20 for the 2D one
        cfg["mat.min_viscosity"] = 1e19
        cfg["mat.max_viscosity"] = 2e20 if nd == 3 else 1e20
        perturb = 0
    elif name.startswith("bc"):
        # Winkler bottom, water-loaded top, lithostatic and free-slip side walls,
        # elastic foundation, damping option 3
        cfg.update({"bc.has_winkler_foundation": True, "bc.winkler_delta_rho": 10.0,
                    "bc.has_water_loading": True, "bc.has_elastic_foundation": True,
                    "bc.elastic_foundation_constant": 1e7,
                    "bc.vbc_x0": 0, "bc.vbc_x1": 1, "bc.vbc_val_x1": 3 * V1,
                    "control.damping_option": 3})
        if nd == 3:
            cfg.update({"bc.vbc_y0": 2, "bc.vbc_y1": 0})
    elif name.startswith("damp2"):
        cfg.update({"control.damping_option": 2, "bc.vbc_x0": 3, "bc.vbc_x1": 3,
                    "bc.vbc_val_x1": 2 * V1, "bc.vbc_z0": 3})
        if nd == 3:
            cfg.update({"bc.vbc_y0": 5, "bc.vbc_val_y0": V1, "bc.vbc_y1": 7, "bc.vbc_val_y1": 0.0,
                        "bc.vbc_x0": 4, "bc.vbc_val_x0": 0.5 * V1})
    elif name.startswith("mat") or name.startswith("dynmat"):
        # two materials, PREM reference pressure for initial stress; dynmat:
        # the same without quasi-static mass scaling (compute_dt's elastic
        # term then reads bulkm and rho of element 0)
        if name.startswith("dynmat"):
            cfg.update({"control.is_quasi_static": False, "control.damping_option": 0})
        cfg.update({"mat.num_materials": 2, "mat.rho0": [2700.0, 3300.0],
                    "mat.bulk_modulus": [40e9, 80e9], "mat.shear_modulus": [25e9, 45e9],
                    "mat.cohesion0": [10e6, 20e6], "mat.cohesion1": [1e6, 2e6],
                    "mat.friction_angle0": [30.0, 25.0], "mat.friction_angle1": [5.0, 10.0],
                    "mat.visc_exponent": [3.0, 3.5], "mat.alpha": [3e-5, 2e-5],
                    "mat.therm_cond": [2.5, 3.3], "mat.heat_capacity": [900.0, 1100.0],
                    "control.ref_pressure_option": 1,
                    "bc.vbc_y0": 0, "bc.vbc_y1": 0} if nd == 3 else
                   {"mat.num_materials": 2, "mat.rho0": [2700.0, 3300.0],
                    "mat.bulk_modulus": [40e9, 80e9], "mat.shear_modulus": [25e9, 45e9],
                    "mat.cohesion0": [10e6, 20e6], "mat.cohesion1": [1e6, 2e6],
                    "mat.friction_angle0": [30.0, 25.0], "mat.friction_angle1": [5.0, 10.0],
                    "mat.visc_exponent": [3.0, 3.5], "mat.alpha": [3e-5, 2e-5],
                    "mat.therm_cond": [2.5, 3.3], "mat.heat_capacity": [900.0, 1100.0],
                    "control.ref_pressure_option": 2, "bc.vbc_z0": 0,
                    "bc.has_winkler_foundation": True})
        kw["elemmarkers"] = _layers
    elif name.startswith("fields"):
        kw["elem_visc"] = (21.0, 25.0, 2)
        kw["elem_cohesion"] = ("normal", 10e6, 1e6, 2)
        kw["plstrain_noise"] = (0.02, 1)
    elif name.startswith("dyn"):
        cfg.update({"control.is_quasi_static": False, "control.damping_option": 0})
    elif name.startswith("slant"):
        # (visco-elastic, so that trajectories do not amplify rounding) slant
        # x walls: n0 (outer, former x1) pulled out along its normal at 1 cm/yr, n1 (former x0) held by the horizontal-normal rule (type
        # 11; slanttb: v along the normal, type 3, and horizontal, type 13);
        # free-slip y walls and bottom meet them at edges
        t0, t1 = (3, 13) if name.startswith("slanttb") else (1, 11)
        cfg.update({"mat.rheology_type": "maxwell", "mat.min_viscosity": 1e19, "mat.max_viscosity": 1e25,
                    "bc.vbc_x0": 0, "bc.vbc_x1": 0, "bc.vbc_n0": t0, "bc.vbc_val_n0": V1,
                    "bc.vbc_n1": t1, "bc.vbc_val_n1": V1, "bc.vbc_z0": 1, "bc.vbc_val_z0": 0.0})
    else:
        raise ValueError(name)
    m = M.Model(cfg, nd, nelem_target=ne, **kw, **mesh_kw)
    if name.startswith("slant"):
        _slant(m, 0.3)
    if perturb:
        _perturb_stress(m, perturb, 11)
    return m


SCENARIOS = ["evp3", "evp2", "ep3", "ps2", "psevp2", "maxwell3", "maxwell2", "elastic3",
             "elastic2", "viscous3", "viscous2", "bc3", "bc2", "damp23", "damp22", "mat3",
             "mat2", "fields3", "dyn3", "dyn2", "dynmat3", "dynmat2", "slant3", "slant2", "slanttb3",
             "slanttb2"]


def rel_err(a, b):
    """max |a-b| / max |b| (normwise relative error of a field)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if b.size else 0.0
    if s == 0.0:
        return d
    return d / s
