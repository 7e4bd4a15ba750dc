"""Restart of a device model from a frame and its checkpoint, following the
reference's restart() (dynearthsol.cxx:104-218) then the compute_dt that
precedes the main loop (dynearthsol.cxx:316):

  1. steps, nnode, nelem, nseg from the frame's .info line;
  2. coordinates and connectivity from .save, boundary segments from .chkpt;
     the boundary flags ("bcflag" of .save), boundary facets and normals from
     the mesh (create_boundary_flags/facets/normals);
  3. coord0 from .save; volume computed, volume_old from .chkpt, then mass,
     thermal mass and nodal volume computed while the temperature is still
     zero (the reference reads the temperature only afterwards), shape
     functions, and the velocity BCs applied to the zero velocity;
  4. velocity, temperature, strain-rate, strain, stress, plastic strain (and
     in plane strain stressyy) from the files, time and compensation pressure
     from .chkpt, force from .save;
  5. compute_dt.

Marker sets stay with the host: the element marker counts are passed in
(elemmarkers, for several materials), as the host's markers give them.
"""
import numpy as np

from . import abi, model as M, output
from .lib import DesModel


class RestartedModel:
    """What DesModel needs of a model, from the files."""

    def __init__(self, params, coord, conn, bcflag, markers_per_element, elemmarkers):
        self.params = params
        self.ndims = nd = params.ndims
        self.npe, self.nstr = nd + 1, 6 if nd == 3 else 3
        self.coord = np.ascontiguousarray(coord, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.nnode, self.nelem = len(self.coord), len(self.conn)
        self.bcflag = np.ascontiguousarray(bcflag, dtype=np.uint32)
        self.bfacets = M.boundary_facets(self.conn, self.bcflag, nd)
        self.bnormals, self.edge_vectors, self.edge_valid = M.boundary_normals(
            self.coord, self.conn, self.bfacets, nd)
        self.markers_per_element = markers_per_element
        self.elemmarkers = None if elemmarkers is None else np.ascontiguousarray(elemmarkers, dtype=np.int32)
        self.elem_visc = self.elem_cohesion = None

    mesh_struct = M.Model.mesh_struct


def read_restart_files(modelname, frame):
    """(info row, .save arrays, .chkpt arrays) of one frame."""
    rows = [r for r in output.read_info(f"{modelname}.info") if r[0] == frame]
    if not rows:
        raise ValueError(f"frame {frame} is not in {modelname}.info")
    info = rows[-1]
    nnode, nelem = info[5], info[6]
    _, _, save = output.read_frame(f"{modelname}.save.{frame:06d}", nnode, nelem)
    chk = output.read_checkpoint(f"{modelname}.chkpt.{frame:06d}")
    return info, save, chk


def restart(params, modelname, frame, markers_per_element=4, elemmarkers=None, elem_visc=None,
            elem_cohesion=None):
    """A DesModel on the current device in the state restart() leaves, and
    the boundary segments and raw marker arrays of the checkpoint (the host's).
    elem_visc / elem_cohesion: the per-element fields of a model that has them
    (the files do not hold them)."""
    info, save, chk = read_restart_files(modelname, frame)
    steps = info[1]
    m = RestartedModel(params, save["coordinate"], save["connectivity"], save["bcflag"].view(np.uint32),
                       markers_per_element, elemmarkers)
    g = DesModel(m, init=False)
    g.upload(coord0=save["coord0"], elem_visc=elem_visc, elem_cohesion=elem_cohesion)
    g.call("compute_volume")
    g.upload(volume_old=chk["volume_old"])
    g.call("compute_mass")            # temperature still zero, as in the reference
    g.call("compute_shape_fn")
    g.call("apply_vbcs")
    fields = {"vel": save["velocity"], "temperature": save["temperature"], "strain_rate": save["strain-rate"],
              "strain": save["strain"], "stress": save["stress"], "plstrain": save["plastic strain"],
              "force": save["force"]}
    if params.is_plane_strain:
        fields["stressyy"] = chk["stressyy"]
    g.upload(**fields)
    time_, cp = chk["time compensation_pressure"]
    g._check(g.lib.des_set_scalars(g.ctx, 0.0, float(time_), int(steps), float(cp)))
    g.compute_dt()
    extra = {k: v for k, v in chk.items() if k not in output.CHKPT_DTYPES}
    return g, {"segment": chk["segment"], "segflag": chk["segflag"], "markers": extra, "steps": steps}
