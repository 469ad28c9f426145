"""Scene flattener: turns an AOpticsManager's volume tree into the flat
structure-of-arrays scene of include/rbgpu.h (rb_scene_desc).

Every placed node instance (a path through the TGeo tree) becomes one
rb_node_desc with its global matrix; node 0 is the top volume. Volumes,
shapes (composite expressions flattened into boolean shape records) and
TGraph tables are de-duplicated by identity.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from . import geometry as G
from .materials import ARefractiveIndex, ASchottFormula, ASellmeierFormula

# exact-class classification, as AOpticsManager::Is*() compares IsA()
# (include/AOpticsManager.h:76-90): subclasses do not match
_KIND = {
    G.ALens: A.RB_KLENS,
    G.AObscuration: A.RB_KOBS,
    G.AMirror: A.RB_KMIRROR,
    G.AFocalSurface: A.RB_KFOCUS,
    G.AOpticalComponent: A.RB_KOPT,
}


def volume_kind(vol):
    return _KIND.get(type(vol), A.RB_KOTHER)


class FlatScene:
    """Owns the numpy arrays behind an rb_scene_desc."""

    def __init__(self, nodes, children, volumes, shapes, params, tables, table_xy, node_names, volume_objs,
                 borders=None, grids=None):
        if borders is None:
            borders = np.zeros(0, dtype=A.BORDER_DTYPE)
        if grids is None:
            grids = np.zeros(0, dtype=A.GRID_DTYPE)
        self.borders = borders
        self.grids = grids
        self.nodes = nodes
        self.children = children
        self.volumes = volumes
        self.shapes = shapes
        self.params = params
        self.tables = tables
        self.table_xy = table_xy
        self.node_names = node_names
        self.volume_objs = volume_objs
        self.coating_grid = None  # (nx, lambda lo, lambda hi) of default coating tables, if any
        d = A.rb_scene_desc()
        d.n_nodes, d.nodes = len(nodes), A.ptr(nodes)
        d.n_children, d.children = len(children), A.ptr(children)
        d.n_volumes, d.volumes = len(volumes), A.ptr(volumes)
        d.n_shapes, d.shapes = len(shapes), A.ptr(shapes)
        d.n_params, d.params = len(params), A.ptr(params)
        d.n_tables, d.tables = len(tables), A.ptr(tables)
        d.n_table_xy, d.table_xy = len(table_xy), A.ptr(table_xy)
        d.n_borders, d.borders = len(borders), (A.ptr(borders) if len(borders) else None)
        d.n_grids, d.grids = len(grids), (A.ptr(grids) if len(grids) else None)
        self.desc = d

    def desc_ptr(self):
        return C.byref(self.desc)


def flatten(top_volume, lambda_range=None):
    """The closed geometry under top_volume as an rb_scene_desc (FlatScene).
    lambda_range (cm): the wavelengths the default tables of coatings that
    were not precalculated must cover (multilayer.default_grid)."""
    shapes, shape_ids, params = [], {}, []
    tables, table_ids, table_xy = [], {}, []
    volumes, volume_ids, volume_objs = [], {}, []
    grid_recs, grid_ids = [], {}

    def add_grid(key, nx, xlo, xhi, ny, ylo, yhi, r_values, t_values=None):
        """A TH2-style table of bin-centre values (multilayer R and T, or a
        mirror's TH2 reflectance), shared by every user of the same object."""
        if key in grid_ids:
            return grid_ids[key]
        ro = add_params(np.ravel(r_values))
        to = ro if t_values is None else add_params(np.ravel(t_values))
        grid_ids[key] = len(grid_recs)
        grid_recs.append((nx, xlo, xhi, ny, ylo, yhi, ro, to))
        return grid_ids[key]

    def add_params(vals):
        off = len(params)
        params.extend(float(v) for v in vals)
        return off

    def add_table(g):
        if g is None:
            return -1
        if id(g) in table_ids:
            return table_ids[id(g)]
        xs, ys = g.sorted_points()
        idx = len(tables)
        tables.append((len(table_xy) // 2, len(xs)))
        for x, y in zip(xs, ys):
            table_xy.extend((x, y))
        for x, y in g.extrapolation_points():
            table_xy.extend((x, y))
        table_ids[id(g)] = idx
        return idx

    def add_shape(sh):
        if id(sh) in shape_ids:
            return shape_ids[id(sh)]
        p = sh.params()
        if sh.shape_type == A.RB_SHAPE_COMPOSITE:
            left, right = sh.children()
            p[1] = float(add_shape(left))
            p[2] = float(add_shape(right))
        idx = len(shapes)
        org, half = sh.local_bbox()
        shapes.append((sh.shape_type, add_params(p), len(p), org, half))
        shape_ids[id(sh)] = idx
        return idx

    def add_volume(vol):
        if id(vol) in volume_ids:
            return volume_ids[id(vol)]
        kind = volume_kind(vol)
        rec = dict(refl_const=1.0, shape=add_shape(vol.shape), kind=kind, refl_table=-1,
                   index_model=A.RB_INDEX_NONE, index_ref=-1, k_table=-1,
                   qe_lambda_table=-1, qe_angle_table=-1, refl_grid=-1)
        if kind == A.RB_KMIRROR:
            rec["refl_const"] = vol.reflectance
            rec["refl_table"] = add_table(vol.reflectance_graph)
            h = getattr(vol, "reflectance_th2", None)
            if h is not None:
                rec["refl_grid"] = add_grid(id(h), h.nx, h.xlo, h.xhi, h.ny, h.ylo, h.yhi, h.contents())
        elif kind == A.RB_KLENS and vol.index is not None:
            ix = vol.index
            if isinstance(ix, ASellmeierFormula):
                rec["index_model"] = A.RB_INDEX_SELLMEIER
                rec["index_ref"] = add_params(ix.par)
            elif isinstance(ix, ASchottFormula):
                rec["index_model"] = A.RB_INDEX_SCHOTT
                rec["index_ref"] = add_params(ix.par)
            elif isinstance(ix, ARefractiveIndex):
                if ix.n_graph is not None:
                    rec["index_model"] = A.RB_INDEX_TABLE
                    rec["index_ref"] = add_table(ix.n_graph)
                else:
                    # an index object without an n graph evaluates to n = 1
                    rec["index_model"] = A.RB_INDEX_TABLE
                    one = _const_graph(1.0)
                    rec["index_ref"] = add_table(one)
            else:
                raise TypeError(f"unsupported refractive-index model {type(ix).__name__}")
            rec["k_table"] = add_table(ix.k_graph)
        elif kind == A.RB_KFOCUS:
            rec["qe_lambda_table"] = add_table(vol.qe_lambda)
            rec["qe_angle_table"] = add_table(vol.qe_angle)
        idx = len(volumes)
        volumes.append(rec)
        volume_ids[id(vol)] = idx
        volume_objs.append(vol)
        return idx

    # breadth-first over placed nodes so each node's children are contiguous
    node_recs = []  # (volume idx, parent, global matrix, name)
    child_lists = []
    queue = [(top_volume, -1, G.IDENTITY, top_volume.name + "_1")]
    node_recs.append((add_volume(top_volume), -1, G.IDENTITY, top_volume.name + "_1"))
    child_lists.append([])
    vols_of = [top_volume]
    head = 0
    while head < len(node_recs):
        vol = vols_of[head]
        gmat = node_recs[head][2]
        for (dvol, copy_no, mat) in vol.nodes:
            g = gmat.compose(mat)
            idx = len(node_recs)
            node_recs.append((add_volume(dvol), head, g, f"{dvol.name}_{copy_no}"))
            vols_of.append(dvol)
            child_lists.append([])
            child_lists[head].append(idx)
        head += 1
    del queue

    nodes = np.zeros(len(node_recs), dtype=A.NODE_DTYPE)
    children = []
    for i, (v, parent, g, _name) in enumerate(node_recs):
        nodes[i]["rot"] = g.rot.reshape(9)
        nodes[i]["tr"] = g.tr
        nodes[i]["volume"] = v
        nodes[i]["parent"] = parent
        nodes[i]["first_child"] = len(children)
        nodes[i]["n_children"] = len(child_lists[i])
        children.extend(child_lists[i])
    children = np.array(children if children else [0], dtype=np.int32)

    vol_arr = np.zeros(len(volumes), dtype=A.VOLUME_DTYPE)
    for i, rec in enumerate(volumes):
        for k, v in rec.items():
            vol_arr[i][k] = v

    # border-surface conditions, per component1 volume in registration order;
    # a component2 that is not placed can never be entered (-2 matches nothing).
    # A multilayer becomes a (wavelength, angle) table of R and T in params,
    # shared by every condition that holds the same AMultilayer object.
    border_recs = []
    coating_grid = None  # the default grid of the coatings that were not precalculated
    for i, vol in enumerate(volume_objs):
        conds = getattr(vol, "border_conditions", [])
        vol_arr[i]["border_first"] = len(border_recs)
        vol_arr[i]["border_count"] = len(conds)
        for c in conds:
            v2 = -1 if c.component2 is None else volume_ids.get(id(c.component2), -2)
            g = -1
            if c.multilayer is not None:
                g = grid_ids.get(id(c.multilayer))
                if g is None:
                    nx, xlo, xhi, ny, ylo, yhi, R, T = c.multilayer.device_table(lambda_range)
                    g = add_grid(id(c.multilayer), nx, xlo, xhi, ny, ylo, yhi, R, T)
                    if c.multilayer.table is None:
                        coating_grid = (nx, xlo, xhi)
            border_recs.append((v2, 1 if c.lambertian else 0, c.sigma, g))
    borders = np.zeros(len(border_recs), dtype=A.BORDER_DTYPE)
    for i, (v2, lam, sig, g) in enumerate(border_recs):
        borders[i]["volume2"], borders[i]["lambertian"], borders[i]["sigma"] = v2, lam, sig
        borders[i]["multilayer"] = g
    grids = np.zeros(len(grid_recs), dtype=A.GRID_DTYPE)
    for i, (nx, xlo, xhi, ny, ylo, yhi, ro, to) in enumerate(grid_recs):
        grids[i]["nx"], grids[i]["xlo"], grids[i]["xhi"] = nx, xlo, xhi
        grids[i]["ny"], grids[i]["ylo"], grids[i]["yhi"] = ny, ylo, yhi
        grids[i]["r_offset"], grids[i]["t_offset"] = ro, to

    shp = np.zeros(len(shapes), dtype=A.SHAPE_DTYPE)
    for i, (t, off, n, org, half) in enumerate(shapes):
        shp[i]["type"], shp[i]["param_offset"], shp[i]["n_params"] = t, off, n
        shp[i]["bbox_origin"], shp[i]["bbox_half"] = org, half

    tab = np.zeros(max(len(tables), 1), dtype=A.TABLE_DTYPE)
    for i, (off, n) in enumerate(tables):
        tab[i]["offset"], tab[i]["n"] = off, n

    par = np.array(params if params else [0.0], dtype=np.float64)
    txy = np.array(table_xy if table_xy else [0.0, 0.0], dtype=np.float64)
    fs = FlatScene(nodes, children, vol_arr, shp, par, tab, txy, [r[3] for r in node_recs], volume_objs, borders,
                   grids)
    fs.desc.n_tables = len(tables)
    fs.desc.n_table_xy = len(table_xy)
    fs.coating_grid = coating_grid
    return fs


def _const_graph(v):
    from .materials import TGraph
    g = TGraph()
    g.SetPoint(0, 0.0, v)
    return g
