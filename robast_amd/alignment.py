"""Differentiable Davies-Cotton alignment.

The reference fits mirror alignment by re-tracing the whole geometry for
every trial parameter set under Minuit, with no gradient
(tutorials/Optimize.C:76-129, loss Sum GetSpotSize^2 at :32-48,128). Here the
same kind of fit gets an exact gradient: the loss is the sum over incidence
angles of the focal-plane spot RMS^2 (x, y about the spot centroid), traced
on the GPU, and its gradient with respect to

  * per-facet tip and tilt (radians; a rotation Rx(tip) Ry(tilt) applied in the
    facet's own frame after its Davies-Cotton placement),
  * per-facet radius of curvature (cm),
  * the camera z offset (cm),

comes from the `rbgpu_adjoint` kernel, which recomputes each focused ray's
bounce chain and back-propagates dL/dhit through the reflections.

Parameter vector layout (256 for the 85-facet array): p[2i] = tip_i,
p[2i+1] = tilt_i, p[2F + i] = R_i, p[3F] = camera dz.

`SpotLoss` is a torch.autograd.Function: its forward runs one adjoint launch
over all angles that returns the loss and the gradient together (the moments
form of rbgpu_adjoint, one group of sums per angle); backward hands the
gradient on. Per angle k with N_k focused rays, centroid c_k and hits h,
L = sum_k (1/N_k) sum |h - c_k|^2, so dL/dh = (2/N_k)(h - c_k) (the centroid's
own dependence drops out because sum (h - c_k) = 0).
"""
import math

import numpy as np
import torch

from . import _abi as A
from . import configs
from ._lib import Context, make_rays_device

C5_ANGLES_DEG = tuple(0.5 * k for k in range(8))   # DaviesCotton.C:185,196 (0 .. 3.5 deg)
C5_RAYS_PER_ANGLE = 1_250_000
C5_NFACETS = configs.C2_NFACETS
C5_NPARAMS = 3 * C5_NFACETS + 1
C5_LAMBDA = 400 * configs.nm

_FIELDS_F64 = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
_FIELDS_I32 = ("status", "npoints", "last_node")


def c5_initial_params(seed=0xC5, nfacets=C5_NFACETS):
    """Tip/tilt ~ N(0, 0.5 mrad), R = 3200 cm (1 + U[-1%, 1%]), camera dz = 0."""
    rng = np.random.Generator(np.random.Philox(seed))
    p = np.zeros(3 * nfacets + 1)
    p[: 2 * nfacets] = rng.normal(0.0, 0.5e-3, 2 * nfacets)
    p[2 * nfacets: 3 * nfacets] = configs.C2_MIRROR_R * (1.0 + rng.uniform(-0.01, 0.01, nfacets))
    return p


def c5_param_scale(nfacets=C5_NFACETS):
    """Natural size of each parameter (1 mrad, 1% of R, 1 cm), used to
    precondition the optimizer."""
    s = np.empty(3 * nfacets + 1)
    s[: 2 * nfacets] = 1e-3
    s[2 * nfacets: 3 * nfacets] = 0.01 * configs.C2_MIRROR_R
    s[3 * nfacets] = 1.0
    return s


def c5_scene(params, nfacets=C5_NFACETS):
    """The C2 Davies-Cotton with per-facet tip/tilt/R and camera dz taken from
    `params`, unit reflectance (deterministic)."""
    p = np.asarray(params, dtype=float)
    tt = p[: 2 * nfacets].reshape(nfacets, 2)
    return configs.c2_scene(deterministic=True, per_facet=True, facet_radius=list(p[2 * nfacets: 3 * nfacets]),
                            facet_tilt=[(float(a), float(b)) for a, b in tt], camera_dz=float(p[3 * nfacets]))


def c5_raygen(angle_deg, index):
    """Parallel beam at `angle_deg` off axis in the x-z plane, uniform over a
    12.5 m disk, 400 nm; one Philox seed per angle."""
    g = configs.c2_raygen(seed=0xC5_00 + index)
    g.theta_min = g.theta_max = angle_deg * math.pi / 180.0
    g.lambda_min = g.lambda_max = C5_LAMBDA
    return g


class FacetSceneUpdater:
    """Moves an existing per-facet scene to new parameter values by patching
    its flattened arrays in place (facet rotations, sphere radii, the
    sphere-to-cap translation and the affected bounding boxes), through the
    same shape code a rebuild uses, so the arrays equal a rebuild's bit for
    bit. Tip/tilt act as M = M0 Rx(tip) Ry(tilt) on the facet's placement."""

    def __init__(self, mgr, nfacets=C5_NFACETS):
        self.flat = flat = mgr.flat()
        self.nfacets = nfacets
        index = {name: k for k, name in enumerate(flat.node_names)}
        self.nodes = [index[f"mirror{i}_{i + 1}"] for i in range(nfacets)]
        self.base_rot = np.stack([np.asarray(configs._dc_facet_matrix(x, y).rot) for x, y in configs.dc_facet_positions()])
        self.comp = []      # (composite object, composite shape index, sphere shape index)
        for k in self.nodes:
            v = int(flat.nodes[k]["volume"])
            cs = int(flat.volumes[v]["shape"])
            sph = int(flat.params[flat.shapes[cs]["param_offset"] + 1])
            self.comp.append((flat.volume_objs[v].shape, cs, sph))
        self.radius = np.array([comp.node.left.rmin for comp, _, _ in self.comp])  # the spheres' current R

    def _put_shape(self, idx, sh, params):
        f = self.flat
        off, n = int(f.shapes[idx]["param_offset"]), int(f.shapes[idx]["n_params"])
        assert len(params) == n
        for j, v in enumerate(params):
            if v is not None:
                f.params[off + j] = v
        org, half = sh.local_bbox()
        f.shapes[idx]["bbox_origin"] = org
        f.shapes[idx]["bbox_half"] = half

    def update(self, params):
        p = np.asarray(params, dtype=float)
        nf = self.nfacets
        f = self.flat
        f.nodes["rot"][self.nodes] = configs.facet_rotations(self.base_rot[:nf], p[:2 * nf]).reshape(nf, 9)
        for i in np.nonzero(p[2 * nf:3 * nf] != self.radius)[0]:  # the facets whose R changed
            comp, cs, sph = self.comp[i]
            s = comp.node.left
            R = float(p[2 * nf + i])
            s.rmin, s.rmax = R, R + configs.C2_MIRROR_T
            s.theta1 = 180.0 - math.asin(configs.C2_MIRROR_D / math.sqrt(3) / R) * 180.0 / math.pi
            comp.node.lmat.tr = np.array([0.0, 0.0, R])
            self._put_shape(sph, s, s.params())
            self._put_shape(cs, comp, comp.params())
            self.radius[i] = R
        cam = f.node_names.index("focalPlane_1")
        obs = f.node_names.index("cameraBox_1")
        zc = configs.C2_F + float(p[3 * nf])
        f.nodes[cam]["tr"] = [0.0, 0.0, zc + 1 * configs.mm]
        f.nodes[obs]["tr"] = [0.0, 0.0, zc + 2 * configs.mm + 1 * configs.cm + 5 * configs.cm]
        return f


def node_param_table(flat, nfacets=C5_NFACETS):
    """[n_nodes][3] int32: the (tip, tilt, R) parameter indices of each facet
    node, -1 elsewhere; plus the focal-plane node index."""
    t = np.full((len(flat.nodes), 3), -1, dtype=np.int32)
    index = {name: k for k, name in enumerate(flat.node_names)}
    for i in range(nfacets):
        k = index[f"mirror{i}_{i + 1}"]
        t[k] = (2 * i, 2 * i + 1, 2 * nfacets + i)
    return t, index["focalPlane_1"]


def adjoint_desc(node_param, values, n_params, z_param, focal_node, cx, cy, weight, mode=A.RB_ADJ_CENTROID):
    """rb_adjoint_desc over caller-owned buffers (host or device pointers, as
    the tensors are)."""
    idx = node_param.reshape(-1)
    if idx.numel() and (int(idx.max()) >= n_params or int(idx.min()) < -1):
        raise ValueError(f"adjoint_desc: a node parameter index lies outside [-1, {n_params})")
    if not (-1 <= z_param < n_params):
        raise ValueError(f"adjoint_desc: z_param {z_param} outside [-1, {n_params})")
    d = A.rb_adjoint_desc()
    d.mode = mode
    d.node_param = node_param.data_ptr()
    d.param_values = values.data_ptr()
    d.n_params = n_params
    d.z_param = z_param
    d.focal_node = focal_node
    d.centroid_x, d.centroid_y, d.weight = float(cx), float(cy), float(weight)
    return d


class DaviesCottonAlignment:
    """Spot-size loss and its adjoint gradient over a fixed set of rays.

    Rays are generated once on the device (they do not depend on the
    parameters) and kept pristine; each evaluation traces a copy."""

    def __init__(self, ctx=None, angles_deg=C5_ANGLES_DEG, rays_per_angle=C5_RAYS_PER_ANGLE, precision="fp64",
                 nfacets=C5_NFACETS):
        self.ctx = ctx if ctx is not None else Context(0)
        self.angles = tuple(angles_deg)
        self.n = int(rays_per_angle)
        self.nfacets = nfacets
        self.n_params = 3 * nfacets + 1
        self.z_param = 3 * nfacets
        self.mixed = precision == "mixed"
        dev = torch.device("cuda")
        self.dev = dev
        # all angles' rays in one buffer (angle k at [k n, (k + 1) n)), so one
        # adjoint launch covers them; self.src[k] are views of angle k
        self.all = self._alloc(len(self.angles) * self.n)
        self.src = []
        for k, a in enumerate(self.angles):
            t = {f: v[k * self.n:(k + 1) * self.n] for f, v in self.all.items()}
            self.ctx.generate(c5_raygen(a, k), make_rays_device(t), self.n)
            self.src.append(t)
        self.work = self._alloc()
        self.scene = None
        self.mgr = None
        self.c0 = None
        self.last_adjoint_ms = 0.0
        self.last_trace_ms = 0.0

    def _alloc(self, n=None):
        n = self.n if n is None else n
        t = {f: torch.empty(n, dtype=torch.float64, device=self.dev) for f in _FIELDS_F64}
        for f in _FIELDS_I32:
            t[f] = torch.empty(n, dtype=torch.int32, device=self.dev)
        return t

    def _opts(self):
        self.mgr.SetPrecision(self.mixed)
        return self.mgr.trace_opts(device_ptrs=1)

    def set_params(self, params):
        p = np.asarray(params, dtype=float)
        assert p.shape == (self.n_params,)
        self.params = p.copy()
        if self.mgr is None:
            self.mgr = c5_scene(p, self.nfacets)
            self.updater = FacetSceneUpdater(self.mgr, self.nfacets)
            table, self.focal_node = node_param_table(self.mgr.flat(), self.nfacets)
            self.node_param = torch.from_numpy(table).to(self.dev)
        flat = self.updater.update(p)
        self.scene = self.ctx.upload(flat)
        self.values = torch.from_numpy(p).to(self.dev)

    def loss(self, params):
        """Sum over angles of the spot RMS^2 [cm^2]; keeps the per-angle
        centroids and counts for the adjoint."""
        self.set_params(params)
        opts = self._opts()
        self.stats = []
        total = 0.0
        self.last_trace_ms = 0.0
        for src in self.src:
            for f, v in src.items():
                self.work[f].copy_(v)
            self.ctx.trace(self.scene, make_rays_device(self.work), self.n, opts)
            self.last_trace_ms += self.ctx.last_trace_ms()
            m = (self.work["status"] == A.RB_FOCUS) & (self.work["last_node"] == self.focal_node)
            cnt = int(m.sum())
            if cnt == 0:
                self.stats.append((0, 0.0, 0.0))
                continue
            hx, hy = self.work["x"][m], self.work["y"][m]
            cx, cy = float(hx.mean()), float(hy.mean())
            rms2 = float(((hx - cx) ** 2 + (hy - cy) ** 2).mean())
            self.stats.append((cnt, cx, cy))
            total += rms2
        return total

    def loss_and_gradient(self, params):
        """Loss and gradient from one adjoint launch over all angles (no
        forward trace): per angle, the kernel accumulates the hit moments and
        sum (h - c0).J, sum J about a reference point c0 (the previous
        evaluation's centroid, so the FP32 products stay small), and the
        centroid terms are formed here in FP64."""
        if self.c0 is None:
            self.loss(params)
            self.c0 = [(cx, cy) for _, cx, cy in self.stats]
        self.set_params(params)
        opts = self._opts()
        P, G = self.n_params, len(self.src)
        acc = torch.zeros((G, 3 * P + 4), dtype=torch.float64, device=self.dev)
        ref = torch.tensor(self.c0, dtype=torch.float64, device=self.dev)
        d = adjoint_desc(self.node_param, self.values, P, self.z_param, self.focal_node, 0.0, 0.0, 0.0,
                         A.RB_ADJ_MOMENTS)
        d.ref_xy = ref.data_ptr()
        d.group_size = self.n
        d.n_groups = G
        self.ctx.adjoint(self.scene, make_rays_device(self.all), G * self.n, opts, d, acc.data_ptr())
        self.last_adjoint_ms = self.ctx.last_trace_ms()
        a = acc.cpu().numpy()
        loss, grad, stats = 0.0, np.zeros(P), []
        for k in range(len(self.src)):
            n, sx, sy, s2 = a[k, 3 * P:]
            if n == 0:
                stats.append((0, 0.0, 0.0))
                continue
            ux, uy = sx / n, sy / n
            loss += s2 / n - ux * ux - uy * uy
            grad += 2.0 / n * (a[k, :P] - ux * a[k, P:2 * P] - uy * a[k, 2 * P:3 * P])
            stats.append((int(n), self.c0[k][0] + ux, self.c0[k][1] + uy))
        self.stats = stats
        self.c0 = [(cx, cy) if cnt else c for (cnt, cx, cy), c in zip(stats, self.c0)]
        return loss, grad

    def gradient(self):
        """dL/dparams for the parameters of the last loss() call."""
        opts = self._opts()
        grad = torch.zeros(self.n_params, dtype=torch.float64, device=self.dev)
        self.last_adjoint_ms = 0.0
        for src, (cnt, cx, cy) in zip(self.src, self.stats):
            if cnt == 0:
                continue
            d = adjoint_desc(self.node_param, self.values, self.n_params, self.z_param, self.focal_node, cx, cy,
                             2.0 / cnt)
            self.ctx.adjoint(self.scene, make_rays_device(src), self.n, opts, d, grad.data_ptr())
            self.last_adjoint_ms += self.ctx.last_trace_ms()
        return grad.cpu().numpy()


class SpotLoss(torch.autograd.Function):
    """loss = SpotLoss.apply(params, model): params a float64 tensor of the
    model's parameter layout; backward runs the adjoint kernel."""

    @staticmethod
    def forward(ctx, params, model):
        loss, grad = model.loss_and_gradient(params.detach().cpu().numpy())
        ctx.grad = torch.from_numpy(grad).to(params.device)
        return torch.tensor(loss, dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(ctx, gout):
        return gout * ctx.grad, None


def adam_fit(model, p0, steps=200, lr=0.05, scale=None, callback=None):
    """Adam on the preconditioned coordinates u = (p - p0) / scale.
    Returns (final params, loss history)."""
    scale = c5_param_scale(model.nfacets) if scale is None else np.asarray(scale, dtype=float)
    p0 = torch.from_numpy(np.asarray(p0, dtype=float).copy())
    s = torch.from_numpy(scale)
    u = torch.zeros_like(p0, requires_grad=True)
    opt = torch.optim.Adam([u], lr=lr)
    history = []
    for it in range(steps):
        opt.zero_grad()
        loss = SpotLoss.apply(p0 + s * u, model)
        loss.backward()
        opt.step()
        history.append(float(loss.detach()))
        if callback is not None:
            callback(it, float(loss))
    return (p0 + s * u.detach()).numpy(), history
