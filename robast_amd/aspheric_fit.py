"""Differentiable figure fit of aspheric mirrors: the Schwarzschild-Couder
telescope (primary and secondary AGeoAsphericDisk mirrors, curved focal
surface).

The reference fits optics by re-tracing the geometry for every trial
parameter set under Minuit, with no gradient (tutorials/Optimize.C:76-129).
Here the figure gets an exact gradient from the `rbgpu_adjoint` kernel,
which recomputes each focused ray's two reflections on the disk surfaces and
back-propagates the loss through them and through the crossing with the
curved focal surface.

Parameter vector (21 entries): the primary's curvature c (1/cm), kappa
(= conic + 1, so d/dkappa is d/dconic) and K_1..K_8 (the r^2k coefficients,
cm^(1-2k)) -- both surfaces of the 1 um thick disk carry the same values, as
AGeoAsphericDisk::SetConicConstants / SetPolynomials set them
(src/AGeoAsphericDisk.cxx:1007-1013, :1187-1205) -- then the secondary's ten,
then the focal surface's z offset (cm).

Losses, per incidence angle k with N_k focused rays and hits h:
  "spot": the spot RMS^2 about the centroid c_k, (1/N_k) sum |h - c_k|^2;
  "ee":   minus the smooth encircled-energy fraction within r0 of a fixed
          point t_k (the first evaluation's centroid),
          -(1/N_k) sum sigmoid((r0^2 - |h - t_k|^2) / (2 r0 w)),
summed over the angles. `FigureLoss` wraps either as a
torch.autograd.Function; the spot loss then takes one adjoint launch per
evaluation (the moments form), which also returns the loss.
"""
import math

import numpy as np
import torch

from . import _abi as A
from . import configs
from ._lib import Context, make_rays_device

NPOL = configs.C3_NPOL
PER_MIRROR = 2 + NPOL
NPARAM = 2 * PER_MIRROR + 1
Z_PARAM = 2 * PER_MIRROR
FIT_ANGLES_DEG = (0.0, 1.5, 3.0, 4.5)
FIT_LAMBDA = 400 * configs.nm

_FIELDS_F64 = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
_FIELDS_I32 = ("status", "npoints", "last_node")


def figure_params(seed=0xC3):
    """The seeded C3 figure and focal offset 0 as a parameter vector."""
    (cp, kp, polp), (cs, ks, pols) = configs.c3_figure(seed)
    return np.array([cp, kp, *polp, cs, ks, *pols, 0.0])


def param_scale(params):
    """Natural size of each parameter (1e-3 of the curvature, 1e-2 of kappa,
    the seeded draws' 1e-3 / R^(2k-1) for K_k, 1 cm), used to precondition
    the optimiser and to size finite-difference steps."""
    p = np.asarray(params, dtype=float)
    s = np.empty(NPARAM)
    for m, rmax in enumerate((configs.C3_RP_MAX, configs.C3_RS_MAX)):
        o = m * PER_MIRROR
        s[o] = 1e-3 * abs(p[o])
        s[o + 1] = 1e-2
        s[o + 2:o + PER_MIRROR] = [1e-3 / rmax ** (2 * k - 1) for k in range(1, NPOL + 1)]
    s[Z_PARAM] = 1.0
    return s


def figure_scene(params):
    p = np.asarray(params, dtype=float)
    fig = tuple((float(p[o]), float(p[o + 1]), [float(v) for v in p[o + 2:o + PER_MIRROR]])
                for o in (0, PER_MIRROR))
    return configs.c3_scene(figure=fig, focal_dz=float(p[Z_PARAM]))


class FigureSceneUpdater:
    """Moves an existing Schwarzschild-Couder scene to new figure parameters
    by patching its flattened arrays in place: both mirror disks' parameters
    and bounding boxes, through the same shape code a rebuild uses, and the
    focal surface's z offset, so the arrays equal a rebuild's bit for bit."""

    def __init__(self, mgr):
        self.flat = flat = mgr.flat()
        index = {name: k for k, name in enumerate(flat.node_names)}
        self.mirrors = []       # (AGeoAsphericDisk, its shape index)
        for name in ("primaryMirror_1", "secondaryMirror_1"):
            v = int(flat.nodes[index[name]]["volume"])
            self.mirrors.append((flat.volume_objs[v].shape, int(flat.volumes[v]["shape"])))
        self.focal = index["focalPlane_1"]

    def update(self, params):
        p = np.asarray(params, dtype=float)
        f = self.flat
        for m, (sh, idx) in enumerate(self.mirrors):
            o = m * PER_MIRROR
            c, conic, pol = float(p[o]), float(p[o + 1]) - 1.0, [float(v) for v in p[o + 2:o + PER_MIRROR]]
            sh.c1 = sh.c2 = c
            sh.SetConicConstants(conic, conic)
            sh.SetPolynomials(len(pol), pol, len(pol), pol)
            off, n = int(f.shapes[idx]["param_offset"]), int(f.shapes[idx]["n_params"])
            vals = sh.params()
            assert len(vals) == n
            f.params[off:off + n] = vals
            org, half = sh.local_bbox()
            f.shapes[idx]["bbox_origin"] = org
            f.shapes[idx]["bbox_half"] = half
        f.nodes[self.focal]["tr"] = [0.0, 0.0, float(p[Z_PARAM])]
        return f


def surface_table(flat):
    """[n_nodes][2] int32 (first, count) of each mirror's parameter block, -1
    elsewhere; plus the focal node index."""
    t = np.full((len(flat.nodes), 2), -1, dtype=np.int32)
    index = {name: k for k, name in enumerate(flat.node_names)}
    t[index["primaryMirror_1"]] = (0, PER_MIRROR)
    t[index["secondaryMirror_1"]] = (PER_MIRROR, PER_MIRROR)
    return t, index["focalPlane_1"]


def raygen(angle_deg, index):
    """Parallel beam at `angle_deg` off axis over the 10 m disk, 400 nm."""
    g = configs.c3_raygen(seed=0xC3_00 + index)
    g.theta_min = g.theta_max = angle_deg * math.pi / 180.0
    g.lambda_min = g.lambda_max = FIT_LAMBDA
    return g


def figure_desc(surface_param, values, focal_node, mode, cx, cy, weight, ee_radius=0.0, ee_width=0.0):
    """rb_adjoint_desc for the figure parameters over caller-owned buffers."""
    d = A.rb_adjoint_desc()
    d.mode = mode
    d.node_param = None
    d.surface_param = surface_param.data_ptr() if hasattr(surface_param, "data_ptr") else surface_param.ctypes.data
    d.param_values = values.data_ptr() if hasattr(values, "data_ptr") else values.ctypes.data
    d.n_params = NPARAM
    d.z_param = Z_PARAM
    d.focal_node = focal_node
    d.centroid_x, d.centroid_y, d.weight = float(cx), float(cy), float(weight)
    d.ee_radius, d.ee_width = float(ee_radius), float(ee_width)
    return d


class AsphericFigureFit:
    """Figure loss and its adjoint gradient over a fixed set of rays, traced
    on the GPU. loss="spot" or "ee" (ee_radius r0 and ee_width w in cm)."""

    def __init__(self, ctx=None, angles_deg=FIT_ANGLES_DEG, rays_per_angle=200_000, loss="spot", ee_radius=1.0,
                 ee_width=0.2):
        if loss not in ("spot", "ee"):
            raise ValueError(f"AsphericFigureFit: unknown loss {loss!r}")
        self.ctx = ctx if ctx is not None else Context(0)
        self.angles = tuple(angles_deg)
        self.n = int(rays_per_angle)
        self.kind = loss
        self.ee_radius, self.ee_width = float(ee_radius), float(ee_width)
        self.dev = torch.device("cuda")
        # all angles' rays in one buffer (angle k at [k n, (k + 1) n)), so one
        # adjoint launch can cover them; self.src[k] are views of angle k
        self.all = self._alloc(len(self.angles) * self.n)
        self.src = []
        for k, a in enumerate(self.angles):
            t = {f: v[k * self.n:(k + 1) * self.n] for f, v in self.all.items()}
            self.ctx.generate(raygen(a, k), make_rays_device(t), self.n)
            self.src.append(t)
        self.work = self._alloc()
        self.updater = None
        self.targets = None
        self.table = None
        self.c0 = None
        self.last_adjoint_ms = 0.0

    def _alloc(self, n=None):
        n = self.n if n is None else n
        t = {f: torch.empty(n, dtype=torch.float64, device=self.dev) for f in _FIELDS_F64}
        for f in _FIELDS_I32:
            t[f] = torch.empty(n, dtype=torch.int32, device=self.dev)
        return t

    def set_params(self, params):
        p = np.asarray(params, dtype=float)
        if p.shape != (NPARAM,):
            raise ValueError(f"AsphericFigureFit: {NPARAM} parameters expected, got {p.shape}")
        self.params = p.copy()
        if self.updater is None:
            self.mgr = figure_scene(p)
            self.updater = FigureSceneUpdater(self.mgr)
            flat = self.mgr.flat()
        else:
            flat = self.updater.update(p)
        if self.table is None:
            table, self.focal_node = surface_table(flat)
            self.table = torch.from_numpy(table).to(self.dev)
        self.scene = self.ctx.upload(flat)
        self.values = torch.from_numpy(p).to(self.dev)

    def _hits(self, src):
        for f, v in src.items():
            self.work[f].copy_(v)
        self.ctx.trace(self.scene, make_rays_device(self.work), self.n, self.mgr.trace_opts(device_ptrs=1))
        m = (self.work["status"] == A.RB_FOCUS) & (self.work["last_node"] == self.focal_node)
        return self.work["x"][m], self.work["y"][m]

    def loss(self, params):
        """The loss over all angles; keeps the per-angle counts and centroids
        (and, on the first call, the encircled-energy target points)."""
        self.set_params(params)
        hits = []
        self.stats = []
        for src in self.src:
            hx, hy = self._hits(src)
            hx, hy = hx.clone(), hy.clone()
            cnt = int(hx.numel())
            self.stats.append((cnt, float(hx.mean()) if cnt else 0.0, float(hy.mean()) if cnt else 0.0))
            hits.append((hx, hy))
        if self.kind == "ee" and self.targets is None:
            self.targets = [(cx, cy) for _, cx, cy in self.stats]
        total = 0.0
        for k, ((hx, hy), (cnt, cx, cy)) in enumerate(zip(hits, self.stats)):
            if cnt == 0:
                continue
            if self.kind == "spot":
                total += float(((hx - cx) ** 2 + (hy - cy) ** 2).mean())
            else:
                tx, ty = self.targets[k]
                r0, w = self.ee_radius, self.ee_width
                u = (r0 * r0 - (hx - tx) ** 2 - (hy - ty) ** 2) / (2 * r0 * w)
                total -= float(torch.sigmoid(u).mean())
        return total

    def loss_and_gradient(self, params):
        """Spot loss and gradient from one adjoint launch over all angles (no
        forward trace): the moments form of rbgpu_adjoint, one group of sums
        per angle about the previous evaluation's centroid, the centroid
        terms formed here in FP64. The "ee" loss takes loss() and gradient()."""
        if self.kind != "spot":
            return self.loss(params), self.gradient()
        if self.c0 is None:
            self.loss(params)
            self.c0 = [(cx, cy) for _, cx, cy in self.stats]
        self.set_params(params)
        P, G = NPARAM, len(self.src)
        acc = torch.zeros((G, 3 * P + 4), dtype=torch.float64, device=self.dev)
        ref = torch.tensor(self.c0, dtype=torch.float64, device=self.dev)
        d = figure_desc(self.table, self.values, self.focal_node, A.RB_ADJ_MOMENTS, 0.0, 0.0, 0.0)
        d.ref_xy = ref.data_ptr()
        d.group_size = self.n
        d.n_groups = G
        self.ctx.adjoint(self.scene, make_rays_device(self.all), G * self.n, self.mgr.trace_opts(device_ptrs=1), d,
                         acc.data_ptr())
        self.last_adjoint_ms = self.ctx.last_trace_ms()
        a = acc.cpu().numpy()
        loss, grad, stats = 0.0, np.zeros(P), []
        for k in range(G):
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
        """dL/dparams at the parameters of the last loss() call."""
        opts = self.mgr.trace_opts(device_ptrs=1)
        P = NPARAM
        acc = torch.zeros(P + 1, dtype=torch.float64, device=self.dev)
        self.last_adjoint_ms = 0.0
        for k, (src, (cnt, cx, cy)) in enumerate(zip(self.src, self.stats)):
            if cnt == 0:
                continue
            if self.kind == "spot":
                d = figure_desc(self.table, self.values, self.focal_node, A.RB_ADJ_CENTROID, cx, cy, 2.0 / cnt)
            else:
                tx, ty = self.targets[k]
                d = figure_desc(self.table, self.values, self.focal_node, A.RB_ADJ_ENCIRCLED, tx, ty, -1.0 / cnt,
                                self.ee_radius, self.ee_width)
            self.ctx.adjoint(self.scene, make_rays_device(src), self.n, opts, d, acc.data_ptr())
            self.last_adjoint_ms += self.ctx.last_trace_ms()
        self.last_adjoint_loss = float(acc[P]) if self.kind == "ee" else None
        return acc[:P].cpu().numpy()


class FigureLoss(torch.autograd.Function):
    """loss = FigureLoss.apply(params, model): params a float64 tensor of the
    figure layout; backward hands on the adjoint gradient."""

    @staticmethod
    def forward(ctx, params, model):
        loss, grad = model.loss_and_gradient(params.detach().cpu().numpy())
        ctx.grad = torch.from_numpy(grad).to(params.device)
        return torch.tensor(loss, dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(ctx, gout):
        return gout * ctx.grad, None


def adam_fit(model, p0, steps=100, lr=0.05, scale=None, mask=None, callback=None):
    """Adam on the preconditioned coordinates u = (p - p0) / scale; `mask`
    (bool per parameter) freezes the entries it leaves False. Returns
    (final params, loss history)."""
    p0n = np.asarray(p0, dtype=float)
    scale = param_scale(p0n) if scale is None else np.asarray(scale, dtype=float)
    if mask is not None:
        scale = np.where(np.asarray(mask, dtype=bool), scale, 0.0)
    p0t = torch.from_numpy(p0n.copy())
    s = torch.from_numpy(scale)
    u = torch.zeros_like(p0t, requires_grad=True)
    opt = torch.optim.Adam([u], lr=lr)
    history = []
    for it in range(steps):
        opt.zero_grad()
        loss = FigureLoss.apply(p0t + s * u, model)
        loss.backward()
        opt.step()
        history.append(float(loss.detach()))
        if callback is not None:
            callback(it, float(loss))
    return (p0t + s * u.detach()).numpy(), history
