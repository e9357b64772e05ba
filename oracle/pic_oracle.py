"""FP64 NumPy oracle for the PyPIC3D per-step PIC hot path (one tile = one device).

TEST INFRASTRUCTURE ONLY -- the checker, never the product (see ``oracle/__init__``).

Layout: the single-tile layout of the reference, without the leading
``(ntx, nty, ntz) = (1, 1, 1)`` tile axes.  A field component is an
``(Nx+2g, Ny+2g, Nz+2g)`` float64 array; particle state is ``x, u`` of shape
``(S, P, 3)`` and ``active`` of shape ``(S, P)``.

Parity pins (checks that do not retype the formula; tests/test_oracle_*.py):

* grids / anchors / offsets: hand-computed coordinates, tile-local index map
  l <-> physical index l-g (grids.py:80-111), half-to-even rounding for TSC.
* shapes: partition of unity, exact values at delta=0 (shapes.py:5-54).
* gather: constant and linear fields are reproduced exactly (boris_test.py:69-129),
  reduced-axis gather equals the 1-D gather (boris_test.py:92-129).
* Boris: pure-B gyration preserves |v| and rotates by 2*atan(|t|); pure-E adds
  q E dt / m; relativistic Boris against the exact gyration of
  pusher_schmitz_test.py:69-84 and the boosted crossed-field case :160-183.
* Esirkepov: discrete continuity d(rho)/dt + div J = 0 to 1e-12 in 1-D, 2-D and
  3-D for both shape orders (extends esirkepov_test.py:700-744 to every axis),
  and a brute-force pure-Python single-particle restatement.
* J_from_rhov / rho: closed-form single-particle stencils
  (single_particle_pipeline_test.py:229-320), total charge = q*w.
* ghost refresh/fold: explicit slab bookkeeping on tiny arrays
  (distributed_ghost_cells_test.py:74-221 style), fold conserves the sum.
* Yee: E = -dt*J/eps from zero fields (esirkepov_test.py:510-529: J=4, eps=2,
  dt=0.25 -> E=-0.5), B_y from E_z = sin(x) converges at second order with the
  half-step factor dt/2 that first_order_yee.py:116 applies.
* CG: single Fourier mode sin(x+y+z) (electrostatic_yee_test.py:92-119) and a
  scipy.sparse direct solve of the same 7-point system.
* centred gradient: exact cos(theta) sin(d)/d of a single mode per axis, and
  the half difference on conducting wall planes of a linear phi; the rolled
  gradient of calculate_electrostatic_fields wraps the interior instead
  (a(n-2)/2 on the wall planes of a linear phi).
* zero / constant boundaries and the phi boundary sequence: the known answers
  of ghost_cells_test.py:97-160 and electrostatic_yee_test.py:150-222, plus an
  element-wise statement of the result with two conducting axes (only the
  last conducting axis keeps copied ghosts).
* J_from_rhov face weights: the second (CIC) and third (TSC) central moments
  of the face stencil equal those of the extrapolating weights from the node
  anchor (J_from_rhov.py:139-141), not of a bracketing stencil.
* Esirkepov reduced-axis currents: totals and the moments x y (2-D null axis,
  the 1/3-1/6 time average of x(t) y(t)) and x (1-D transverse, the path
  midpoint).
* mode="drop": TSC deposits with g = 1 whose stencil passes the upper end
  (dropped) or index -1 (wrapped to L-1).
* sort: ``np.argsort(kind="stable")`` is the specification itself
  (particle_initialization.py:159); checked against a brute-force bucket sort.
* energy, momentum, charge: closed forms on uniform fields / single
  particles; the charge of a deposit equals sum q w.
* PML: profile values at hand-computed indices and the validation errors of
  PML.py:17-81; the ADE memory under a constant derivative follows the closed
  form m_n = (b^n - 1) d, so the stretched derivative decays as b^n d;
  sigma = 0 leaves the Yee updates unchanged; a 1-D pulse between PML walls
  loses most of its energy (pml_test.py:527-617 setup) while conducting walls
  alone keep it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Tuple

import numpy as np

BC_PERIODIC = 0      # grid_and_stencil.py:10
BC_CONDUCTING = 1    # grid_and_stencil.py:11
PBC_PERIODIC, PBC_REFLECTING, PBC_ABSORBING = 0, 1, 2   # initialization.py:55-66


# --------------------------------------------------------------------------
# parameters and grids
# --------------------------------------------------------------------------

@dataclass
class Params:
    """The subset of StaticParameters/DynamicParameters the hot path reads
    (parameters.py:14-53), for one tile covering the whole domain."""

    N: Tuple[int, int, int] = (8, 6, 4)
    wind: Tuple[float, float, float] = (4.0, 3.0, 2.0)
    d: Tuple[float, float, float] | None = None
    g: int = 2
    dt: float = 0.05
    C: float = 1.0
    eps: float = 1.0
    mu: float = 1.0
    alpha: float = 1.0
    shape_factor: int = 1
    relativistic: bool = True
    particle_pusher: str = "boris"
    current_deposition: str = "esirkepov"
    current_filter: str = "none"
    bc: Tuple[int, int, int] = (0, 0, 0)
    particle_bc: Tuple[int, int, int] = (0, 0, 0)
    electrostatic: bool = False

    def __post_init__(self):
        if self.d is None:
            self.d = tuple(float(w) / int(n) for w, n in zip(self.wind, self.N))
        self.N = tuple(int(n) for n in self.N)
        self.wind = tuple(float(w) for w in self.wind)
        self.d = tuple(float(v) for v in self.d)

    @property
    def shape(self):
        return tuple(n + 2 * self.g for n in self.N)

    @property
    def reduced(self):
        """ghost_cells.py:42-49 with one tile: an axis of one cell is reduced."""
        return tuple(n == 1 for n in self.N)

    @property
    def active_axes(self):
        """particle_push.py:38-42 / Esirkepov.py:92-94: axis active iff N > 1."""
        return tuple(n > 1 for n in self.N)


def tiled_axis(p: Params, axis: int, kind: str) -> np.ndarray:
    """Tile-local coordinate line of one tile covering the domain.

    build_collocated_axis / build_staggered_axis (grid_and_stencil.py:242-286)
    give the global line's first entry (``linspace`` returns ``start`` at index
    0); _tile_grid_axis (grids.py:98-104) then places local index l at
    ``global[0] + (l - (g-1)) * d``.  In the electrodynamic solver "center" is
    collocated and "vertex" staggered (grids.py:42-77); the electrostatic solver
    uses collocated lines for both (initialization.py:~310, grids.py:14-39).
    """
    w, d, n, g = p.wind[axis], p.d[axis], p.N[axis], p.g
    minimum = -w / 2
    if kind == "center" or p.electrostatic:
        start = minimum - d
    elif kind == "vertex":
        start = minimum - 0.5 * d
    else:
        raise ValueError(kind)
    offsets = np.arange(n + 2 * g, dtype=np.float64)
    return start + (offsets + 0.0 * n - (g - 1)) * d


def axis_origin_spacing(p: Params, axis: int, kind: str):
    line = tiled_axis(p, axis, kind)
    return float(line[0]), float(line[1] - line[0])


# --------------------------------------------------------------------------
# stencil primitives (grid_and_stencil.py, shapes.py)
# --------------------------------------------------------------------------

def compute_anchor(pos, origin, spacing, shape_factor):
    """grid_and_stencil.py:92-118: floor for CIC, round-half-even for TSC."""
    t = (np.asarray(pos, dtype=np.float64) - origin) / spacing
    if shape_factor == 1:
        return np.floor(t).astype(np.int64)
    return np.round(t).astype(np.int64)


def axis_offset(pos, anchor, origin, spacing):
    """grid_and_stencil.py:121-138."""
    return np.asarray(pos, dtype=np.float64) - (anchor * spacing + origin)


def shape_weights(delta, d, shape_factor):
    """shapes.py:5-54; returns weights for stencil points (-1, 0, +1)."""
    r = np.asarray(delta, dtype=np.float64) / d
    if shape_factor == 1:
        return np.stack([np.zeros_like(r), 1 - r, r])
    return np.stack([0.5 * (0.5 - r) ** 2, 0.75 - r ** 2, 0.5 * (0.5 + r) ** 2])


def scatter_add_drop(arr, ix, iy, iz, vals):
    """``arr.at[ix, iy, iz].add(vals, mode="drop")``: negative indices are first
    normalised NumPy-style (index + size), then anything still outside
    ``[0, size)`` is dropped (Esirkepov.py:238, J_from_rhov.py:187-198, rho.py:147)."""
    ix = np.asarray(ix, dtype=np.int64)
    iy = np.asarray(iy, dtype=np.int64)
    iz = np.asarray(iz, dtype=np.int64)
    vals = np.broadcast_to(np.asarray(vals, dtype=np.float64), np.broadcast(ix, iy, iz).shape)
    ix, iy, iz = np.broadcast_arrays(ix, iy, iz)
    sx, sy, sz = arr.shape
    ix = np.where(ix < 0, ix + sx, ix)
    iy = np.where(iy < 0, iy + sy, iy)
    iz = np.where(iz < 0, iz + sz, iz)
    ok = (ix >= 0) & (ix < sx) & (iy >= 0) & (iy < sy) & (iz >= 0) & (iz < sz)
    np.add.at(arr, (ix[ok], iy[ok], iz[ok]), vals[ok])


# --------------------------------------------------------------------------
# ghost cells, one device (ghost_cells.py)
# --------------------------------------------------------------------------

def _sl(axis, s):
    out = [slice(None)] * 3
    out[axis] = s
    return tuple(out)


def refresh(field, p: Params, bcs=None):
    """update_tiled_ghost_cells for one tile (ghost_cells.py:142-215, 615-650).

    Per axis x -> y -> z: a reduced axis copies interior[g] into every ghost
    (periodic) or zeroes them; a periodic axis self-exchanges
    (lower ghost <- upper interior, upper ghost <- lower interior, the
    ((0,0),) permutation of :72-76); a non-periodic axis of one device receives
    zeros from the empty permutation.
    """
    f = np.array(field, dtype=np.float64, copy=True)
    g = p.g
    bcs = p.bc if bcs is None else bcs
    for axis in range(3):
        bc = bcs[axis]
        if p.reduced[axis]:
            interior = f[_sl(axis, slice(g, g + 1))]
            if bc == BC_PERIODIC:
                f[_sl(axis, slice(0, g))] = interior
                f[_sl(axis, slice(-g, None))] = interior
            else:
                f[_sl(axis, slice(0, g))] = 0.0
                f[_sl(axis, slice(-g, None))] = 0.0
        else:
            if bc == BC_PERIODIC:
                lower_vals = f[_sl(axis, slice(-2 * g, -g))].copy()
                upper_vals = f[_sl(axis, slice(g, 2 * g))].copy()
            else:
                lower_vals = 0.0
                upper_vals = 0.0
            f[_sl(axis, slice(0, g))] = lower_vals
            f[_sl(axis, slice(-g, None))] = upper_vals
    return f


def fold(field, p: Params, bcs=None):
    """fold_tiled_ghost_cells for one tile (ghost_cells.py:218-316, 703-736).

    Per axis x -> y -> z: periodic adds the lower ghost slab into the upper
    interior slab and the upper ghost into the lower interior; code 1
    (conducting / reflecting) subtracts each ghost slab from the adjacent
    interior slab (:238-260); a reduced axis sums all ghosts into interior[g]
    (sign - for code 1, dropped for other non-periodic codes).  Ghosts -> 0.
    """
    f = np.array(field, dtype=np.float64, copy=True)
    g = p.g
    bcs = p.bc if bcs is None else bcs
    for axis in range(3):
        bc = bcs[axis]
        lo, hi = _sl(axis, slice(0, g)), _sl(axis, slice(-g, None))
        if p.reduced[axis]:
            ghost_sum = f[lo].sum(axis=axis, keepdims=True) + f[hi].sum(axis=axis, keepdims=True)
            if bc == BC_PERIODIC:
                f[_sl(axis, slice(g, g + 1))] += ghost_sum
            elif bc == BC_CONDUCTING:
                f[_sl(axis, slice(g, g + 1))] += -ghost_sum
        else:
            lower_vals = f[lo].copy()
            upper_vals = f[hi].copy()
            if bc == BC_PERIODIC:
                f[_sl(axis, slice(-2 * g, -g))] += lower_vals
                f[_sl(axis, slice(g, 2 * g))] += upper_vals
            if bc == BC_CONDUCTING:
                f[_sl(axis, slice(g, 2 * g))] += -lower_vals
                f[_sl(axis, slice(-2 * g, -g))] += -upper_vals
        f[lo] = 0.0
        f[hi] = 0.0
    return f


def zero_boundary(field, p: Params, axis):
    """apply_tiled_zero_boundary (ghost_cells.py:344-362, 653-672)."""
    if p.bc[axis] != BC_CONDUCTING:
        return refresh(field, p)
    f = np.array(field, dtype=np.float64, copy=True)
    f[_sl(axis, p.g)] = 0.0
    f[_sl(axis, -p.g - 1)] = 0.0
    return refresh(f, p)


def constant_boundary(field, p: Params, axis):
    """apply_tiled_constant_boundary (ghost_cells.py:365-386, 675-700)."""
    f = refresh(field, p)
    if p.bc[axis] != BC_CONDUCTING:
        return f
    g = p.g
    f[_sl(axis, slice(0, g))] = f[_sl(axis, slice(g, g + 1))]
    f[_sl(axis, slice(-g, None))] = f[_sl(axis, slice(-g - 1, -g))]
    return f


def phi_boundaries(field, p: Params):
    """_apply_tiled_phi_constant_boundaries (electrostatic_yee.py:26-43)."""
    applied = False
    f = field
    for axis in range(3):
        if p.bc[axis] == BC_CONDUCTING:
            f = constant_boundary(f, p, axis)
            applied = True
    if not applied:
        f = refresh(f, p)
    return f


# --------------------------------------------------------------------------
# filters (filters.py)
# --------------------------------------------------------------------------

def _conv3_interior(field, kernel, g):
    """VALID 3x3x3 correlation of the stencil window g-1 .. L-g (filters.py:6-44,
    208-215), written back onto the interior g:-g."""
    f = np.array(field, dtype=np.float64, copy=True)
    X, Y, Z = f.shape
    out = np.zeros((X - 2 * g, Y - 2 * g, Z - 2 * g))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                w = kernel[a, b, c]
                if w == 0.0:
                    continue
                out += w * field[g - 1 + a:X - g - 1 + a, g - 1 + b:Y - g - 1 + b, g - 1 + c:Z - g - 1 + c]
    f[g:-g, g:-g, g:-g] = out
    return f


def digital_filter(field, alpha, g):
    """filters.py:259-291: centre alpha, six faces (1-alpha)/6."""
    k = np.zeros((3, 3, 3))
    k[1, 1, 1] = alpha
    nw = (1.0 - alpha) / 6.0
    for idx in ((0, 1, 1), (2, 1, 1), (1, 0, 1), (1, 2, 1), (1, 1, 0), (1, 1, 2)):
        k[idx] = nw
    return _conv3_interior(field, k, g)


def bilinear_filter(field, g):
    """filters.py:234-256: (1,2,1)^3 / 64."""
    k1 = np.array([1.0, 2.0, 1.0])
    k = k1[:, None, None] * k1[None, :, None] * k1[None, None, :] / 64.0
    return _conv3_interior(field, k, g)


# --------------------------------------------------------------------------
# field gather + Boris (pusher/boris.py, pusher/particle_push.py)
# --------------------------------------------------------------------------

def _axis_stencil(pos, line, shape_factor, active, g):
    """prepare_particle_axis_stencil with BC_PERIODIC + collapse (boris.py:338-409)."""
    origin, spacing = float(line[0]), float(line[1] - line[0])
    a = compute_anchor(pos, origin, spacing, shape_factor)
    delta = axis_offset(pos, a, origin, spacing)
    pts = np.mod(a[None, :] + np.array([-1, 0, 1])[:, None], len(line))
    w = shape_weights(delta, spacing, shape_factor)
    if not active:
        pts = np.full((1, pts.shape[1]), g, dtype=np.int64)
        w = np.sum(w, axis=0, keepdims=True)
    return pts, w


def interpolate(field, x, y, z, lines, p: Params):
    """interpolate_field_to_particles with ghost_cells=True, explicit active
    axes and inactive index g (boris.py:302-433; particle_push.py:38-94)."""
    xp, wx = _axis_stencil(x, lines[0], p.shape_factor, p.active_axes[0], p.g)
    yp, wy = _axis_stencil(y, lines[1], p.shape_factor, p.active_axes[1], p.g)
    zp, wz = _axis_stencil(z, lines[2], p.shape_factor, p.active_axes[2], p.g)
    out = np.zeros(np.asarray(x).shape, dtype=np.float64)
    for i in range(xp.shape[0]):
        for j in range(yp.shape[0]):
            for k in range(zp.shape[0]):
                out += field[xp[i], yp[j], zp[k]] * wx[i] * wy[j] * wz[k]
    return out


def component_lines(p: Params):
    """Yee staggering of the pushed components (particle_push.py:63-68)."""
    c = [tiled_axis(p, a, "center") for a in range(3)]
    v = [tiled_axis(p, a, "vertex") for a in range(3)]
    return {
        "Ex": (v[0], c[1], c[2]),
        "Ey": (c[0], v[1], c[2]),
        "Ez": (c[0], c[1], v[2]),
        "Bx": (c[0], v[1], v[2]),
        "By": (v[0], c[1], v[2]),
        "Bz": (v[0], v[1], c[2]),
    }


def boris(v, E, B, q, m, dt):
    """boris_single_particle (boris.py:14-56), vectorised; v, E, B: (3, N)."""
    h = q * dt / (2 * m)
    vminus = v + h * E
    t = h * B
    vprime = vminus + np.cross(vminus, t, axis=0)
    s = 2 * t / (1 + t[0] ** 2 + t[1] ** 2 + t[2] ** 2)
    vplus = vminus + np.cross(vprime, s, axis=0)
    return vplus + h * E


def relativistic_boris(v, E, B, q, m, dt, C):
    """relativistic_boris_single_particle (boris.py:60-124)."""
    gamma = 1 / np.sqrt(1 - ((v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / C ** 2))
    h = q * dt / (2 * m)
    uminus = v * gamma + h * E
    gamma_minus = np.sqrt(1 + ((uminus[0] ** 2 + uminus[1] ** 2 + uminus[2] ** 2) / C ** 2))
    t = h * B / gamma_minus
    uprime = uminus + np.cross(uminus, t, axis=0)
    s = 2 * t / (1 + t[0] ** 2 + t[1] ** 2 + t[2] ** 2)
    uplus = uminus + np.cross(uprime, s, axis=0)
    newu = uplus + h * E
    new_gamma = np.sqrt(1 + ((newu[0] ** 2 + newu[1] ** 2 + newu[2] ** 2) / C ** 2))
    return newu / new_gamma


def higuera_cary(v, E, B, q, m, dt, C):
    """higuera_cary_single_particle (higuera_cary.py:5-113), vectorised; (3, N)."""
    u = v / np.sqrt(1.0 - (v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / C ** 2)
    h = q * dt / (2.0 * m)
    eps = h * E
    beta = h * B
    ue = u + eps
    beta2 = beta[0] ** 2 + beta[1] ** 2 + beta[2] ** 2
    ustar = (ue[0] * beta[0] + ue[1] * beta[1] + ue[2] * beta[2]) / C
    sigma = (1.0 + (ue[0] ** 2 + ue[1] ** 2 + ue[2] ** 2) / C ** 2) - beta2
    gamma = np.sqrt((sigma + np.sqrt(sigma ** 2 + 4.0 * (beta2 + ustar ** 2))) / 2.0)
    t = beta / gamma
    s = 1.0 / (1.0 + t[0] ** 2 + t[1] ** 2 + t[2] ** 2)
    uet = ue[0] * t[0] + ue[1] * t[1] + ue[2] * t[2]
    um = s * (ue + uet * t + np.cross(ue, t, axis=0))
    newu = um + eps + np.cross(um, t, axis=0)
    return newu / np.sqrt(1.0 + (newu[0] ** 2 + newu[1] ** 2 + newu[2] ** 2) / C ** 2)


@dataclass
class Species:
    """SpeciesConfig (particle_class.py:5-14)."""
    charge: np.ndarray
    mass: np.ndarray
    weight: np.ndarray
    update_x: np.ndarray   # (S, 3) bool
    update_u: np.ndarray   # (S, 3) bool


def particle_push(x, u, active, sp: Species, E, B, p: Params):
    """particle_push (particle_push.py:13-175) for one tile; returns new u."""
    if p.particle_pusher not in ("boris", "higuera_cary"):
        raise ValueError(f"Unknown particle_pusher: {p.particle_pusher}")
    S, P, _ = x.shape
    lines = component_lines(p)
    xs, ys, zs = (x[..., a].reshape(-1) for a in range(3))
    Ef = np.stack([interpolate(E[0], xs, ys, zs, lines["Ex"], p),
                   interpolate(E[1], xs, ys, zs, lines["Ey"], p),
                   interpolate(E[2], xs, ys, zs, lines["Ez"], p)])
    Bf = np.stack([interpolate(B[0], xs, ys, zs, lines["Bx"], p),
                   interpolate(B[1], xs, ys, zs, lines["By"], p),
                   interpolate(B[2], xs, ys, zs, lines["Bz"], p)])
    v = np.stack([u[..., a].reshape(-1) for a in range(3)])
    q = np.repeat(np.asarray(sp.charge, dtype=np.float64), P)
    m = np.repeat(np.asarray(sp.mass, dtype=np.float64), P)
    if p.particle_pusher == "higuera_cary":        # particle_push.py:124-131
        newv = higuera_cary(v, Ef, Bf, q, m, p.dt, p.C)
    elif p.relativistic:
        newv = relativistic_boris(v, Ef, Bf, q, m, p.dt, p.C)
    else:
        newv = boris(v, Ef, Bf, q, m, p.dt)
    act = np.asarray(active, dtype=bool).reshape(-1)
    out = np.array(u, dtype=np.float64, copy=True)
    for a in range(3):
        upd = np.repeat(np.asarray(sp.update_u, dtype=bool)[:, a], P)
        out[..., a] = np.where(act & upd, newv[a], v[a]).reshape(S, P)
    return out


# --------------------------------------------------------------------------
# deposition (deposition/Esirkepov.py, J_from_rhov.py, rho.py)
# --------------------------------------------------------------------------

def _flat_species(x, u, active, sp: Species):
    S, P, _ = x.shape
    xs = [np.asarray(x[..., a], dtype=np.float64).reshape(-1) for a in range(3)]
    us = [np.asarray(u[..., a], dtype=np.float64).reshape(-1) for a in range(3)]
    act = np.asarray(active, dtype=bool).reshape(-1)
    qw = np.repeat(np.asarray(sp.charge, dtype=np.float64) * np.asarray(sp.weight, dtype=np.float64), P)
    upd_x = [np.repeat(np.asarray(sp.update_x, dtype=bool)[:, a], P) for a in range(3)]
    return xs, us, act, qw, upd_x


def _shift_old(w5, shift):
    """shift_old_stencil (Esirkepov.py:17-25): roll by -shift per particle."""
    idx = (np.arange(5)[:, None] + shift[None, :]) % 5
    return np.take_along_axis(w5, idx, axis=0)


def esirkepov_raw(x, u, active, sp: Species, p: Params):
    """Tile deposit of Esirkepov_current before the fold (Esirkepov.py:105-331)."""
    g = p.g
    dt = p.dt
    dx, dy, dz = p.d
    xs, us, act, q, upd_x = _flat_species(x, u, active, sp)
    act = act.astype(np.float64)
    lines = [tiled_axis(p, a, "center") for a in range(3)]
    shape = p.shape
    new = [xs[a] + np.where(upd_x[a], us[a] * dt, 0.0) for a in range(3)]   # :125-127
    pts, nw, ow = [], [], []
    for a in range(3):
        origin, spacing = float(lines[a][0]), float(lines[a][1] - lines[a][0])
        a_new = compute_anchor(new[a], origin, spacing, p.shape_factor)
        a_old = compute_anchor(xs[a], origin, spacing, p.shape_factor)
        d_new = axis_offset(new[a], a_new, origin, spacing)
        d_old = axis_offset(xs[a], a_old, origin, spacing)
        w_new = shape_weights(d_new, p.d[a], p.shape_factor)       # dynamic d (:164-175)
        w_old = shape_weights(d_old, p.d[a], p.shape_factor)
        zero = np.zeros((1, w_new.shape[1]))
        w_new5 = np.concatenate([zero, w_new, zero])
        w_old5 = _shift_old(np.concatenate([zero, w_old, zero]), a_new - a_old)
        pt = a_new[None, :] + np.arange(-2, 3)[:, None]
        if not p.active_axes[a]:                                      # :28-45
            pt = np.full_like(pt, shape[a] // 2)
            tot_n = w_new5.sum(axis=0)
            tot_o = w_old5.sum(axis=0)
            w_new5 = np.zeros_like(w_new5)
            w_old5 = np.zeros_like(w_old5)
            w_new5[2] = tot_n
            w_old5[2] = tot_o
        pts.append(pt)
        nw.append(w_new5)
        ow.append(w_old5)
    xa, ya, za = p.active_axes
    dV = dx * dy * dz
    dJ = [act * (-(q / (dy * dz)) / dt) if xa else act * q * us[0] / dV,
          act * (-(q / (dx * dz)) / dt) if ya else act * q * us[1] / dV,
          act * (-(q / (dx * dy)) / dt) if za else act * q * us[2] / dV]
    J = [np.zeros(shape) for _ in range(3)]
    xw, yw, zw = nw
    oxw, oyw, ozw = ow
    n_active = sum(p.active_axes)
    if n_active == 3:                                                 # :222-240
        Wx = (xw[:, None, None] - oxw[:, None, None]) * (
            1 / 3 * (yw[None, :, None] * zw[None, None, :] + oyw[None, :, None] * ozw[None, None, :])
            + 1 / 6 * (yw[None, :, None] * ozw[None, None, :] + oyw[None, :, None] * zw[None, None, :]))
        Wy = (yw[None, :, None] - oyw[None, :, None]) * (
            1 / 3 * (xw[:, None, None] * zw[None, None, :] + oxw[:, None, None] * ozw[None, None, :])
            + 1 / 6 * (xw[:, None, None] * ozw[None, None, :] + oxw[:, None, None] * zw[None, None, :]))
        Wz = (zw[None, None, :] - ozw[None, None, :]) * (
            1 / 3 * (xw[:, None, None] * yw[None, :, None] + oxw[:, None, None] * oyw[None, :, None])
            + 1 / 6 * (xw[:, None, None] * oyw[None, :, None] + oxw[:, None, None] * yw[None, :, None]))
        Jx_loc = np.cumsum(dJ[0] * Wx, axis=0)
        Jy_loc = np.cumsum(dJ[1] * Wy, axis=1)
        Jz_loc = np.cumsum(dJ[2] * Wz, axis=2)
        for i in range(5):
            for j in range(5):
                for k in range(5):
                    ix, iy, iz = pts[0][i], pts[1][j], pts[2][k]
                    scatter_add_drop(J[0], ix, iy, iz, Jx_loc[i, j, k])
                    scatter_add_drop(J[1], ix, iy, iz, Jy_loc[i, j, k])
                    scatter_add_drop(J[2], ix, iy, iz, Jz_loc[i, j, k])
    elif n_active == 2:                                               # :241-287, :438-504
        null = [a for a in range(3) if not p.active_axes[a]][0]
        a1, a2 = [a for a in range(3) if p.active_axes[a]]
        w1, w2, o1, o2 = nw[a1], nw[a2], ow[a1], ow[a2]
        # plane weights indexed [slot along a1, slot along a2]
        W_a1 = 0.5 * (w1[:, None] - o1[:, None]) * (w2[None, :] + o2[None, :])
        W_a2 = 0.5 * (w2[None, :] - o2[None, :]) * (w1[:, None] + o1[:, None])
        W_null = (1 / 3 * (w1[:, None] * w2[None, :] + o1[:, None] * o2[None, :])
                  + 1 / 6 * (w1[:, None] * o2[None, :] + o1[:, None] * w2[None, :]))
        W = [None, None, None]
        W[a1], W[a2], W[null] = W_a1, W_a2, W_null
        F = [dJ[c] * W[c] for c in range(3)]
        loc = [None, None, None]
        loc[a1] = np.cumsum(F[a1], axis=0)
        loc[a2] = np.cumsum(F[a2], axis=1)
        loc[null] = F[null]
        for i in range(5):
            for j in range(5):
                idx = [None, None, None]
                idx[a1] = pts[a1][i]
                idx[a2] = pts[a2][j]
                idx[null] = pts[null][2]
                for c in range(3):
                    scatter_add_drop(J[c], idx[0], idx[1], idx[2], loc[c][i, j])
    elif n_active == 1:                                               # :288-329, :412-435
        a0 = [a for a in range(3) if p.active_axes[a]][0]
        w, o = nw[a0], ow[a0]
        loc = [None, None, None]
        for c in range(3):
            if c == a0:
                loc[c] = np.cumsum(dJ[c] * (w - o), axis=0)
            else:
                loc[c] = dJ[c] * ((w + o) / 2)
        for i in range(5):
            idx = [pts[0][2], pts[1][2], pts[2][2]]
            idx[a0] = pts[a0][i]
            for c in range(3):
                scatter_add_drop(J[c], idx[0], idx[1], idx[2], loc[c][i])
    else:
        raise ValueError("Esirkepov needs at least one active axis")
    return J


def esirkepov_current(x, u, active, sp: Species, p: Params):
    """Esirkepov_current (Esirkepov.py:48-362): deposit, fold, refresh (bc_type=1)."""
    J = esirkepov_raw(x, u, active, sp, p)
    J = [fold(c, p, p.particle_bc) for c in J]
    return [refresh(c, p, p.particle_bc) for c in J]


def _node_face_stencils(xs, p: Params):
    """prepare_particle_axis_stencil with local_bc=2 plus face offsets
    (J_from_rhov.py:110-174, rho.py:88-136)."""
    g = p.g
    out = []
    for a in range(3):
        line = tiled_axis(p, a, "center")
        origin, spacing = float(line[0]), float(line[1] - line[0])
        anc = compute_anchor(xs[a], origin, spacing, p.shape_factor)
        d_node = axis_offset(xs[a], anc, origin, spacing)
        pts = anc[None, :] + np.array([-1, 0, 1])[:, None]       # local_bc=2: no wrap
        d_face = (xs[a] - line[0]) - (anc + 0.5) * p.d[a]
        w_node = shape_weights(d_node, p.d[a], p.shape_factor)
        w_face = shape_weights(d_face, p.d[a], p.shape_factor)
        if p.reduced[a]:
            pts = np.full((1, pts.shape[1]), g, dtype=np.int64)
            w_node = w_node.sum(axis=0, keepdims=True)
            w_face = w_face.sum(axis=0, keepdims=True)
        out.append((pts, w_node, w_face))
    return out


def j_from_rhov_raw(x, u, active, sp: Species, p: Params):
    xs, us, act, q, _ = _flat_species(x, u, active, sp)
    dq = act.astype(np.float64) * q / (p.d[0] * p.d[1] * p.d[2])
    st = _node_face_stencils(xs, p)
    J = [np.zeros(p.shape) for _ in range(3)]
    (xp, xn, xf), (yp, yn, yf), (zp, zn, zf) = st
    for i in range(xp.shape[0]):
        for j in range(yp.shape[0]):
            for k in range(zp.shape[0]):
                scatter_add_drop(J[0], xp[i], yp[j], zp[k], dq * us[0] * xf[i] * yn[j] * zn[k])
                scatter_add_drop(J[1], xp[i], yp[j], zp[k], dq * us[1] * xn[i] * yf[j] * zn[k])
                scatter_add_drop(J[2], xp[i], yp[j], zp[k], dq * us[2] * xn[i] * yn[j] * zf[k])
    return J


def j_from_rhov(x, u, active, sp: Species, p: Params):
    """J_from_rhov (J_from_rhov.py:31-260): deposit, fold, refresh, filter."""
    J = j_from_rhov_raw(x, u, active, sp, p)
    J = [refresh(fold(c, p, p.particle_bc), p, p.particle_bc) for c in J]
    if p.current_filter == "bilinear":
        J = [refresh(bilinear_filter(c, p.g), p, p.particle_bc) for c in J]
    elif p.current_filter == "digital":
        J = [refresh(digital_filter(c, p.alpha, p.g), p, p.particle_bc) for c in J]
    return J


def rho_raw(x, active, sp: Species, p: Params):
    S, P, _ = x.shape
    xs = [np.asarray(x[..., a], dtype=np.float64).reshape(-1) for a in range(3)]
    act = np.asarray(active, dtype=bool).reshape(-1).astype(np.float64)
    qd = np.repeat(np.asarray(sp.charge, dtype=np.float64) * np.asarray(sp.weight, dtype=np.float64)
                   / (p.d[0] * p.d[1] * p.d[2]), P)
    st = _node_face_stencils(xs, p)
    rho = np.zeros(p.shape)
    (xp, xn, _), (yp, yn, _), (zp, zn, _) = st
    for i in range(xp.shape[0]):
        for j in range(yp.shape[0]):
            for k in range(zp.shape[0]):
                scatter_add_drop(rho, xp[i], yp[j], zp[k], act * qd * xn[i] * yn[j] * zn[k])
    return rho


def compute_rho(x, active, sp: Species, p: Params):
    """compute_rho (rho.py:29-196)."""
    rho = refresh(fold(rho_raw(x, active, sp, p), p, p.particle_bc), p, p.particle_bc)
    if p.current_filter == "digital":
        rho = refresh(digital_filter(rho, p.alpha, p.g), p, p.particle_bc)
    return rho


# --------------------------------------------------------------------------
# Yee solver (solvers/first_order_yee.py)
# --------------------------------------------------------------------------

# --------------------------------------------------------------------------
# PML (boundary_conditions/PML.py)
# --------------------------------------------------------------------------

PML_WALLS = ("-x", "+x", "-y", "+y", "-z", "+z")


def pml_layers(raw, p: Params):
    """load_pml_from_toml's validation and sigma_max (PML.py:17-81): raw is a
    dict or a list of dicts {wall, thickness, order=3, target_reflection=1e-8,
    sigma_max (optional)}; returns [(wall, axis, thickness, order, sigma_max)]."""
    raw = [] if raw is None else ([raw] if isinstance(raw, dict) else list(raw))
    out, seen = [], set()
    for r in raw:
        wall = r.get("wall")
        if wall not in PML_WALLS:
            raise ValueError(f"Invalid PML wall: {wall}. Expected one of {list(PML_WALLS)}")
        if wall in seen:
            raise ValueError(f"Duplicate PML wall: {wall}")
        seen.add(wall)
        axis = "xyz".index(wall[1])
        thickness = int(r.get("thickness", 0))
        if thickness <= 0:
            raise ValueError(f"PML thickness for {wall} must be positive")
        if thickness > p.N[axis]:
            raise ValueError(f"PML thickness for {wall} exceeds active cells on {wall[1]}: "
                             f"{thickness} > {p.N[axis]}")
        order = float(r.get("order", 3.0))
        if "sigma_max" in r:
            smax = float(r["sigma_max"])
        else:
            smax = -((order + 1.0) * p.C * math.log(float(r.get("target_reflection", 1.0e-8)))) / (
                2.0 * thickness * p.d[axis])
        out.append((wall, axis, thickness, order, smax))
    return out


def pml_profiles(layers, p: Params):
    """build_pml (PML.py:84-122): (sigma_x, sigma_y, sigma_z) of shape N+2
    per axis; layer cell i (0 at the interface) of thickness T holds
    sigma_max ((i+1)/T)^order, counted from the interface outwards."""
    shape = tuple(n + 2 for n in p.N)
    prof = [np.zeros(shape) for _ in range(3)]
    for wall, axis, T, order, smax in layers:
        for i in range(T):
            idx = T - i if wall[0] == "-" else shape[axis] - 1 - T + i
            sl = [slice(None)] * 3
            sl[axis] = idx
            prof[axis][tuple(sl)] = smax * ((i + 1) / T) ** order
    return prof


def pml_ghosted(profile, p: Params):
    """_tile_scalar_profile for one tile (PML.py:158-189): the one-ghost
    profile placed at offset g-1 in a field-shaped array (zeros elsewhere)."""
    g = p.g
    out = np.zeros(p.shape)
    n = profile.shape
    out[g - 1:g - 1 + n[0], g - 1:g - 1 + n[1], g - 1:g - 1 + n[2]] = profile
    return out


def pml_stretch(deriv, memory, sigma, dt):
    """stretch_spatial_derivative (PML.py:251-267): returns (deriv', memory')."""
    b = np.exp(-sigma * dt)
    m = b * memory + (b - 1.0) * deriv
    return deriv + m, m


# derivative order of the memory tuples and the axis of each derivative's
# stretch (PML.py:270-359): (d?z/dy, d?y/dz, d?x/dz, d?z/dx, d?y/dx, d?x/dy)
PML_DERIV_AXIS = (1, 2, 2, 0, 0, 1)


def _pml_curl(derivs, memory, profiles, p: Params):
    g = p.g
    A = slice(g, -g)
    sig = [pr[A, A, A] for pr in profiles]
    new_d, new_m = [], []
    for dv, m, ax in zip(derivs, memory, PML_DERIV_AXIS):
        a, b = pml_stretch(dv, m, sig[ax], p.dt)
        new_d.append(a)
        new_m.append(b)
    d = new_d
    return (d[0] - d[1], d[2] - d[3], d[4] - d[5]), new_m


def update_B(E, B, p: Params, do_filter=False, pml=None):
    """update_B (first_order_yee.py:96-162).  pml: None or (b_memory (6
    interior arrays), ghosted profiles (3)); with it returns (B, b_memory')."""
    g = p.g
    Ex, Ey, Ez = [refresh(c, p) for c in E]
    dt = p.dt / 2
    dx, dy, dz = p.d
    A = slice(g, -g)
    F = slice(g + 1, None if g == 1 else -g + 1)
    dEz_dy = (Ez[A, F, A] - Ez[A, A, A]) / dy
    dEy_dz = (Ey[A, A, F] - Ey[A, A, A]) / dz
    dEx_dz = (Ex[A, A, F] - Ex[A, A, A]) / dz
    dEx_dy = (Ex[A, F, A] - Ex[A, A, A]) / dy
    dEz_dx = (Ez[F, A, A] - Ez[A, A, A]) / dx
    dEy_dx = (Ey[F, A, A] - Ey[A, A, A]) / dx
    derivs = (dEz_dy, dEy_dz, dEx_dz, dEz_dx, dEy_dx, dEx_dy)
    if pml is None:
        curl = (dEz_dy - dEy_dz, dEx_dz - dEz_dx, dEy_dx - dEx_dy)
    else:
        curl, mem = _pml_curl(derivs, pml[0], pml[1], p)
    out = []
    for c in range(3):
        b = np.array(B[c], dtype=np.float64, copy=True)
        b[A, A, A] = b[A, A, A] - dt * curl[c]
        out.append(b)
    if do_filter:
        out = [digital_filter(refresh(b, p), p.alpha, g) for b in out]
    out = [refresh(b, p) for b in out]
    return out if pml is None else (out, mem)


def update_E(E, B, J, p: Params, pml=None):
    """update_E (first_order_yee.py:12-93).  pml: None or (e_memory (6
    interior arrays), ghosted profiles (3)); with it returns (E, e_memory')."""
    g = p.g
    Bx, By, Bz = [refresh(c, p) for c in B]
    dx, dy, dz = p.d
    A = slice(g, -g)
    K = slice(g - 1, -g - 1)
    dBz_dy = (Bz[A, A, A] - Bz[A, K, A]) / dy
    dBy_dz = (By[A, A, A] - By[A, A, K]) / dz
    dBx_dz = (Bx[A, A, A] - Bx[A, A, K]) / dz
    dBx_dy = (Bx[A, A, A] - Bx[A, K, A]) / dy
    dBz_dx = (Bz[A, A, A] - Bz[K, A, A]) / dx
    dBy_dx = (By[A, A, A] - By[K, A, A]) / dx
    if pml is None:
        curl = (dBz_dy - dBy_dz, dBx_dz - dBz_dx, dBy_dx - dBx_dy)
    else:
        curl, mem = _pml_curl((dBz_dy, dBy_dz, dBx_dz, dBz_dx, dBy_dx, dBx_dy), pml[0], pml[1], p)
    out = []
    for c in range(3):
        e = np.array(E[c], dtype=np.float64, copy=True)
        e[A, A, A] = e[A, A, A] + (p.C ** 2 * curl[c] - J[c][A, A, A] / p.eps) * p.dt
        out.append(e)
    out = [refresh(e, p) for e in out]
    out = [digital_filter(e, p.alpha, g) for e in out]
    Ex, Ey, Ez = out
    if p.bc[0] == BC_CONDUCTING:
        Ey, Ez = zero_boundary(Ey, p, 0), zero_boundary(Ez, p, 0)
    if p.bc[1] == BC_CONDUCTING:
        Ex, Ez = zero_boundary(Ex, p, 1), zero_boundary(Ez, p, 1)
    if p.bc[2] == BC_CONDUCTING:
        Ex, Ey = zero_boundary(Ex, p, 2), zero_boundary(Ey, p, 2)
    out = [refresh(e, p) for e in (Ex, Ey, Ez)]
    return out if pml is None else (out, mem)


# --------------------------------------------------------------------------
# electrostatic solver (solvers/electrostatic_yee.py)
# --------------------------------------------------------------------------

def laplacian_interior(f, p: Params):
    """lapl (electrostatic_yee.py:98-103)."""
    g = p.g
    A = slice(g, -g)
    F = slice(g + 1, None if g == 1 else -g + 1)
    K = slice(g - 1, -g - 1)
    dx, dy, dz = p.d
    return ((f[F, A, A] + f[K, A, A] - 2.0 * f[A, A, A]) / (dx * dx)
            + (f[A, F, A] + f[A, K, A] - 2.0 * f[A, A, A]) / (dy * dy)
            + (f[A, A, F] + f[A, A, K] - 2.0 * f[A, A, A]) / (dz * dz))


def solve_poisson_cg(rho, phi, p: Params, tol=1e-12, max_iter=5000, return_iterations=False):
    """solve_poisson_with_conjugate_gradient (electrostatic_yee.py:71-156)."""
    g = p.g
    A = slice(g, -g)
    apply_bc = lambda f: phi_boundaries(f, p)   # noqa: E731
    phi = apply_bc(np.asarray(phi, dtype=np.float64))
    r = np.asarray(rho, dtype=np.float64)[A, A, A] / p.eps + laplacian_interior(phi, p)
    pvec = np.zeros_like(phi)
    pvec[A, A, A] = r
    pvec = apply_bc(pvec)
    k = 0
    while k < max_iter and np.sum(r * r) > tol ** 2:
        lapl_p = -laplacian_interior(pvec, p)
        alpha = np.sum(r * r) / np.sum(pvec[A, A, A] * lapl_p)
        phi = phi.copy()
        phi[A, A, A] += alpha * pvec[A, A, A]
        phi = apply_bc(phi)
        r_next = r - alpha * lapl_p
        beta = np.sum(r_next * r_next) / np.sum(r * r)
        pn = pvec.copy()
        pn[A, A, A] = r_next + beta * pvec[A, A, A]
        pvec = apply_bc(pn)
        r = r_next
        k += 1
    phi = apply_bc(phi)
    return (phi, k) if return_iterations else phi


def centered_gradient_E(phi, p: Params):
    """_centered_tiled_electrostatic_gradient (electrostatic_yee.py:212-246)."""
    g = p.g
    A = slice(g, -g)
    F = slice(g + 1, None if g == 1 else -g + 1)
    K = slice(g - 1, -g - 1)
    phi = phi_boundaries(phi, p)
    dx, dy, dz = p.d
    E = [np.zeros_like(phi) for _ in range(3)]
    E[0][A, A, A] = -1.0 * (phi[F, A, A] - phi[K, A, A]) / (2.0 * dx)
    E[1][A, A, A] = -1.0 * (phi[A, F, A] - phi[A, K, A]) / (2.0 * dy)
    E[2][A, A, A] = -1.0 * (phi[A, A, F] - phi[A, A, K]) / (2.0 * dz)
    return [refresh(e, p) for e in E]


def electrostatic_fields(x, active, sp: Species, rho, phi, p: Params, tol=1e-12, max_iter=5000):
    """calculate_tiled_electrostatic_fields (electrostatic_yee.py:249-276)."""
    del rho   # used only as a template by the reference (rho.py:62)
    rho = compute_rho(x, active, sp, p)
    phi = solve_poisson_cg(rho, phi, p, tol=tol, max_iter=max_iter)
    phi = phi_boundaries(phi, p)
    phi = digital_filter(phi, p.alpha, p.g)
    phi = phi_boundaries(phi, p)
    E = centered_gradient_E(phi, p)
    return E, phi, rho


def roll_gradient_E(phi, p: Params):
    """_centered_finite_difference_gradient of -phi on the interior
    (electrostatic_yee.py:46-60, 197-207): jnp.roll wraps the interior on every
    axis whatever the field BC; the result is placed in zero ghosted arrays and
    refreshed with the field codes."""
    g = p.g
    A = slice(g, -g)
    f = -1.0 * np.asarray(phi, dtype=np.float64)[A, A, A]
    E = []
    for a in range(3):
        grad = (np.roll(f, -1, axis=a) - np.roll(f, 1, axis=a)) / (2.0 * p.d[a])
        full = np.zeros(p.shape)
        full[A, A, A] = grad
        E.append(refresh(full, p))
    return E


def electrostatic_fields_from_rho(rho, phi, p: Params, tol=1e-12, max_iter=5000):
    """calculate_electrostatic_fields (electrostatic_yee.py:159-209): CG on the
    given rho (no deposit), phi boundaries, digital filter, phi boundaries,
    rolled centred gradient.  Returns (E, phi, rho)."""
    phi = solve_poisson_cg(rho, phi, p, tol=tol, max_iter=max_iter)
    phi = phi_boundaries(phi, p)
    phi = digital_filter(phi, p.alpha, p.g)
    phi = phi_boundaries(phi, p)
    return roll_gradient_E(phi, p), phi, rho


# --------------------------------------------------------------------------
# particle positions and boundaries (particles/particle_tile_communication.py)
# --------------------------------------------------------------------------

def update_positions(x, u, active, sp: Species, dt):
    """update_tiled_particle_positions (particle_tile_communication.py:82-99)."""
    S, P, _ = x.shape
    out = np.array(x, dtype=np.float64, copy=True)
    act = np.asarray(active, dtype=bool)
    for a in range(3):
        upd = np.asarray(sp.update_x, dtype=bool)[:, a][:, None]
        step = act.astype(np.float64) * np.asarray(u[..., a], dtype=np.float64) * dt
        out[..., a] = np.where(act & upd, out[..., a] + step, out[..., a])
    return out


def wrap_periodic(x, wind):
    """wrap_periodic_position (grid_and_stencil.py:14-35)."""
    half = 0.5 * wind
    wrapped = np.mod(x + half, wind) - half
    return np.where((wrapped == -half) & (x >= half), half, wrapped)


def refresh_particles(x, u, active, p: Params):
    """refresh_tiled_particle_tiles on one device (particle_tile_communication.py:
    41-59, 233-289, 297-375): axis boundaries, then zero x,u of inactive slots.
    Returns (x, u, active, overflow) with overflow False (no tile moves)."""
    x = np.array(x, dtype=np.float64, copy=True)
    u = np.array(u, dtype=np.float64, copy=True)
    act = np.asarray(active, dtype=bool).copy()
    for a in range(3):
        half = 0.5 * p.wind[a]
        xa, ua = x[..., a].copy(), u[..., a].copy()
        bc = p.particle_bc[a]
        if bc == PBC_PERIODIC:
            x[..., a] = wrap_periodic(xa, p.wind[a])
        elif bc == PBC_REFLECTING:
            x[..., a] = np.where(xa > half, 2.0 * half - xa, np.where(xa < -half, -2.0 * half - xa, xa))
            u[..., a] = np.where((xa >= half) | (xa <= -half), -ua, ua)
        elif bc == PBC_ABSORBING:
            act = act & (xa <= half) & (xa >= -half)
    x = np.where(act[..., None], x, 0.0)
    u = np.where(act[..., None], u, 0.0)
    return x, u, act, False


def cell_keys(x, p: Params):
    """Flat cell key (ix*Ny + iy)*Nz + iz of _particle_tile_indices with one
    cell per tile (particle_initialization.py:89-102, 140-160)."""
    cells = []
    for a in range(3):
        c = np.floor((np.asarray(x[..., a], dtype=np.float64) + p.wind[a] / 2) / p.d[a]).astype(np.int64)
        cells.append(np.clip(c, 0, p.N[a] - 1))
    return (cells[0] * p.N[1] + cells[1]) * p.N[2] + cells[2]


def sort_permutation(keys):
    """np.argsort(kind="stable") -- the packing order (particle_initialization.py:159)."""
    return np.argsort(np.asarray(keys), kind="stable")


# --------------------------------------------------------------------------
# diagnostics (utils.py)
# --------------------------------------------------------------------------

def compute_energy(u, active, sp: Species, E, B, p: Params):
    """compute_energy (utils.py:108-187): (U_E, U_B, KE)."""
    g = p.g
    A = (slice(g, -g),) * 3
    dV = p.d[0] * p.d[1] * p.d[2]
    E2 = sum(np.sum(np.asarray(c, dtype=np.float64)[A] ** 2) for c in E) * dV
    B2 = sum(np.sum(np.asarray(c, dtype=np.float64)[A] ** 2) for c in B) * dV
    uE = 0.5 * p.eps * E2
    uB = 0.5 / p.mu * B2
    v2 = sum(np.asarray(u[..., a], dtype=np.float64) ** 2 for a in range(3))
    act = np.asarray(active, dtype=np.float64)
    mass = (np.asarray(sp.mass, dtype=np.float64) * np.asarray(sp.weight, dtype=np.float64))[:, None]
    mass = np.broadcast_to(mass, v2.shape)
    gamma = 1.0 / np.sqrt(1 - v2 / p.C ** 2)
    mom2 = np.square(mass * gamma) * v2
    ke = np.sum(act * (np.sqrt(mom2 * p.C ** 2 + mass ** 2 * p.C ** 4) - mass * p.C ** 2))
    return uE, uB, ke


def compute_total_momentum(u, active, sp: Species):
    """compute_total_momentum (utils.py:190-202): sum active |v| mass weight."""
    v = np.sqrt(sum(np.asarray(u[..., a], dtype=np.float64) ** 2 for a in range(3)))
    mass = (np.asarray(sp.mass, dtype=np.float64) * np.asarray(sp.weight, dtype=np.float64))[:, None]
    return float(np.sum(np.asarray(active, dtype=np.float64) * v * mass))


def total_charge(rho, p: Params):
    """Sum of rho over the interior times the cell volume."""
    g = p.g
    return float(np.sum(np.asarray(rho, dtype=np.float64)[g:-g, g:-g, g:-g]) * p.d[0] * p.d[1] * p.d[2])


def courant_condition(cfl, p: Params):
    """courant_condition (utils.py:761-801): dt = cfl / (C * sum_{N>1} 1/d)."""
    inv = sum(1 / d for d, n in zip(p.d, p.N) if n > 1)
    return cfl / (p.C * inv)


# --------------------------------------------------------------------------
# one step (evolve.py)
# --------------------------------------------------------------------------

@dataclass
class State:
    x: np.ndarray
    u: np.ndarray
    active: np.ndarray
    E: list
    B: list
    J: list = None
    rho: np.ndarray = None
    phi: np.ndarray = None
    E_ext: list = None
    B_ext: list = None

    def copy(self):
        cp = lambda v: None if v is None else (   # noqa: E731
            [np.array(c, dtype=np.float64) for c in v] if isinstance(v, (list, tuple))
            else np.array(v, dtype=np.float64 if v.dtype != bool else bool))
        return State(cp(self.x), cp(self.u), np.array(self.active, dtype=bool), cp(self.E), cp(self.B),
                     cp(self.J), cp(self.rho), cp(self.phi), cp(self.E_ext), cp(self.B_ext))


def time_loop_electrodynamic(st: State, sp: Species, p: Params):
    """time_loop_electrodynamic (evolve.py:16-103); returns a new State."""
    E, B = st.E, st.B
    pE = E if st.E_ext is None else [e + ee for e, ee in zip(E, st.E_ext)]
    pB = B if st.B_ext is None else [b + bb for b, bb in zip(B, st.B_ext)]
    u = particle_push(st.x, st.u, st.active, sp, pE, pB, p)
    x, act = st.x, st.active
    if p.current_deposition == "esirkepov":
        J = esirkepov_current(x, u, act, sp, p)
        x = update_positions(x, u, act, sp, p.dt)
        x, u, act, _ = refresh_particles(x, u, act, p)
    else:
        x = update_positions(x, u, act, sp, p.dt / 2)
        x, u, act, _ = refresh_particles(x, u, act, p)
        J = j_from_rhov(x, u, act, sp, p)
        x = update_positions(x, u, act, sp, p.dt / 2)
        x, u, act, _ = refresh_particles(x, u, act, p)
    B = update_B(E, B, p, do_filter=False)
    E = update_E(E, B, J, p)
    B = update_B(E, B, p, do_filter=True)
    return State(x, u, act, E, B, J, st.rho, st.phi, st.E_ext, st.B_ext)


def time_loop_electrostatic(st: State, sp: Species, p: Params, tol=1e-12, max_iter=5000):
    """time_loop_electrostatic (evolve.py:106-161)."""
    pE = st.E if st.E_ext is None else [e + ee for e, ee in zip(st.E, st.E_ext)]
    pB = st.B if st.B_ext is None else [b + bb for b, bb in zip(st.B, st.B_ext)]
    u = particle_push(st.x, st.u, st.active, sp, pE, pB, p)
    x = update_positions(st.x, u, st.active, sp, p.dt)
    x, u, act, _ = refresh_particles(x, u, st.active, p)
    E, phi, rho = electrostatic_fields(x, act, sp, st.rho, st.phi, p, tol=tol, max_iter=max_iter)
    return State(x, u, act, E, st.B, st.J, rho, phi, st.E_ext, st.B_ext)


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("annotations", "math", "np",
                                                                     "dataclass", "field", "replace",
                                                                     "Tuple")]
