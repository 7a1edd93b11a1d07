"""Initial conditions for the hydro path's workloads, generated on the host.

One generator feeds both the HIP path and the CPU checks, so both start from the same bits. Each
problem returns primitives on cell centres, ``(i + 1/2) dx + lo`` with ``lo = -box/2``
(src/domain.ts:1-40), and :func:`cons_from_prims` turns them into the reference's state layout:
``u1 = (rho, rho v)`` and ``u2 = (E, rho e, rho s, mark)``, one float4 pair per cell, cell
``((k ny + j) nx + i)`` (src/sim.ts:3754-3757). The conversion is done in float64 and rounded
once, as the reference's Problem.init fills Float32 arrays (src/sim.ts:2704-2825).

Problems:

- :func:`sod_tube`: the reference's sod-x/y/z presets (src/presets/sod-x.json).
- :func:`kelvin_helmholtz`: 2D KH on the unit square, periodic, gamma 5/3, slab |y - 1/2| < 1/4.
- :func:`turbulence`: 3D decaying turbulence, random-phase solenoidal field, P(k) ~ k^-2 on
  1 <= |k| <= 8, rms Mach 1.
- :func:`sedov`: 3D Sedov blast, E = 1 in r < 3 cells, ambient rho 1, P 1e-5, noise 1e-4.
- :func:`riemann_quadrants`: 2D four-quadrant Riemann problem with seeded states.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    """A workload: grid, boundaries, gas, and solver settings (the SimConfig fields it sets)."""

    id: str
    n: tuple            # cells along x, y, z
    box: float          # side of the domain along x
    gamma: float
    bc_lo: tuple = (0, 0, 0)
    bc_hi: tuple = (0, 0, 0)
    t_end: float = float("inf")
    cfl: float = 0.8
    limiter: str = "koren"
    riemann: str = "hllc"
    isentropic: bool = True
    rho_floor: float = 1e-10
    p_floor: float = 1e-14
    prims: dict = field(default_factory=dict)   # rho, vx, vy, vz, p, s: arrays of n[2]*n[1]*n[0]
    dims: int = 0       # the run's number of dimensions (0: 2 if n[2] == 1 else 3)

    @property
    def ndim(self):
        if self.dims:
            return self.dims
        return 2 if self.n[2] == 1 else 3

    @property
    def dx(self):
        return self.box / self.n[0]

    @property
    def ncell(self):
        return self.n[0] * self.n[1] * self.n[2]

    def state(self, dtype=np.float32):
        return cons_from_prims(self.prims, self.gamma, dtype)


def centers(n, box):
    """Cell-centre coordinates (x, y, z), each of shape (nz, ny, nx), box centred on 0."""
    dx = box / n[0]
    ax = [(np.arange(n[a]) + 0.5) * dx - 0.5 * n[a] * dx for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return x, y, z


def cons_from_prims(pr, gamma, dtype=np.float32):
    """(u1, u2) as (ncell, 4) arrays from primitives, in float64 then rounded once."""
    rho0 = np.asarray(pr["rho"], np.float64)
    rho = rho0.reshape(-1)

    def field(key):
        return np.broadcast_to(np.asarray(pr.get(key, 0.0), np.float64), rho0.shape).reshape(-1)

    vx, vy, vz, p, s = field("vx"), field("vy"), field("vz"), field("p"), field("s")
    eint = p / (gamma - 1.0)
    u1 = np.stack([rho, rho * vx, rho * vy, rho * vz], axis=1)
    E = eint + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    u2 = np.stack([E, eint, rho * s, np.zeros_like(rho)], axis=1)
    return np.ascontiguousarray(u1, dtype), np.ascontiguousarray(u2, dtype)


def sod_tube(n: int, axis: int = 0, ndim: int = 3, transverse: int = 1):
    """Sod shock tube along `axis` (src/presets/sod-x.json): gamma 1.4, outflow, left
    (rho, p, s) = (1, 1, 1) below the midplane, right (0.125, 0.1, 0), t_end 0.2, p floor 1e-12.
    `transverse` cells across the tube (1 gives a 1D line with the run's ndim kept)."""
    dims = [transverse] * 3
    dims[axis] = n
    if ndim == 2 and axis != 2:
        dims[2] = 1
    box = 1.0
    x, y, z = centers_dx(dims, box / n)
    s = (x, y, z)[axis]
    left = s < 0
    pr = {"rho": np.where(left, 1.0, 0.125), "p": np.where(left, 1.0, 0.1), "s": np.where(left, 1.0, 0.0)}
    bc = [0, 0, 0]
    bc[axis] = 1
    return Problem(id=f"sod-{'xyz'[axis]}", n=tuple(dims), box=box * dims[0] / n, gamma=1.4, bc_lo=tuple(bc),
                   bc_hi=tuple(bc), t_end=0.2, p_floor=1e-12, prims=pr, dims=ndim)


def centers_dx(n, dx):
    ax = [(np.arange(n[a]) + 0.5) * dx - 0.5 * n[a] * dx for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return x, y, z


def kelvin_helmholtz(n: int = 256, seed: int = 42, cfl: float = 0.4, limiter="mc", isentropic=False):
    """2D Kelvin-Helmholtz on the unit square, periodic, gamma 5/3: rho 2 and vx +0.5 where
    |y - 1/2| < 1/4, rho 1 and vx -0.5 elsewhere, P 2.5; vy = 0.01 sin(4 pi x) plus seeded
    uniform noise of amplitude 1e-3; the slab carries the passive scalar s = 1."""
    dims = (n, n, 1)
    x, y, _ = centers(dims, 1.0)
    xu, yu = x + 0.5, y + 0.5              # positions on [0, 1]
    inside = np.abs(yu - 0.5) < 0.25
    rng = np.random.default_rng(seed)
    noise = rng.uniform(-1e-3, 1e-3, size=x.shape)
    pr = {
        "rho": np.where(inside, 2.0, 1.0),
        "vx": np.where(inside, 0.5, -0.5),
        "vy": 0.01 * np.sin(4 * np.pi * xu) + noise,
        "p": np.full(x.shape, 2.5),
        "s": np.where(inside, 1.0, 0.0),
    }
    return Problem(id="kelvin-helmholtz", n=dims, box=1.0, gamma=5.0 / 3.0, cfl=cfl, limiter=limiter,
                   isentropic=isentropic, prims=pr)


def turbulence(n: int = 256, seed: int = 42, kmin=1.0, kmax=8.0, mach=1.0, limiter="mc", isentropic=False,
               cfl=0.4):
    """3D decaying turbulence, periodic unit box, gamma 5/3, rho 1, P 1: a Gaussian random-phase
    velocity field with power spectrum P(k) ~ k^-2 on kmin <= |k| <= kmax (k in units of the
    box's fundamental), projected solenoidal, scaled to rms Mach `mach` (v_rms = c_s)."""
    gamma = 5.0 / 3.0
    dims = (n, n, n)
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz = np.fft.rfftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, kz, indexing="ij")
    K = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    band = (K >= kmin) & (K <= kmax)
    # amplitude: |v_k|^2 4 pi k^2 dk ~ P(k) = k^-2  =>  |v_k| ~ k^-2
    amp = np.zeros_like(K)
    amp[band] = K[band] ** -2.0
    comps = []
    shape = K.shape
    for _ in range(3):
        ph = rng.uniform(0.0, 2 * np.pi, size=shape)
        g = rng.normal(size=shape)
        comps.append(amp * g * np.exp(1j * ph))
    vk = np.stack(comps)
    # solenoidal projection: v_k - k (k . v_k) / k^2
    Ksafe = np.where(K > 0, K, 1.0)
    kv = (KX * vk[0] + KY * vk[1] + KZ * vk[2]) / (Ksafe * Ksafe)
    vk[0] -= KX * kv
    vk[1] -= KY * kv
    vk[2] -= KZ * kv
    v = [np.fft.irfftn(vk[a], s=(n, n, n)) for a in range(3)]
    # array axes are (ix, iy, iz); the state is indexed (k, j, i)
    v = [np.ascontiguousarray(np.transpose(c, (2, 1, 0))) for c in v]
    vrms = np.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    cs = np.sqrt(gamma * 1.0 / 1.0)
    scale = mach * cs / vrms
    pr = {"rho": np.ones(dims[::-1]), "vx": v[0] * scale, "vy": v[1] * scale, "vz": v[2] * scale,
          "p": np.ones(dims[::-1])}
    return Problem(id="turbulence", n=dims, box=1.0, gamma=gamma, cfl=cfl, limiter=limiter, isentropic=isentropic,
                   prims=pr)


def sedov(n: int = 128, seed: int = 42, gamma: float = 1.4, energy: float = 1.0, radius_cells: float = 3.0,
          p_amb: float = 1e-5, noise: float = 1e-4, limiter="mc", isentropic=False, cfl=0.4, t_end=0.05):
    """3D Sedov blast on the unit box: ambient rho 1 (times 1 + noise U(-1, 1)) and P 1e-5; energy
    E deposited as thermal energy, evenly by volume, in the cells within `radius_cells` cells of
    the centre. Periodic: the blast stays well inside the box until t = 0.05."""
    dims = (n, n, n)
    dx = 1.0 / n
    x, y, z = centers(dims, 1.0)
    r = np.sqrt(x * x + y * y + z * z)
    inside = r < radius_cells * dx
    vol = inside.sum() * dx ** 3
    rng = np.random.default_rng(seed)
    rho = 1.0 + noise * rng.uniform(-1.0, 1.0, size=x.shape)
    p = np.where(inside, (gamma - 1.0) * energy / vol, p_amb)
    pr = {"rho": rho, "p": p}
    return Problem(id="sedov", n=dims, box=1.0, gamma=gamma, cfl=cfl, limiter=limiter, isentropic=isentropic,
                   t_end=t_end, prims=pr)


def quadrant_states(seed: int):
    """Four seeded quadrant states: rho and P log-uniform in [0.1, 10], |v| <= 1 in a uniform
    direction. Returns a (4, 4) array of (rho, vx, vy, p) for quadrants (x<0,y<0), (x>0,y<0),
    (x<0,y>0), (x>0,y>0)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((4, 4))
    for q in range(4):
        rho = 10.0 ** rng.uniform(-1.0, 1.0)
        p = 10.0 ** rng.uniform(-1.0, 1.0)
        speed = rng.uniform(0.0, 1.0)
        ang = rng.uniform(0.0, 2 * np.pi)
        out[q] = (rho, speed * np.cos(ang), speed * np.sin(ang), p)
    return out


def riemann_quadrants(n: int, seed: int, gamma: float = 1.4, limiter="mc", isentropic=False, cfl=0.4):
    """2D four-quadrant Riemann problem on the unit square with outflow boundaries."""
    dims = (n, n, 1)
    x, y, _ = centers(dims, 1.0)
    st = quadrant_states(seed)
    q = (x > 0).astype(int) + 2 * (y > 0).astype(int)
    pr = {"rho": st[q, 0], "vx": st[q, 1], "vy": st[q, 2], "p": st[q, 3]}
    return Problem(id=f"quadrants-{seed}", n=dims, box=1.0, gamma=gamma, bc_lo=(1, 1, 0), bc_hi=(1, 1, 0), cfl=cfl,
                   limiter=limiter, isentropic=isentropic, prims=pr)


def ensemble_sizes(count: int = 64, seed: int = 42, choices=(64, 128, 256, 512)):
    """Seeded grid sizes for the uneven-size ensemble."""
    rng = np.random.default_rng(seed)
    return [int(c) for c in rng.choice(np.asarray(choices), size=count)]


def kh_adaptive(root: int = 16, max_level: int = 4, buffer: float = 0.5):
    """The 2D Kelvin-Helmholtz shear layer (kelvin_helmholtz's state, without the noise) on a
    quadtree: root^2 blocks of 16^2 cells (256^2 at the defaults), refined `max_level` times
    (finest 4096^2) wherever a block grown by `buffer` blocks on each side holds one of the two
    interfaces y = 1/4, 3/4. The defaults give 4096 blocks, 3136 leaves, 802,816 leaf cells, the
    size of the reference's adaptive KH figure at 4096 finest and 4 levels (772,864 leaf cells,
    refined by its own flagging). Returns (blocks, dx0, prim_fn)."""
    from .amr_tree import BS, build_tree
    dx0 = 1.0 / (root * BS)

    def refine(level, b):
        size = BS * dx0 / (1 << level)
        lo = b[1] * size - 0.5
        glo, ghi = lo - buffer * size, lo + (1 + buffer) * size
        return any(glo <= yi <= ghi for yi in (-0.25, 0.25))
    blocks = build_tree((root, root, 1), max_level, refine, ndim=2)

    def prims(x, y, z):
        inside = np.abs(y) < 0.25
        return {"rho": np.where(inside, 2.0, 1.0), "vx": np.where(inside, 0.5, -0.5),
                "vy": 0.01 * np.sin(4 * np.pi * (x + 0.5)), "vz": 0.0 * x, "p": np.full(x.shape, 2.5),
                "s": np.where(inside, 1.0, 0.0)}
    return blocks, dx0, prims


def amr_blast(seed: int = 7, root: int = 8, max_level: int = 3, radius: float = 0.1, p_in: float = 100.0,
              p_out: float = 1.0, gamma: float = 1.4):
    """The block-AMR blast workload: the unit box, root^3 blocks of 16^3 cells, up to `max_level`
    refinements around a sphere of `radius` at a seeded offset centre (uniform in +-0.05 per axis)
    holding P = p_in against rho = 1, P = p_out. The initial density is uniform, so the refinement
    indicator (a relative jump above 0.3 across a block) is evaluated on the pressure: blocks
    straddling the sphere, grown by half a block (amr_tree.blast_tree). Returns (blocks, dx0,
    prim_fn, centre)."""
    from .amr_tree import blast_tree
    c = np.random.default_rng(seed).uniform(-0.05, 0.05, 3)
    blocks, dx0 = blast_tree(root=root, max_level=max_level, radius=radius, center=tuple(c), p_in=p_in, p_out=p_out)

    def prims(x, y, z):
        r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        return {"rho": np.ones_like(x), "p": np.where(r2 < radius * radius, p_in, p_out)}
    return blocks, dx0, prims, c
