"""Host FP64 diagnostics of a state read back from the GPU: the 24 channels of
hydro/diagnostics.wgsl:13-98 (uniform grid) and amr/diagnostics.wgsl:13-117 (leaves of a block
grid), summed with numpy in float64. The GPU reduction (shx_diagnostics) is checked against these
independent sums; conservation claims rest on them."""
import numpy as np

SUM, MAX, MIN = 0, 1, 2
OPS = [SUM] * 7 + [MAX, SUM, SUM] + [MIN, MAX] * 5 + [SUM] * 4


def channels(u1, u2, gamma, rho_floor, p_floor, dv, x, y, z):
    """The 24 channels over cells (u1, u2: (n, 4)), each of volume dv (scalar or (n,)) centred at
    (x, y, z) (each (n,)). Solid cells (mark > 1/2) hold no gas."""
    u1 = np.asarray(u1, np.float64)
    u2 = np.asarray(u2, np.float64)
    gas = u2[:, 3] <= 0.5
    u1, u2 = u1[gas], u2[gas]
    dv = np.broadcast_to(np.asarray(dv, np.float64), gas.shape)[gas]
    x, y, z = (np.broadcast_to(np.asarray(a, np.float64), gas.shape)[gas] for a in (x, y, z))
    gm1 = gamma - 1.0
    rho = np.maximum(u1[:, 0], rho_floor)
    v = u1[:, 1:4] / rho[:, None]
    eint = u2[:, 1]
    p = np.maximum(gm1 * eint, p_floor)
    cs = np.sqrt(gamma * p / rho)
    v2 = (v * v).sum(1)
    speed = np.sqrt(v2)
    T = gm1 * eint / rho
    s = u2[:, 2] / rho
    m = rho * dv
    r = np.zeros(24)
    scale = np.zeros(24)          # sum of |terms| of each summed channel
    r[0] = m.sum()
    r[1:4] = (u1[:, 1:4] * dv[:, None]).sum(0)
    scale[1:4] = np.abs(u1[:, 1:4] * dv[:, None]).sum(0)
    r[4] = (0.5 * rho * v2 * dv).sum()
    eint_tot = (u2[:, 0] - 0.5 * rho * v2) * dv
    r[5] = eint_tot.sum()
    scale[5] = np.abs(eint_tot).sum()
    r[7] = (speed / cs).max()
    r[8] = ((u1[:, 0] <= rho_floor * 1.0001) | (p <= p_floor * 1.0001)).sum()
    for k, q in enumerate((rho, T, s, p, speed)):
        r[10 + 2 * k] = q.min()
        r[11 + 2 * k] = q.max()
    r[20] = (m * x).sum()
    r[21] = (m * y).sum()
    r[22] = (m * z).sum()
    scale[20:23] = [np.abs(m * x).sum(), np.abs(m * y).sum(), np.abs(m * z).sum()]
    for c in range(24):
        scale[c] = max(scale[c], abs(r[c]))
    return r, scale


def uniform(cfg, u1, u2):
    """Channels of a uniform grid (a shaderous_amd.sim.SimConfig) from its read-back state."""
    nx, ny, nz = cfg.n
    dx = cfg.dx
    org = cfg.org if cfg.org is not None else (-0.5 * nx * dx,) * 3
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = (i.ravel() + 0.5) * dx + org[0]
    y = (j.ravel() + 0.5) * dx + org[1]
    z = np.zeros_like(x) if nz == 1 else (k.ravel() + 0.5) * dx + org[2]
    dv = dx ** cfg.ndim
    return channels(u1, u2, cfg.gamma, cfg.rho_floor, cfg.p_floor, dv, x, y, z)


def blocks(cfg, blocks_, leaf, u1, u2):
    """Channels over the leaves of a block grid: blocks_ (nblocks, 4) of (level, ix, iy, iz),
    leaf (nblocks,) bool, u1, u2 of nblocks * 4096 cells; cfg.dx the level-0 cell size."""
    bs = 16
    org = cfg.org if cfg.org is not None else (-0.5 * cfg.n[0] * cfg.dx,) * 3
    c = np.arange(bs ** 3)
    li, lj, lk = c % bs, (c // bs) % bs, c // (bs * bs)
    sel = np.flatnonzero(leaf)
    b = np.asarray(blocks_, np.int64)[sel]
    h = cfg.dx / np.exp2(b[:, 0])
    x = ((b[:, 1:2] * bs + li[None, :]) + 0.5) * h[:, None] + org[0]
    y = ((b[:, 2:3] * bs + lj[None, :]) + 0.5) * h[:, None] + org[1]
    z = ((b[:, 3:4] * bs + lk[None, :]) + 0.5) * h[:, None] + org[2]
    dv = np.repeat(h ** 3, bs ** 3)
    rows = (sel[:, None] * bs ** 3 + c[None, :]).ravel()
    return channels(u1[rows], u2[rows], cfg.gamma, cfg.rho_floor, cfg.p_floor, dv, x.ravel(), y.ravel(), z.ravel())


def compare(gpu, host, rtol=1e-12):
    """Channels where the GPU's value and the host's (a (values, scales) pair from channels())
    differ by more than rtol of the channel's scale (the sum of |terms| for sums, |value| for
    extrema and counts). Returns [(channel, gpu, host)]."""
    vals, scale = host
    bad = []
    for c in range(24):
        g, h = gpu[c], vals[c]
        if abs(g - h) > rtol * max(scale[c], 1e-300):
            bad.append((c, g, h))
    return bad
