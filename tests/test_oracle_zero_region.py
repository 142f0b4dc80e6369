"""The premise of the forward step's zero-tile skip, on the kernel-order
oracle (the arithmetic the GPU matches bit for bit): where every input within
the step's reach (the velocities read the pressure 10 cells out, the pressure
the velocities 10 cells out) is +0, every output is +0 -- the bit pattern 0,
not -0 -- in the pressure, both velocities and the absorbing layer's split
pair, inside and outside that layer.  A skipped tile leaves exactly these +0
in place, so skipping it changes no bit.  Two steps from a wavefield confined
to a box: outside the box grown by 20 cells per step nothing but +0."""
import numpy as np

from oracle import oracle as O


def test_zero_neighbourhood_steps_to_plus_zero():
    nz, nx, npml = 60, 140, 20
    rng = np.random.default_rng(31)
    vp = rng.uniform(1800, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2300, (nz, nx)).astype(np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=npml, nt=10, flag_layout=2, spx0=500.0, dt=1e-3)
    g = O.geometry(P, 1)
    F = O.alloc_fields(g, npml)
    z0, z1, x0, x1 = 45, 52, 90, 101                      # padded cells holding the wavefield
    for k in ("p", "vx", "vz"):
        F[k][z0:z1 + 1, x0:x1 + 1] = rng.standard_normal((z1 - z0 + 1, x1 - x0 + 1)).astype(np.float32)
    nsteps, reach = 2, 20
    O.fd_steps(P, 1, vp, rho, F, nsteps, korder=True)
    d = nsteps * reach
    za, zb, xa, xb = z0 - d, z1 + d, x0 - d, x1 + d

    def far(shape, zoff=0, xoff=0):
        zz = np.arange(shape[0])[:, None] + zoff
        xx = np.arange(shape[1])[None, :] + xoff
        return (zz < za) | (zz > zb) | (xx < xa) | (xx > xb)

    nzp, nxp, nxl = g.nzp, g.nxp, g.nxl
    spread = 0
    for k, (zoff, xoff) in {"p": (0, 0), "vx": (0, 0), "vz": (0, 0),
                            "pl1": (0, 0), "pl2": (0, 0), "pr1": (0, nxp - npml), "pr2": (0, nxp - npml),
                            "pt1": (0, npml), "pt2": (0, npml),
                            "pb1": (nzp - npml, npml), "pb2": (nzp - npml, npml)}.items():
        a = F[k]
        m = far(a.shape, zoff, xoff)
        assert m.any(), k
        bits = a.view(np.uint32)[m]
        assert np.all(bits == 0), (k, int(np.count_nonzero(bits)))
        spread += int(np.count_nonzero(a[~m]))
    # the wavefield did move (not vacuous), into the absorbing layer too
    assert spread > 3 * (z1 - z0 + 1) * (x1 - x0 + 1)
    assert sum(np.count_nonzero(F[k]) for k in ("pl1", "pl2", "pr1", "pr2", "pt1", "pt2", "pb1", "pb2")) > 0


def test_one_step_reaches_19_rows_and_columns():
    """The reach the skip's neighbourhood must cover: one kernel-order step
    from a single non-zero pressure cell in the interior leaves non-zero values
    exactly 19 rows and 19 columns out (velocities 10 / 9 cells from the
    pressure, the pressure 9 / 10 cells from them), and +0 beyond"""
    nz, nx, npml = 80, 120, 20
    vp = np.full((nz, nx), 2500.0, np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=npml, nt=10, flag_layout=2, spx0=600.0, dt=1e-3)
    g = O.geometry(P, 1)
    F = O.alloc_fields(g, npml)
    r, c = 60, 70
    F["p"][r, c] = 1.0
    O.fd_steps(P, 1, vp, rho, F, 1, korder=True)
    rows = np.nonzero(F["p"].any(1))[0]
    cols = np.nonzero(F["p"].any(0))[0]
    assert rows.min() == r - 19 and rows.max() == r + 19
    assert cols.min() == c - 19 and cols.max() == c + 19
    for k in ("vx", "vz"):
        a = F[k]
        zr, xr = np.nonzero(a.any(1))[0], np.nonzero(a.any(0))[0]
        assert r - 19 <= zr.min() and zr.max() <= r + 19 and c - 19 <= xr.min() and xr.max() <= c + 19


def _tile_any(m, tz):
    nzt, nxt = -(-m.shape[0] // tz), -(-m.shape[1] // 64)
    return np.array([[m[a * tz:(a + 1) * tz, b * 64:(b + 1) * 64].any() for b in range(nxt)] for a in range(nzt)])


def test_skip_rule_by_tile_height():
    """The forward step's skip rule on exact kernel-order fields, tile by tile
    as the kernel applies it (tiles of tz rows and 64 columns from the padded
    origin; a tile skips step it when its words and its 8 neighbours' at the
    level read and its own at the level written are zero; source and receiver
    tiles and the first two steps never skip).  A very strong source placed so
    that the wavefield front sits in the last rows of a tile at step 2: the
    front then moves 18 rows in one step, across a whole 16-row tile into the
    next, so with 16-row tiles the rule would skip a tile whose exact output is
    not +0 -- the engine therefore runs the skip on 32- and 64-row tiles only,
    where the rule never fails (the 19-row reach stays within the adjacent
    tiles)."""
    nz, nx, npml = 200, 400, 40
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[100:] = 2600.0
    rho = np.full((nz, nx), 2000.0, np.float32)
    P = O.make_params(nshot=1, flag_layout=2, nx=nx, nz=nz, nt=40, dx=10.0, dz=10.0, dt=1e-3, npml=npml,
                      fdom=20.0, spx0=2000.0, spz0=150.0, offsetmax=4000.0, amp=1e30, alp=0.05)
    g = O.geometry(P, 1)
    F = O.alloc_fields(g, npml)
    nzp, nxp, nxl = g.nzp, g.nxp, g.nxl

    def nonzero_cells():
        m = np.zeros((nzp, nxp), bool)
        m |= F["p"].view(np.uint32) != 0
        m[:, :nxp - 1] |= F["vx"].view(np.uint32) != 0
        m[:nzp - 1] |= F["vz"].view(np.uint32) != 0
        for k, (z, x) in {"pl": (slice(None), slice(0, npml)), "pr": (slice(None), slice(nxp - npml, None)),
                          "pt": (slice(0, npml), slice(npml, npml + nxl)),
                          "pb": (slice(nzp - npml, None), slice(npml, npml + nxl))}.items():
            m[z, x] |= (F[k + "1"].view(np.uint32) != 0) | (F[k + "2"].view(np.uint32) != 0)
        return m

    nsteps = 8
    states = [nonzero_cells()]
    for it in range(nsteps):
        O.fd_steps(P, 1, vp, rho, F, 1, it0=it, with_source=True, korder=True)
        states.append(nonzero_cells())
    # the source box: cells where amp * exp(-alp^2 d^2) is not 0 in float32
    zz, xx = np.meshgrid(np.arange(nzp), np.arange(nxp), indexing="ij")
    d2 = ((xx - g.nsx) ** 2 * 100.0 + (zz - g.nsz) ** 2 * 100.0).astype(np.float32)
    sz, sx = np.nonzero(np.float32(P.amp) * np.exp(np.float32(-P.alp * P.alp) * d2) != 0)
    fails = {}
    for tz in (16, 32, 64):
        W = [_tile_any(s, tz) for s in states]
        nzt, nxt = W[0].shape
        keep = np.zeros((nzt, nxt), bool)
        for a in range(nzt):
            for b in range(nxt):
                z0, z1, x0, x1 = a * tz, a * tz + tz - 1, b * 64, b * 64 + 63
                keep[a, b] = (z0 <= sz.max() and z1 >= sz.min() and x0 <= sx.max() and x1 >= sx.min()) or \
                             (z0 <= npml <= z1 and x0 <= g.rcol0 + g.nrec - 1 and x1 >= g.rcol0)
        fails[tz] = 0
        for it in range(2, nsteps):
            pad = np.zeros((nzt + 2, nxt + 2), bool)
            pad[1:-1, 1:-1] = W[it]
            around = np.zeros((nzt, nxt), bool)
            for dz in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    around |= pad[1 + dz:1 + dz + nzt, 1 + dx:1 + dx + nxt]
            skip = ~keep & ~around & ~W[it - 1]
            fails[tz] += int((skip & W[it + 1]).sum())
    print("rule failures by tile height", fails)
    assert fails[16] > 0, fails
    assert fails[32] == 0 and fails[64] == 0, fails
