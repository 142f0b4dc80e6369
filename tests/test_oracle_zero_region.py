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
