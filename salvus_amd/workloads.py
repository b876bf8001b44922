"""Synthetic inputs of the five benchmark configurations (seeded).

Every choice the reference does not fix is marked "chosen here".  The
arrays follow the reference's input conventions (cells in QuadP1/HexP1
vertex order; per-element, per-vertex material in the reference's
parameter order).

  cfg1  2-D acoustic, 50x50 quads over 1000 m, order 4, vp 4000, FP64 case
  cfg2  2-D elastic, 500x500 quads of 10 m, vp 3000 / vs 1732 / rho 2500
        with a Gaussian-correlated +-5 % perturbation (50 m), vertical
        Ricker 30 Hz, 100 receivers, dt 2e-4, 5000 steps
  cfg3  3-D acoustic, 128^3 hexes of 15 m, 8 layers 2000->5000 m/s +-3 %,
        Ricker 20 Hz, 1024 surface receivers, 2000 steps at CFL 0.4
  cfg4  3-D elastic shots on 32^3, 48^3, 64^3, 96^3 meshes (8 each),
        moment-tensor Ricker 15 Hz, 1500 steps
  cfg5  serving mix (see salvus_amd.serve)
"""
import numpy as np

from . import meshgen

GLL4_MIN_SPACING = (1.0 - np.sqrt(3.0 / 7.0)) / 2.0  # smallest GLL gap / h at order 4


def gaussian_field(shape, sigma_cells, rng):
    """White noise smoothed by a Gaussian of sigma_cells grid cells, scaled
    to max |value| = 1 (chosen here: noise on the vertex grid)."""
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), sigma_cells, mode="reflect")
    return f / np.abs(f).max()


def cfg1(nsteps=1000):
    cells, coords = meshgen.box_quad(50, 50, 1000.0, 1000.0)
    par = np.full((len(cells), 4, 1), 4000.0)
    f0 = 100.0
    return dict(name="cfg1: 2-D acoustic 50x50 quads, order 4", dim=2, physics=0, order=4,
                cells=cells, coords=coords, par=par, dt=2e-4, nsteps=nsteps,
                sources=[dict(kind="force", loc=(500.0, 500.0), amp=1.0, f0=f0, t0=1.5 / f0,
                              dir=(1.0,))],
                receivers=[(x, 700.0) for x in (300.0, 400.0, 500.0, 600.0, 700.0)])


def cfg2(n=500, h=10.0, nsteps=5000, nrec=100, seed=2):
    """2-D elastic; n and nsteps can be reduced for parity samples (same h,
    material law, source and receiver line scaled to the domain)."""
    rng = np.random.default_rng(seed)
    L = n * h
    cells, coords = meshgen.box_quad(n, n, L, L)
    d = 0.05 * gaussian_field((n + 1, n + 1), 50.0 / h, rng).ravel()  # +-5 %, 50 m
    vp, vs, rho = 3000.0 * (1 + d), 1732.0 * (1 + d), 2500.0 * (1 + d)
    # Elastic2D.cpp:53-82: C11 = C33 = rho vp^2, C13 = rho (vp^2 - 2 vs^2), C55 = rho vs^2
    c11 = rho * vp ** 2
    par = np.stack([rho, c11, rho * (vp ** 2 - 2 * vs ** 2), c11, rho * vs ** 2], axis=1)[cells]
    f0 = 30.0
    src = dict(kind="force", loc=(0.5 * L + 0.37 * h, 0.5 * L + 0.41 * h), amp=1e9, f0=f0,
               t0=1.2 / f0, dir=(0.0, 1.0))  # vertical force (the reference's CLI cannot set a direction)
    recs = [(0.1 * L + 0.8 * L * k / (nrec - 1), 0.7 * L) for k in range(nrec)]
    return dict(name="cfg2: 2-D elastic %dx%d quads of %g m, order 4" % (n, n, h), dim=2,
                physics=1, order=4, cells=cells, coords=coords, par=par, dt=2e-4,
                nsteps=nsteps, sources=[src], receivers=recs)


def cfg3(n=128, h=15.0, nsteps=2000, nrec_side=32, seed=3):
    rng = np.random.default_rng(seed)
    L = n * h
    cells, coords = meshgen.box_hex(n, n, n, L, L, L)
    z = coords[:, 2]
    layer = np.minimum((z / L * 8).astype(int), 7)
    vp_v = (2000.0 + 3000.0 * layer / 7.0) * (1 + rng.uniform(-0.03, 0.03, len(coords)))
    par = vp_v[cells][..., None]
    dt = 0.4 * h * GLL4_MIN_SPACING / vp_v.max()  # CFL number chosen here (the reference has no CFL logic)
    f0 = 20.0
    c = 0.5 * L
    src = dict(kind="force", loc=(c + 0.37 * h, c + 0.41 * h, c + 0.29 * h), amp=1e6, f0=f0,
               t0=1.2 / f0, dir=(1.0,))
    xs = (np.arange(nrec_side) + 0.5) / nrec_side * L
    recs = [(x, y, L) for y in xs for x in xs]
    return dict(name="cfg3: 3-D acoustic %d^3 hexes of %g m, order 4" % (n, h), dim=3,
                physics=0, order=4, cells=cells, coords=coords, par=par, dt=dt, nsteps=nsteps,
                sources=[src], receivers=recs)


def elastic3d_mesh(n, h=15.0, seed=4):
    """cfg4 mesh: seeded smooth vp around 3000 m/s (100 m Gaussian, chosen
    here), vs = vp/1.73, rho = 310 vp^0.25 (Gardner, chosen here)."""
    rng = np.random.default_rng(seed + n)
    L = n * h
    cells, coords = meshgen.box_hex(n, n, n, L, L, L)
    d = 0.05 * gaussian_field((n + 1,) * 3, 100.0 / h, rng).ravel()
    vp = 3000.0 * (1 + d)
    vs = vp / 1.73
    rho = 310.0 * vp ** 0.25
    par = np.stack([rho, vp, vp, vs, vs, np.ones_like(vp)], axis=1)[cells]  # isotropic: VPV = VPH, VSV = VSH, ETA = 1 (Elastic3D.cpp:42-52)
    dt = 0.4 * h * GLL4_MIN_SPACING / vp.max()
    return cells, coords, par, dt, L


def moment_tensor(rng, scale=1e15):
    M = rng.normal(0.0, 1.0, size=(3, 3)) * scale
    return 0.5 * (M + M.T)


def cfg4(sizes=(32, 48, 64, 96), shots_per_mesh=8, nsteps=1500, h=15.0, seed=4):
    meshes = []
    for n in sizes:
        cells, coords, par, dt, L = elastic3d_mesh(n, h, seed)
        rng = np.random.default_rng(seed * 1000 + n)
        shots = []
        for _ in range(shots_per_mesh):
            loc = tuple(rng.uniform(0.2, 0.8, 3) * L)
            shots.append(dict(kind="moment", loc=loc, M=moment_tensor(rng), amp=1.0,
                              f0=15.0, t0=1.2 / 15.0))
        recs = [(0.5 * L, 0.5 * L, L), (0.25 * L, 0.75 * L, L), (0.75 * L, 0.25 * L, L)]
        meshes.append(dict(n=n, cells=cells, coords=coords, par=par, dt=dt, shots=shots,
                           receivers=recs))
    return dict(name="cfg4: 3-D elastic multi-shot, meshes %s, %d shots each"
                % ("/".join("%d^3" % n for n in sizes), shots_per_mesh),
                dim=3, physics=1, order=4, meshes=meshes, nsteps=nsteps)


def unique_nodes(dim, n, order=4):
    return (n * order + 1) ** dim
