"""Seeded case generators in the reference's own input formats.

Each generator writes ``<name>.inp`` (key: value input file,
src/input.cpp:167-643), ``<name>.xyz`` (binary multi-block Plot3D,
src/plot3d.cpp:363-445) and, where the initial state is not uniform, a point
cloud at the cell centres (src/utility.cpp:513-605) referenced by
``icState(tag=-1; file=...)``.

Workloads:
  * ``vortex_case``: inviscid Euler, calorically perfect air, single Cartesian
    block n^3 on [0,10]^3, isentropic vortex (strength 5, core radius 1, axis
    along z through the centre) on a Mach 0.5 freestream along x,
    characteristic farfield BCs, Roe flux, MUSCL (kappa = 1/3) with van Albada
    limiter, implicit Euler LU-SGS at CFL 5.  n = 64 and 200 iterations is the
    reference's CPU-runnable benchmark configuration.
  * ``bump_case``: 3D channel with a sin^2 bump on the lower wall, slip walls,
    characteristic inflow/outflow, Roe + MUSCL, LU-SGS: one steady request of
    the serving workload.
"""
import math
import os
import struct

import numpy as np

AIR_N = 2.5
AIR_MOLAR = 28.97e-3
R_UNIVERSAL = 8.3144598


def air_gas_constant():
    return R_UNIVERSAL / AIR_MOLAR


def write_plot3d(path, blocks):
    """blocks: list of (x, y, z) arrays shaped (ni+1, nj+1, nk+1) [i, j, k]."""
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(blocks)))
        for x, _, _ in blocks:
            f.write(struct.pack("<3i", *x.shape))
        for x, y, z in blocks:
            for c in (x, y, z):
                # Plot3D order: i fastest
                f.write(np.ascontiguousarray(c.transpose(2, 1, 0)).astype("<f8").tobytes())


def cell_centres(x, y, z):
    """Cell centroids in the reference's summation order (src/plot3d.cpp:36-44)."""
    def cen(a):
        s = a[:-1, :-1, :-1].copy()
        s = s + a[1:, :-1, :-1]
        s = s + a[:-1, 1:, :-1]
        s = s + a[1:, 1:, :-1]
        s = s + a[:-1, :-1, 1:]
        s = s + a[1:, :-1, 1:]
        s = s + a[:-1, 1:, 1:]
        s = s + a[1:, 1:, 1:]
        return 0.125 * s
    return cen(x), cen(y), cen(z)


def write_cloud(path, xc, yc, zc, rho, u, v, w, p, species="air"):
    """Point-cloud initial condition; values in SI units, %.17g."""
    n = xc.size
    with open(path, "w") as f:
        f.write(f"{n}\n{species}\n")
        cols = [a.transpose(2, 1, 0).reshape(-1) for a in (xc, yc, zc, rho, u, v, w, p)]
        zero = np.zeros(n)
        one = np.ones(n)
        data = np.stack(cols + [zero, zero, one], axis=1)
        np.savetxt(f, data, fmt="%.17g")


def _surfaces_box(ni, nj, nk, types):
    """types: dict surface-name -> (bctype, tag) for imin..kmax."""
    lines = []
    for name, (a, b, c, d, e, f) in (
        ("imin", (0, 0, 0, nj, 0, nk)), ("imax", (ni, ni, 0, nj, 0, nk)),
        ("jmin", (0, ni, 0, 0, 0, nk)), ("jmax", (0, ni, nj, nj, 0, nk)),
        ("kmin", (0, ni, 0, nj, 0, 0)), ("kmax", (0, ni, 0, nj, nk, nk)),
    ):
        bc, tag = types[name]
        lines.append(f"   {bc:<18} {a:5d} {b:5d} {c:5d} {d:5d} {e:5d} {f:5d} {tag:5d}")
    return "2 2 2\n#i_surfaces\n" + "\n".join(lines[0:2]) + "\n#j_surfaces\n" + \
        "\n".join(lines[2:4]) + "\n#k_surfaces\n" + "\n".join(lines[4:6]) + "\n"


def vortex_state(xc, yc, beta=5.0, rc=1.0, x0=5.0, y0=5.0, t_inf=288.15, rho_inf=1.225, mach=0.5):
    """Isentropic vortex on a freestream along x (SI units)."""
    gamma = (AIR_N + 1.0) / AIR_N
    R = air_gas_constant()
    p_inf = rho_inf * R * t_inf
    a_inf = math.sqrt(gamma * R * t_inf)
    u_inf = mach * a_inf
    r2 = ((xc - x0) ** 2 + (yc - y0) ** 2) / (rc * rc)
    tr = 1.0 - (gamma - 1.0) * beta ** 2 / (8.0 * gamma * math.pi ** 2) * np.exp(1.0 - r2)
    rho = rho_inf * tr ** (1.0 / (gamma - 1.0))
    p = p_inf * tr ** (gamma / (gamma - 1.0))
    scale = beta / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2)) * math.sqrt(p_inf / rho_inf)
    u = u_inf - scale * (yc - y0) / rc
    v = scale * (xc - x0) / rc
    return rho, u, v, np.zeros_like(u), p, dict(p_inf=p_inf, rho_inf=rho_inf, u_inf=u_inf, t_inf=t_inf)


def vortex_case(outdir, n=64, iterations=200, name="vortex", length=10.0, cfl=5.0, limiter="vanAlbada",
                recon="thirdOrder", solver="lusgs", sweeps=1, flux="roe", ni=None, nj=None, nk=None):
    """Write the isentropic-vortex case; returns the .inp path."""
    os.makedirs(outdir, exist_ok=True)
    ni = ni or n
    nj = nj or n
    nk = nk or n
    h = length / n
    i, j, k = np.meshgrid(np.arange(ni + 1), np.arange(nj + 1), np.arange(nk + 1), indexing="ij")
    x, y, z = i * h, j * h, k * h
    write_plot3d(os.path.join(outdir, name + ".xyz"), [(x, y, z)])
    xc, yc, zc = cell_centres(x, y, z)
    c = 0.5 * length
    rho, u, v, w, p, fs = vortex_state(xc, yc, x0=c, y0=c)
    write_cloud(os.path.join(outdir, name + ".cloud"), xc, yc, zc, rho, u, v, w, p)
    bcs = _surfaces_box(ni, nj, nk, {s: ("characteristic", 1) for s in
                                     ("imin", "imax", "jmin", "jmax", "kmin", "kmax")})
    inp = f"""# isentropic vortex, {ni}x{nj}x{nk} cells on [0,{length}]^3
gridName: {name}
equationSet: euler
timeIntegration: implicitEuler
cflStart: {cfl}
cflMax: {cfl}
faceReconstruction: {recon}
limiter: {limiter}
inviscidFlux: {flux}
iterations: {iterations}
outputFrequency: {iterations}
restartFrequency: 0
referenceTemperature: {fs['t_inf']!r}
referenceDensity: {fs['rho_inf']!r}
referenceLength: 1.0
initialConditions: <icState(tag=-1; file={name}.cloud)>
matrixSolver: {solver}
matrixSweeps: {sweeps}
matrixRelaxation: 1.0
boundaryStates: <characteristic(tag=1; pressure={fs['p_inf']!r}; density={fs['rho_inf']!r}; velocity=[{fs['u_inf']!r}, 0.0, 0.0])>
boundaryConditions: 1
{bcs}"""
    path = os.path.join(outdir, name + ".inp")
    with open(path, "w") as f:
        f.write(inp)
    return path


def bump_case(outdir, seed, n=None, name=None, iterations=100, cfl=5.0, recon="thirdOrder", viscous=False,
              visc_recon="central", limiter="vanAlbada"):
    """3D channel with a sin^2 bump on the lower wall; parameters seeded."""
    rng = np.random.default_rng(seed)
    n = n or int(rng.integers(24, 97))
    hl = float(rng.uniform(0.02, 0.10))
    mach = float(rng.uniform(0.2, 0.8))
    name = name or f"bump{seed}"
    os.makedirs(outdir, exist_ok=True)
    L, H, W = 3.0, 1.0, 1.0
    ell = 1.0
    x0 = 1.0
    i, j, k = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    x = L * i / n
    xi = np.clip((x - x0) / ell, 0.0, 1.0)
    yw = hl * ell * np.sin(np.pi * xi) ** 2
    y = yw + (H - yw) * j / n
    z = W * k / n
    scale = 2e-4 if viscous else 1.0  # viscous: a 0.6 mm channel, Reynolds number of order 1e3
    write_plot3d(os.path.join(outdir, name + ".xyz"), [(scale * x, scale * y, scale * z)])
    gamma = (AIR_N + 1.0) / AIR_N
    R = air_gas_constant()
    t_inf, rho_inf = 288.15, 1.225
    p_inf = rho_inf * R * t_inf
    u_inf = mach * math.sqrt(gamma * R * t_inf)
    walls = ("viscousWall", 3), ("viscousWall", 4)
    bcs = _surfaces_box(n, n, n, {"imin": ("characteristic", 1), "imax": ("characteristic", 1),
                                  "jmin": walls[0] if viscous else ("slipWall", 2),
                                  "jmax": walls[1] if viscous else ("slipWall", 2),
                                  "kmin": ("slipWall", 2), "kmax": ("slipWall", 2)})
    # viscous: an adiabatic wall on the bump, an isothermal wall moving along
    # x on the top; laminar at a low Reynolds number (small channel)
    eq = "navierStokes" if viscous else "euler"
    extra_states = (f""", viscousWall(tag=3), viscousWall(tag=4; temperature=300.0; velocity=[{0.3 * u_inf!r}, 0.0, 0.0])"""
                    if viscous else "")

    inp = f"""# bump channel seed {seed}: n={n}, h/l={hl:.4f}, Mach={mach:.4f}
gridName: {name}
equationSet: {eq}
viscousFaceReconstruction: {visc_recon}
timeIntegration: implicitEuler
cflStart: {cfl}
cflMax: {cfl}
faceReconstruction: {recon}
limiter: {limiter}
iterations: {iterations}
referenceTemperature: {t_inf!r}
referenceDensity: {rho_inf!r}
initialConditions: <icState(tag=-1; pressure={p_inf!r}; density={rho_inf!r}; velocity=[{u_inf!r}, 0.0, 0.0])>
matrixSolver: lusgs
matrixSweeps: 1
matrixRelaxation: 1.0
boundaryStates: <characteristic(tag=1; pressure={p_inf!r}; density={rho_inf!r}; velocity=[{u_inf!r}, 0.0, 0.0]){extra_states}>
boundaryConditions: 1
{bcs}"""
    path = os.path.join(outdir, name + ".inp")
    with open(path, "w") as f:
        f.write(inp)
    return path, dict(n=n, h_over_l=hl, mach=mach)


def channel_case(outdir, name="channel", refine=1.0, steps=50, subits=20, dt=1.0e-5, seed=0,
                 solver="blusgs", sweeps=1):
    """Batched multiblock laminar channel: 8 point-matched blocks in a
    4 (streamwise) x 2 (wall-normal) layout, columns of 96, 96, 96 and 48
    cells, a 64-cell wall row under a 32-cell row, 32 cells spanwise
    (1,032,192 cells at refine = 1; every count is scaled by `refine`).
    Wall-normal geometric stretching 1.1 from a first cell of 1e-5 L, a 4 L
    long, 0.5 L wide channel; no-slip adiabatic bottom wall, characteristic
    top, inflow and outflow, spanwise periodic.  Mach 0.3, Re_L = 1e6 at
    288.15 K with Sutherland air (the density follows from Re); thermal
    conductivity refit so that k / mu = c_p / 0.72 (a fluid file beside the
    input; the spectral radius keeps the reference's 4 gamma / (9 gamma - 5)).
    AUSMPW+ with MUSCL (kappa = 1/3, van Albada), central viscous flux,
    BLU-SGS, BDF2 with `subits` nonlinear iterations per physical step of
    `dt` seconds; uniform free stream plus seeded 1e-3 random velocity
    perturbations.  Returns the input path and a dict of the sizes."""
    rng = np.random.default_rng(seed)
    os.makedirs(outdir, exist_ok=True)
    cols = [max(2, int(round(c * refine))) for c in (96, 96, 96, 48)]
    rows = [max(2, int(round(r * refine))) for r in (64, 32)]
    nk = max(2, int(round(32 * refine)))
    Lx, Lz, dy1, ratio = 4.0, 0.5, 1.0e-5, 1.1
    nx = sum(cols)
    xs = Lx * np.arange(nx + 1) / nx
    ys = np.concatenate([[0.0], np.cumsum(dy1 * ratio ** np.arange(sum(rows)))])
    zs = Lz * np.arange(nk + 1) / nk
    R = air_gas_constant()
    gamma = (AIR_N + 1.0) / AIR_N
    t_inf = 288.15
    a_inf = math.sqrt(gamma * R * t_inf)
    u_inf = 0.3 * a_inf
    mu_inf = 1.458e-6 * t_inf ** 1.5 / (t_inf + 110.4)
    rho_inf = 1.0e6 * mu_inf / (u_inf * 1.0)
    p_inf = rho_inf * R * t_inf
    cp = R * (AIR_N + 1.0)
    with open(os.path.join(outdir, "air72.dat"), "w") as f:
        f.write(f"""n: {AIR_N}
molarMass: {AIR_MOLAR * 1000.0!r}
sutherlandViscosityC1: 1.458e-6
sutherlandViscosityS: 110.4
sutherlandConductivityC1: {1.458e-6 * cp / 0.72!r}
sutherlandConductivityS: 110.4
""")
    blocks, bcs, cents = [], [], []
    x0s = np.concatenate([[0], np.cumsum(cols)])
    y0s = np.concatenate([[0], np.cumsum(rows)])

    def bid(c, r):
        return c + 4 * r

    for r in range(2):
        for c in range(4):
            ii = np.arange(x0s[c], x0s[c + 1] + 1)
            jj = np.arange(y0s[r], y0s[r + 1] + 1)
            X, Y, Z = np.meshgrid(xs[ii], ys[jj], zs, indexing="ij")
            blocks.append((X, Y, Z))
            cents.append(cell_centres(X, Y, Z))
            types = {
                "imin": ("characteristic", 1) if c == 0 else ("interblock", 2000 + bid(c - 1, r)),
                "imax": ("characteristic", 1) if c == 3 else ("interblock", 1000 + bid(c + 1, r)),
                "jmin": ("viscousWall", 3) if r == 0 else ("interblock", 4000 + bid(c, 0)),
                "jmax": ("characteristic", 1) if r == 1 else ("interblock", 3000 + bid(c, 1)),
                "kmin": ("periodic", 7), "kmax": ("periodic", 8),
            }
            bcs.append(_surfaces_box(cols[c], rows[r], nk, types))
    write_plot3d(os.path.join(outdir, name + ".xyz"), blocks)
    # cloud: every block's cell centres, block order, i fastest
    xc = np.concatenate([cc[0].transpose(2, 1, 0).reshape(-1) for cc in cents])
    yc = np.concatenate([cc[1].transpose(2, 1, 0).reshape(-1) for cc in cents])
    zc = np.concatenate([cc[2].transpose(2, 1, 0).reshape(-1) for cc in cents])
    n = xc.size
    pert = 1.0e-3 * u_inf * rng.uniform(-1.0, 1.0, (3, n))
    with open(os.path.join(outdir, name + ".cloud"), "w") as f:
        f.write(f"{n}\nair72\n")
        data = np.stack([xc, yc, zc, np.full(n, rho_inf), u_inf + pert[0], pert[1], pert[2], np.full(n, p_inf),
                         np.zeros(n), np.zeros(n), np.ones(n)], axis=1)
        np.savetxt(f, data, fmt="%.17g")
    inp = f"""# multiblock laminar channel, refine {refine}, seed {seed}
gridName: {name}
equationSet: navierStokes
timeIntegration: bdf2
timeStep: {dt!r}
nonlinearIterations: {subits}
iterations: {steps}
faceReconstruction: thirdOrder
limiter: vanAlbada
inviscidFlux: ausm
viscousFaceReconstruction: central
referenceTemperature: {t_inf!r}
referenceDensity: {rho_inf!r}
referenceLength: 1.0
fluids: <fluid(name=air72; referenceMassFraction=1.0)>
initialConditions: <icState(tag=-1; file={name}.cloud)>
matrixSolver: {solver}
matrixSweeps: {sweeps}
matrixRelaxation: 1.0
boundaryStates: <characteristic(tag=1; pressure={p_inf!r}; density={rho_inf!r}; velocity=[{u_inf!r}, 0.0, 0.0]), viscousWall(tag=3), periodic(startTag=7; endTag=8; translation=[0.0, 0.0, {Lz!r}])>
boundaryConditions: 8
""" + "".join(bcs)
    path = os.path.join(outdir, name + ".inp")
    with open(path, "w") as f:
        f.write(inp)
    cells = sum(cc * rr * nk for cc in cols for rr in rows)
    return path, dict(cells=cells, blocks=8, cols=cols, rows=rows, nk=nk, steps=steps, subits=subits)
