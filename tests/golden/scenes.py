"""Small scenes shared by the golden-fixture writer and the parity tests.

Each builder takes the photonfdtd-compatible module ``pf`` (the reference
package when fixtures are written, ``photonfdtd_amd`` when they are checked)
and a ``backend`` string, and returns a ``Simulation``. Sizes are chosen so a
float64 NumPy run finishes in well under a second.
"""
import numpy as np

LAM = 1.0e-6


def _f0(pf):
    return pf.C_0 / LAM


def tm2d_box(pf, backend, precision="float64"):
    """2-D TM dipole beside a dielectric box, CPML on x/y, fields + DFT."""
    grid = pf.Grid(size=(2.0e-6, 1.6e-6), cell_size=50e-9, pml_layers=(8, 8, 0))
    box = pf.Box(center=(0.3e-6, 0.0), size=(0.6e-6, 0.5e-6), medium=pf.Medium.from_index(2.0))
    src = pf.PointDipole(position=(-0.4e-6, 0.05e-6), component="Ez",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=6e-15))
    mon = pf.FieldMonitor(name="s", components=("Ez", "Hx", "Hy"), interval=25)
    dft = pf.DFTMonitor(name="d", components=("Ez", "Hy"), freqs=[0.9 * _f0(pf), _f0(pf)])
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[mon, dft],
                         run_time=25e-15, backend=backend, precision=precision)


def te2d_poly(pf, backend, precision="float64"):
    """2-D TE (Ex, Ey, Hz) with a PolySlab scatterer and a y-plane DFT."""
    grid = pf.Grid(size=(1.8e-6, 1.8e-6), cell_size=(40e-9, 40e-9), pml_layers=(8, 8, 0))
    verts = ((-0.3e-6, -0.2e-6), (0.35e-6, -0.25e-6), (0.2e-6, 0.3e-6), (-0.25e-6, 0.15e-6))
    poly = pf.PolySlab(vertices=verts, z_bounds=(-1e-6, 1e-6), medium=pf.Medium(6.0))
    src = pf.PointDipole(position=(-0.5e-6, -0.5e-6), component="Ex",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=5e-15))
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Hz"), times=[10e-15, 18e-15])
    dft = pf.DFTMonitor(name="d", components=("Ex", "Hz"), freqs=[_f0(pf)],
                        plane_axis="y", plane_position=0.4e-6)
    return pf.Simulation(grid, structures=[poly], sources=[src], monitors=[mon, dft],
                         run_time=20e-15, backend=backend, precision=precision)


def box3d(pf, backend, precision="float64"):
    """3-D dipole with CPML on all faces; full snapshots and an x-plane DFT."""
    grid = pf.Grid(size=(1.6e-6, 1.28e-6, 1.12e-6), cell_size=80e-9, pml_layers=(5, 5, 5))
    box = pf.Box(center=(0.2e-6, 0, 0), size=(0.5e-6, 0.5e-6, 0.5e-6),
                 medium=pf.Medium.from_index(1.8))
    src = pf.PointDipole(position=(-0.3e-6, 0, 0), component="Ey",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=5e-15))
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"),
                          times=[9e-15])
    dft = pf.DFTMonitor(name="d", components=("Ey", "Ez", "Hy", "Hz"),
                        freqs=[0.8 * _f0(pf), _f0(pf), 1.2 * _f0(pf)],
                        plane_axis="x", plane_position=0.45e-6)
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[mon, dft],
                         run_time=10e-15, backend=backend, precision=precision)


def lorentz1d(pf, backend, precision="float64"):
    """1-D pulse into a two-pole Lorentz slab; every-node DFT."""
    med = pf.DispersiveMedium.lorentz(2.25, [(1.2, 500e12, 10e12), (0.3, 300e12, 0.0)])
    grid = pf.Grid(size=(6e-6,), cell_size=50e-9, pml_layers=(12,))
    box = pf.Box(center=(1.0e-6,), size=(2.0e-6,), medium=med)
    src = pf.PointDipole(position=(-1.5e-6,), component="Ez",
                         waveform=pf.GaussianPulse(freq0=200e12, fwhm=8e-15))
    mon = pf.FieldMonitor(name="s", components=("Ez", "Hy"), interval=40)
    dft = pf.DFTMonitor(name="d", components=("Ez",), freqs=[150e12, 210e12, 270e12])
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[mon, dft],
                         run_time=60e-15, backend=backend, precision=precision)


def metals_xz(pf, backend, precision="float64"):
    """2-D x-z grid: Drude, 4-pole Lorentz-Drude gold and 3-pole silicon regions, an
    overriding plain box, and two dipoles on one node."""
    grid = pf.Grid(size=(1.0e-6, 0, 0.8e-6), cell_size=20e-9, pml_layers=(8, 0, 8))
    drude = pf.DispersiveMedium.drude(1.0, f_plasma_Hz=150e12, f_collision_Hz=5e12)
    gold = pf.DispersiveMedium(eps_inf=1.0, poles=pf.gold().poles[:4], name="Au LD rows 0-3")
    boxes = [
        pf.Box(center=(-0.2e-6, 0, 0.15e-6), size=(0.2e-6, 1, 0.15e-6), medium=drude),
        pf.Box(center=(0.2e-6, 0, -0.1e-6), size=(0.15e-6, 1, 0.2e-6), medium=gold),
        pf.Box(center=(0.0, 0, -0.25e-6), size=(0.6e-6, 1, 0.06e-6), medium=pf.silicon()),
        pf.Box(center=(0.22e-6, 0, -0.12e-6), size=(0.05e-6, 1, 0.05e-6), medium=pf.Medium(3.0)),
    ]
    w = pf.GaussianPulse(freq0=250e12, fwhm=6e-15)
    srcs = [pf.PointDipole(position=(0.0, 0, 0.05e-6), component="Ey", waveform=w),
            pf.PointDipole(position=(0.0, 0, 0.05e-6), component="Ey", waveform=w, amplitude=0.5),
            pf.PointDipole(position=(-0.1e-6, 0, -0.05e-6), component="Hy", waveform=w,
                           amplitude=2e-3)]
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"), interval=60)
    dft = pf.DFTMonitor(name="d", components=("Ey", "Hx"), freqs=[200e12, 260e12], downsample=2)
    return pf.Simulation(grid, structures=boxes, sources=srcs, monitors=[mon, dft],
                         run_time=20e-15, backend=backend, precision=precision)


def mode3d(pf, backend, precision="float64"):
    """3-D ModeSource with a 2-D Gaussian profile, plane_z / downsampled
    snapshots and a sub-sampled DFT."""
    grid = pf.Grid(size=(1.2e-6, 0.96e-6, 0.8e-6), cell_size=80e-9, pml_layers=(4, 4, 4))
    core = pf.Box(center=(0, 0, 0), size=(2e-6, 0.32e-6, 0.24e-6), medium=pf.Medium(12.0))
    yy = np.linspace(-0.4e-6, 0.4e-6, 21)
    zz = np.linspace(-0.3e-6, 0.3e-6, 13)
    prof = np.exp(-(yy[:, None] / 0.2e-6) ** 2 - (zz[None, :] / 0.15e-6) ** 2)
    src = pf.ModeSource(center=(-0.3e-6, 0, 0), size=(0, 0.8e-6, 0.6e-6), component="Ey",
                        waveform=pf.GaussianPulse(freq0=_f0(pf) * 1.2, fwhm=5e-15),
                        profile=prof, profile_coords=(yy, zz), amplitude=2.0)
    mz = pf.FieldMonitor(name="z", components=("Ey", "Hz"), interval=15, plane_z=0.05e-6)
    md = pf.FieldMonitor(name="ds", components=("Ey",), interval=30, downsample=2)
    dft = pf.DFTMonitor(name="d", components=("Ey", "Hz"), freqs=[_f0(pf), 1.2 * _f0(pf)],
                        interval=3, plane_axis="x", plane_position=0.3e-6)
    return pf.Simulation(grid, structures=[core], sources=[src], monitors=[mz, md, dft],
                         run_time=9e-15, backend=backend, precision=precision)


def flux2d(pf, backend, precision="float64"):
    """2-D TM run with time-domain flux planes on x and y."""
    grid = pf.Grid(size=(1.6e-6, 1.6e-6), cell_size=40e-9, pml_layers=(8, 8, 0))
    src = pf.PointDipole(position=(0, 0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=5e-15))
    fx = pf.FluxMonitor(name="fx", plane_axis="x", plane_position=0.4e-6)
    fy = pf.FluxMonitor(name="fy", plane_axis="y", plane_position=-0.4e-6)
    return pf.Simulation(grid, sources=[src], monitors=[fx, fy], run_time=20e-15,
                         backend=backend, precision=precision)


def subpixel2d(pf, backend, precision="float64"):
    """2-D TE (Ex, Ey, Hz) with subpixel smoothing (factor 3): an oblique
    triangular PolySlab and a box whose walls fall between nodes."""
    grid = pf.Grid(size=(1.8e-6, 1.6e-6), cell_size=(40e-9, 40e-9), pml_layers=(8, 8, 0))
    tri = pf.PolySlab(vertices=((-0.35e-6, -0.3e-6), (0.3e-6, -0.17e-6), (-0.1e-6, 0.33e-6)),
                      z_bounds=(-1e-6, 1e-6), medium=pf.Medium(7.0))
    box = pf.Box(center=(0.41e-6, 0.23e-6), size=(0.27e-6, 0.33e-6), medium=pf.Medium(3.2))
    src = pf.PointDipole(position=(-0.5e-6, 0.4e-6), component="Ex",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=5e-15))
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Hz"), interval=30)
    dft = pf.DFTMonitor(name="d", components=("Ey", "Hz"), freqs=[0.8 * _f0(pf), _f0(pf)])
    return pf.Simulation(grid, structures=[tri, box], sources=[src], monitors=[mon, dft],
                         run_time=18e-15, backend=backend, precision=precision, subpixel=True,
                         subpixel_factor=3)


def subpixel3d(pf, backend, precision="float64"):
    """3-D with subpixel smoothing (factor 2): a box and a z-bounded PolySlab
    with faces between nodes on all three axes; Ez dipole, x-plane DFT."""
    grid = pf.Grid(size=(1.0e-6, 0.88e-6, 0.8e-6), cell_size=40e-9, pml_layers=(5, 5, 5))
    box = pf.Box(center=(0.13e-6, -0.05e-6, 0.07e-6), size=(0.31e-6, 0.23e-6, 0.27e-6),
                 medium=pf.Medium(5.5))
    hexa = pf.PolySlab(vertices=((-0.3e-6, 0.1e-6), (-0.1e-6, 0.05e-6), (0.05e-6, 0.25e-6),
                                 (-0.15e-6, 0.3e-6)), z_bounds=(-0.13e-6, 0.11e-6),
                       medium=pf.Medium(2.7))
    src = pf.PointDipole(position=(-0.2e-6, -0.2e-6, 0.0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=4e-15))
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Ez", "Hx"), times=[6e-15, 11e-15])
    dft = pf.DFTMonitor(name="d", components=("Ey", "Ez", "Hy", "Hz"), freqs=[_f0(pf)],
                        plane_axis="x", plane_position=0.2e-6)
    return pf.Simulation(grid, structures=[box, hexa], sources=[src], monitors=[mon, dft],
                         run_time=12e-15, backend=backend, precision=precision, subpixel=True,
                         subpixel_factor=2)


def mode_strip(pf):
    """Silicon strip (500 x 220 nm) in silica, 1.55 um, 20 nm cells: the two
    highest-index modes of the full-vectorial solver."""
    core = pf.Box(center=(0.0, 0.0), size=(0.5e-6, 0.22e-6), medium=pf.Medium(12.1))
    return pf.ModeSolver(size=(1.2e-6, 0.8e-6), cell_size=20e-9, structures=[core],
                         wavelength=1.55e-6, background_eps=2.085, num_modes=2)


def mode_slab_tm(pf):
    """Asymmetric slab (silicon nitride on silica, air above), TM filter,
    (dy, dz) = (30, 15) nm, dispersive core evaluated at 1.31 um."""
    sub = pf.Box(center=(0.0, -0.35e-6), size=(2.0e-6, 0.6e-6), medium=pf.Medium(2.085))
    core = pf.Box(center=(0.0, 0.0), size=(2.0e-6, 0.1e-6), medium=pf.silicon_nitride())
    return pf.ModeSolver(size=(0.9e-6, 0.9e-6), cell_size=(30e-9, 15e-9), structures=[sub, core],
                         wavelength=1.31e-6, num_modes=1, polarization="tm")


MODE_CASES = {"mode_strip": mode_strip, "mode_slab_tm": mode_slab_tm}


def unimode3d(pf, backend, precision="float64"):
    """3-D silicon strip along x; a solved TE-like mode launched toward +x by a
    UniModeSource (electric + magnetic sheets); DFT port planes before and
    after the source with the tangential (Ey, Ez, Hy, Hz) fields."""
    d = 40e-9
    size = (1.6e-6, 1.2e-6, 0.96e-6)
    grid = pf.Grid(size=size, cell_size=d, pml_layers=(6, 6, 6))
    core = pf.Box(center=(0, 0, 0), size=(10e-6, 0.44e-6, 0.24e-6), medium=pf.Medium(12.1))
    mode = pf.ModeSolver(size=size[1:], cell_size=d, structures=[core], wavelength=1.55e-6,
                         background_eps=1.0, num_modes=1).solve()
    f0 = pf.C_0 / 1.55e-6
    src = pf.UniModeSource(center=(-0.4e-6, 0, 0), size=(0, size[1], size[2]),
                           waveform=pf.GaussianPulse(freq0=f0, fwhm=5e-15), mode=mode)
    ports = [pf.DFTMonitor(name=nm, components=("Ey", "Ez", "Hy", "Hz"),
                           freqs=[0.95 * f0, f0], plane_axis="x", plane_position=x)
             for nm, x in (("back", -0.56e-6), ("fwd", 0.44e-6))]
    return pf.Simulation(grid, structures=[core], sources=[src], monitors=ports,
                         run_time=14e-15, backend=backend, precision=precision)


def compressed2d(pf, backend, precision="float64"):
    """2-D TM dipole beside a box with compressed FieldMonitors: int8 frames
    (zlib) every 15 steps and float16 frames of every 2nd node at set times."""
    grid = pf.Grid(size=(2.0e-6, 1.6e-6), cell_size=50e-9, pml_layers=(8, 8, 0))
    box = pf.Box(center=(0.3e-6, 0.0), size=(0.6e-6, 0.5e-6), medium=pf.Medium.from_index(2.0))
    src = pf.PointDipole(position=(-0.4e-6, 0.05e-6), component="Ez",
                         waveform=pf.GaussianPulse(freq0=_f0(pf), fwhm=6e-15))
    m8 = pf.FieldMonitor(name="q8", components=("Ez", "Hy"), interval=15, compression="zlib",
                         compression_bits=8)
    m16 = pf.FieldMonitor(name="q16", components=("Ez", "Hx"), times=[8e-15, 16e-15, 24e-15],
                          downsample=2, compression="zlib", compression_bits=16)
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[m8, m16],
                         run_time=25e-15, backend=backend, precision=precision)


SCENES = {
    "tm2d_box": tm2d_box,
    "te2d_poly": te2d_poly,
    "box3d": box3d,
    "lorentz1d": lorentz1d,
    "metals_xz": metals_xz,
    "mode3d": mode3d,
    "flux2d": flux2d,
    "subpixel2d": subpixel2d,
    "subpixel3d": subpixel3d,
}
# float32 trajectories pinned as well (the reference's NumPy float32 path)
F32_SCENES = ("tm2d_box", "box3d", "metals_xz")
# Scenes whose fixtures pin one feature (not full-trajectory oracle scenes):
# compressed FieldMonitor frames, a one-way mode source from a solved mode.
FEATURE_SCENES = {"compressed2d": compressed2d, "unimode3d": unimode3d}


def dispersive_slab(pf, backend, shape=(48, 72, 140), steps=300, precision="float64",
                    pml=12, courant=0.9):
    """The 3-D dispersive slab physics on a smaller grid: dx = 20 nm, CPML on
    all six faces, silica cladding (eps 2.085), the silicon Sellmeier box
    200 nm x 500 nm x 8 um (clipped by the domain in z), a 300 nm cube with
    gold's Drude pole (Rakic row 0), an Ey point dipole at 1.55 um (10 fs)
    and two x-plane DFT monitors (Ey, Ez, Hy, Hz) at 32 frequencies in
    [c/1.65 um, c/1.45 um]. Positions scale per axis with the domain, as on
    the 512^3 grid (source at -0.146 Lx, cube at +0.146 (Lx, Ly), planes at
    +-0.2 Lx). Courant 0.9: the explicit ADE of silicon and gold grows without
    bound at 0.99 on a 20 nm 3-D grid."""
    dx = 20e-9
    L = [n * dx for n in shape]
    f = 1.5e-6 / (512 * dx)
    gold_drude = pf.DispersiveMedium(eps_inf=1.0, poles=(pf.gold().poles[0],),
                                     name="Au Drude term (Rakic row 0)")
    structures = [
        pf.Box(center=(0, 0, 0), size=tuple(3 * v for v in L), medium=pf.Medium(2.085)),
        pf.Box(center=(0, 0, 0), size=(200e-9, 500e-9, 8e-6), medium=pf.silicon()),
        pf.Box(center=(f * L[0], f * L[1], 0), size=(300e-9,) * 3, medium=gold_drude),
    ]
    src = pf.PointDipole(position=(-f * L[0], 0, 0), component="Ey",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.55e-6, fwhm=10e-15))
    freqs = np.linspace(pf.C_0 / 1.65e-6, pf.C_0 / 1.45e-6, 32)
    mons = [pf.DFTMonitor(name=nm, components=("Ey", "Ez", "Hy", "Hz"), freqs=freqs,
                          plane_axis="x", plane_position=x)
            for nm, x in (("in", -0.2 * L[0]), ("out", 0.2 * L[0]))]
    grid = pf.Grid(size=tuple(L), cell_size=dx, pml_layers=(pml, pml, pml))
    dt = courant / (pf.C_0 * np.sqrt(3.0 / dx ** 2))
    return pf.Simulation(grid, structures=structures, sources=[src], monitors=mons,
                         run_time=(steps - 0.5) * dt, courant=courant, precision=precision,
                         backend=backend)
