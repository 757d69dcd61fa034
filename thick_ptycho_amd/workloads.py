"""Seeded synthetic inputs for the benchmark and parity workloads.

Objects follow the reference's sign convention n = 1 - sum_b (delta_b + i beta_b) G_b
(absorbing under the PWE sign; generate_data.py:76-78 builds objects the same
way from shapes).  Probes are Gaussians sampled on the pixel grid, with the
two edge samples set to zero so that Dirichlet rows hold zero (the reference
carries boundary rows by (1+C)/(1-C), boundary_conditions.py:57-66).

Physical constants shared by every workload: wavelength 1 (k = 2 pi),
dx = dy = 2, dz = 4, n_medium = 1.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

WAVELENGTH = 1.0
K = 2 * np.pi / WAVELENGTH
DX = 2.0
DZ = 4.0


@dataclass
class Scan2D:
    """One 2D scan: object (nx, nz), probes (P, nx), geometry."""

    nx: int
    nz: int
    n: np.ndarray
    probes: np.ndarray
    bc: str = "dirichlet"
    dx: float = DX
    dz: float = DZ
    k: float = K
    n_medium: float = 1.0
    name: str = ""

    @property
    def steps(self) -> int:
        return self.probes.shape[0] * (self.nz - 1)


@dataclass
class Scan3D:
    """One 3D scan: object (nx, ny, nz), probes (P, nx, ny), geometry."""

    nx: int
    ny: int
    nz: int
    n: np.ndarray
    probes: np.ndarray
    bc: str = "dirichlet"
    dx: float = DX
    dy: float = DX
    dz: float = DZ
    k: float = K
    name: str = ""

    @property
    def steps(self) -> int:
        return self.probes.shape[0] * (self.nz - 1)


@dataclass
class Workload:
    name: str
    scans: list = field(default_factory=list)
    extra: dict = field(default_factory=dict)

    @property
    def steps(self) -> int:
        return sum(s.steps for s in self.scans)


def blob_object_2d(nx: int, nz: int, nblobs: int, rng: np.random.Generator,
                   sigma=(5.0, 20.0), delta=(0.0, 1e-3), beta=(0.0, 1e-5)) -> np.ndarray:
    """n(x, z) = 1 - sum_b (delta_b + i beta_b) exp(-r^2 / (2 sigma_b^2))."""
    x = np.arange(nx)[:, None]
    z = np.arange(nz)[None, :]
    dn = np.zeros((nx, nz), dtype=np.complex128)
    for _ in range(nblobs):
        cx = rng.uniform(0, nx)
        cz = rng.uniform(0, nz)
        s = rng.uniform(*sigma)
        d = rng.uniform(*delta)
        b = rng.uniform(*beta)
        dn += (d + 1j * b) * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * s * s))
    return 1.0 - dn


def ellipsoid_object_3d(nx: int, ny: int, nz: int, nblobs: int, rng: np.random.Generator,
                        semi=(3.0, 12.0), delta=(0.0, 1e-3), beta=(0.0, 1e-5)) -> np.ndarray:
    """n(x, y, z) = 1 - sum_b (delta_b + i beta_b) exp(-sum_i (r_i / a_i)^2 / 2)."""
    x = np.arange(nx)[:, None, None]
    y = np.arange(ny)[None, :, None]
    z = np.arange(nz)[None, None, :]
    dn = np.zeros((nx, ny, nz), dtype=np.complex128)
    for _ in range(nblobs):
        c = [rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz)]
        a = rng.uniform(*semi, size=3)
        d = rng.uniform(*delta)
        b = rng.uniform(*beta)
        r2 = ((x - c[0]) / a[0]) ** 2 + ((y - c[1]) / a[1]) ** 2 + ((z - c[2]) / a[2]) ** 2
        dn += (d + 1j * b) * np.exp(-0.5 * r2)
    return 1.0 - dn


def gaussian_probes_1d(nx: int, centres, waist: float) -> np.ndarray:
    x = np.arange(nx)[None, :]
    c = np.asarray(centres, dtype=float)[:, None]
    p = np.exp(-((x - c) ** 2) / waist**2).astype(np.complex128)
    p[:, 0] = 0
    p[:, -1] = 0
    return p


def gaussian_probes_2d(nx: int, ny: int, centres, waist: float) -> np.ndarray:
    x = np.arange(nx)[None, :, None]
    y = np.arange(ny)[None, None, :]
    c = np.asarray(centres, dtype=float)
    cx = c[:, 0][:, None, None]
    cy = c[:, 1][:, None, None]
    p = np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / waist**2).astype(np.complex128)
    p[:, 0, :] = 0
    p[:, -1, :] = 0
    p[:, :, 0] = 0
    p[:, :, -1] = 0
    return p


def scan_2d(nx: int, nz: int, n_probes: int, centres, seed: int, bc: str = "dirichlet", name: str = "") -> Scan2D:
    rng = np.random.default_rng(seed)
    n = blob_object_2d(nx, nz, 8, rng)
    waist = 24.0 * nx / 256.0
    return Scan2D(nx=nx, nz=nz, n=n, probes=gaussian_probes_1d(nx, centres, waist), bc=bc, name=name)


def config0() -> Workload:
    """Plain 2D forward: Nx = Nz = 256, 16 probes at 8 + 16 s, waist 24 px, seed 0."""
    s = scan_2d(256, 256, 16, [8 + 16 * i for i in range(16)], seed=0, name="cfg0")
    return Workload("2d-256x256-16probes", [s])


def config1(nz: int = 1024, main_probes: int = 512, extra_probes: int = 128) -> Workload:
    """Batched 2D: Nx = 1024 with 512 probes plus Nx in {384, 640, 896} with 128 each."""
    scans = [scan_2d(1024, nz, main_probes, [2 * i + 1 for i in range(main_probes)], seed=10, name="main")]
    for i, nx in enumerate((384, 640, 896)):
        cs = [int(round((s + 0.5) * nx / extra_probes)) for s in range(extra_probes)]
        scans.append(scan_2d(nx, nz, extra_probes, cs, seed=11 + i, name=f"extra{nx}"))
    return Workload("2d-batched-1024x1024x512+384/640/896x128", scans)


def config2(nz: int = 512, probes: int = 256) -> Workload:
    """2D reconstruction gradient: Nx = 1024, Nz = 512, 256 probes at 2 + 4 s.

    ``extra`` holds the true object, the object the measured intensities are
    simulated from (delta-n perturbed by 5 % Gaussian noise, seed 21) and the
    free-space object.  The scan's object, where the gradient is evaluated,
    is a seeded initial guess: 8 blobs from seed 22 with the config-0
    generator.  (At free space the amplitude-misfit gradient is dominated by
    detector bins where |F u_L| is ~1e-6 of its peak while the data are not,
    so its phase there is set by complex64 rounding of u_L.)"""
    nx = 1024
    true = scan_2d(nx, nz, probes, [2 + 4 * i for i in range(probes)], seed=20, name="true")
    rng = np.random.default_rng(21)
    dn = 1.0 - true.n
    n_meas = 1.0 - dn * (1.0 + 0.05 * rng.standard_normal(dn.shape))
    guess = blob_object_2d(nx, nz, 8, np.random.default_rng(22))
    s = Scan2D(nx=nx, nz=nz, n=guess, probes=true.probes, bc=true.bc, name="grad")
    return Workload("2d-gradient-1024x512-256probes", [s],
                    extra={"n_true": true.n, "n_meas_object": n_meas, "n_free": np.ones_like(true.n)})


def raster_centres(n_side: int, step: float, centre: float) -> np.ndarray:
    c = centre + step * (np.arange(n_side) - (n_side - 1) / 2.0)
    X, Y = np.meshgrid(c, c, indexing="ij")
    return np.stack([X.ravel(), Y.ravel()], axis=1)


def config3(nz: int = 256) -> Workload:
    """3D forward: 128 x 128 x 256, 64 probes on an 8 x 8 raster (step 13, waist 16)."""
    rng = np.random.default_rng(30)
    n = ellipsoid_object_3d(128, 128, nz, 32, rng)
    probes = gaussian_probes_2d(128, 128, raster_centres(8, 13.0, 64.0), 16.0)
    return Workload("3d-128x128x256-64probes", [Scan3D(128, 128, nz, n, probes, name="cfg3")])


def config4(nz: int = 128) -> Workload:
    """3D reconstruction: 256 x 256 x 128, 400 probes on a 20 x 20 raster (60 % overlap)."""
    rng = np.random.default_rng(40)
    n = ellipsoid_object_3d(256, 256, nz, 32, rng)
    waist = 16.0
    step = 0.4 * 2 * waist
    probes = gaussian_probes_2d(256, 256, raster_centres(20, step, 128.0), waist)
    return Workload("3d-256x256x128-400probes", [Scan3D(256, 256, nz, n, probes, name="cfg4")])
