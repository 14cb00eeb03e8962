"""Workload registry for the RCA delay-and-sum beamformer.

Each workload is expressed in the reference's own parameterisation — the
``Geometry`` / ``Transmit`` / ``Param`` / ``Options`` dicts of
``bfVirtualSourceGeneric_RCA`` (MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.m:24-57)
— as the experiment drivers build them (MATLAB/script_All.m:128-134, 177-270):

* the grid is a ``meshgrid(x, z, y)`` volume of shape [Nz, Nx, Ny], so voxel
  p = iz + Nz*(ix + Nx*iy) after a column-major flatten (script_All.m:198);
* an RCA frame is two halves: transmit on rows / receive on the columns
  (``flagReceiveX=True``, tilts in ``thetaY``) and transmit on columns /
  receive on the rows (``flagReceiveX=False``, tilts in ``thetaX``); the
  caller adds the two beamformed volumes (script_All.m:366-382).

Workloads (sizes in the order they are quoted):

``cfg0``  64+64 elements, pitch 0.2 mm, 11 row tilts -10..10 deg, 1024 samples,
          64x64x128 voxels at lambda/2, Lanczos a=3, 200 point scatterers.
``cfg1``  128+128 elements, 2 x 16 tilts -12..12 deg, 2048 samples,
          128x128x256 voxels, Lanczos a=4, 1000 scatterers + 20 dB noise.
``cfg2``  cfg1 geometry, 200-frame Doppler ensemble, FP16-I/Q storage.
``cfg4``  96+96 elements, 16 row tilts -10..10 deg, 1024 samples,
          96x96x192 voxels, Lanczos a=3, 8 frames (differentiable beamforming).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

C_SOUND = 1540.0
F0 = 7.5e6
FS = 15e6
PITCH = 0.2e-3
LAMBDA = C_SOUND / F0


@dataclass
class Half:
    """One RCA half: which elements receive, and the tilt angles (rad)."""
    receive_x: bool
    theta: np.ndarray  # tilt along the transmitting axis


@dataclass
class RcaConfig:
    name: str
    n_elem: int          # elements per axis (rows = columns = n_elem)
    n_sample: int        # I/Q samples per trace
    nx: int
    ny: int
    nz: int
    halves: list
    interpolation: str   # 'quadratic' | 'lanczos' (reference) | 'lanczos3' | 'lanczos4'
    n_frame: int = 1
    n_scat: int = 200
    scat_box: tuple = (6.4e-3, 6.4e-3, 3e-3, 23e-3)   # |x|,|y| half-widths, z range
    seeds: tuple = (1000, 1001, 1002)
    snr_db: float | None = None
    z0: float = 5e-3
    dxyz: float = LAMBDA / 2
    fs: float = FS
    fd: float = F0
    c: float = C_SOUND
    t0: float = 0.0
    fnumber: float = 1.0
    iq_dtype: str = "c64"
    extra: dict = field(default_factory=dict)

    @property
    def n_tilt(self) -> int:
        return sum(len(h.theta) for h in self.halves)

    @property
    def n_vox(self) -> int:
        return self.nx * self.ny * self.nz

    def updates_per_frame(self) -> int:
        """Dense voxel*tilt*element updates of one frame (all halves)."""
        return self.n_vox * self.n_elem * self.n_tilt

    def taps(self) -> int:
        return {"quadratic": 3, "lanczos": 6}.get(self.interpolation,
                                                   2 * int(self.interpolation[7:] or 3))


def _deg(a, b, n):
    return np.deg2rad(np.linspace(a, b, n))


CONFIGS = {
    "cfg0": RcaConfig("cfg0", 64, 1024, 64, 64, 128, [Half(True, _deg(-10, 10, 11))],
                      "lanczos3", n_scat=200, seeds=(1000, 1001, 1002)),
    "cfg1": RcaConfig("cfg1", 128, 2048, 128, 128, 256,
                      [Half(True, _deg(-12, 12, 16)), Half(False, _deg(-12, 12, 16))],
                      "lanczos4", n_scat=1000, scat_box=(6.5e-3, 6.5e-3, 5e-3, 31e-3),
                      seeds=(2000, 2001, 2002), snr_db=20.0),
    "cfg2": RcaConfig("cfg2", 128, 2048, 128, 128, 256,
                      [Half(True, _deg(-12, 12, 16)), Half(False, _deg(-12, 12, 16))],
                      "lanczos4", n_frame=200, n_scat=1000,
                      scat_box=(6.5e-3, 6.5e-3, 5e-3, 31e-3), seeds=(3000, 3001, 3002),
                      iq_dtype="f16", extra={"vmax": 5e-3, "frame_dt": 1e-3,
                                             "tissue_z": (10e-3, 11e-3), "tissue_db": 40.0,
                                             "n_tissue": 2000}),
    "cfg4": RcaConfig("cfg4", 96, 1024, 96, 96, 192, [Half(True, _deg(-10, 10, 16))],
                      "lanczos3", n_frame=8, n_scat=200, scat_box=(4.9e-3, 4.9e-3, 5e-3, 24e-3),
                      seeds=(4000, 4001, 4002)),
}


def element_positions(n_elem: int, pitch: float = PITCH) -> np.ndarray:
    return (np.arange(n_elem) - (n_elem - 1) / 2.0) * pitch


def grid_axes(cfg: RcaConfig):
    x = (np.arange(cfg.nx) - (cfg.nx - 1) / 2.0) * cfg.dxyz
    y = (np.arange(cfg.ny) - (cfg.ny - 1) / 2.0) * cfg.dxyz
    z = cfg.z0 + np.arange(cfg.nz) * cfg.dxyz
    return x, y, z


def meshgrid_zxy(x, y, z):
    """[grid.X, grid.Z, grid.Y] = meshgrid(x, z, y) (script_All.m:198): arrays
    of shape [Nz, Nx, Ny]."""
    X = np.broadcast_to(x[None, :, None], (z.size, x.size, y.size))
    Y = np.broadcast_to(y[None, None, :], (z.size, x.size, y.size))
    Z = np.broadcast_to(z[:, None, None], (z.size, x.size, y.size))
    return np.array(X), np.array(Y), np.array(Z)


def geometry(cfg: RcaConfig, half: Half, grid=None) -> dict:
    """``Geometry`` dict of one half (script_All.m:227-263)."""
    if grid is None:
        grid = meshgrid_zxy(*grid_axes(cfg))
    e = element_positions(cfg.n_elem)
    zero = np.zeros_like(e)
    if half.receive_x:   # column elements: lines along y at x = e
        ex, ey = e, zero
    else:                # row elements: lines along x at y = e
        ex, ey = zero, e
    return {"gridBfX": grid[0], "gridBfY": grid[1], "gridBfZ": grid[2],
            "posEleX": ex, "posEleY": ey, "posEleZ": zero.copy(),
            "apertureX": float(ex.max() - ex.min()), "apertureY": float(ey.max() - ey.min()),
            "apertureZ": 0.0}


def transmit(half: Half) -> dict:
    """``Transmit`` dict (script_All.m:240-270): receive on X => tilts along y."""
    th = np.asarray(half.theta, dtype=np.float64)
    z = np.zeros_like(th)
    if half.receive_x:
        return {"thetaX": z, "thetaY": th, "flagReceiveX": True}
    return {"thetaX": th, "thetaY": z, "flagReceiveX": False}


def param(cfg: RcaConfig) -> dict:
    return {"samplingFrequency": cfg.fs, "demodulationFrequency": cfg.fd,
            "speedOfSound": cfg.c, "startTime": cfg.t0, "fnumber": cfg.fnumber}


def options(cfg: RcaConfig, compounding: bool = True) -> dict:
    return {"interpolation": cfg.interpolation, "flagCompounding": compounding,
            "flagMaskTransmit": False}


def virtual_sources(half: Half, n_elem: int):
    """FP64 virtual sources at r = 10 m and their minimal distance to the
    array, as the reference wrapper defines them (RCA.m:527-558)."""
    th = np.asarray(half.theta, dtype=np.float64)
    tx = np.tan(th) if not half.receive_x else np.zeros_like(th)
    ty = np.tan(th) if half.receive_x else np.zeros_like(th)
    vz = -10.0 / np.sqrt(1 + tx ** 2 + ty ** 2)
    vx, vy = vz * tx, vz * ty
    e = element_positions(n_elem)
    ex, ey = (e, np.zeros_like(e)) if half.receive_x else (np.zeros_like(e), e)
    d0 = np.sqrt((ey[None, :] - vx[:, None]) ** 2 + (ex[None, :] - vy[:, None]) ** 2 +
                 vz[:, None] ** 2).min(axis=1)
    return vx, vy, vz, d0


def get(name: str) -> RcaConfig:
    return CONFIGS[name]


def scaled(cfg: RcaConfig, **kw) -> RcaConfig:
    """A copy of ``cfg`` with some fields replaced (smaller grids for tests)."""
    d = dict(cfg.__dict__)
    d.update(kw)
    return RcaConfig(**d)


def sigma_pulse(f0: float = F0) -> float:
    """Gaussian envelope of a 2-cycle (FWHM) pulse."""
    return (2.0 / f0) / (2.0 * math.sqrt(2.0 * math.log(2.0)))
