"""Seeded synthetic scenes for the render path (tests, smoke check, benchmark).

Every field is generated on the host with numpy's PCG64 generator from a fixed
seed, so the HIP renderer and the CPU oracle consume identical bytes.  Voxel i
holds the field evaluated at position i + 0.5, matching the sampler's -0.5
shift (lib/isaac/isaac_texture.hpp:243).  Arrays are laid out x-fastest,
shape (Z, Y, X[, C]), the layout the reference's example sources index with
x + y*X + z*X*Y (example/example.cpp:39-143).

Workloads (sizes in voxels / pixels):
  small    64^3 scalar Gaussian mixture, 512x512, trilinear, step 0.5, 4-point TF
  prod     256^3: density (Gaussian mixture + 3-octave value noise), float3
           curl-noise vector field shown through `length`, 2,000,000 particles
           on a 64^3 cell grid; iso 0.5, one clipping plane, SSAO, 1920x1080
  bricks   the small scene's field at 512^3, 3840x2160, step 0.25
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

RGBA = Tuple[float, float, float, float]


# --------------------------------------------------------------------------- scene description
@dataclass
class VolumeSourceSpec:
    name: str
    data: np.ndarray  # (Z, Y, X) or (Z, Y, X, C) float32, x fastest
    chain: str = "idem"
    tf_points: List[Tuple[int, RGBA]] = field(default_factory=list)
    weight: float = 2.0
    iso: float = 0.0
    persistent: bool = True

    @property
    def feature_dim(self) -> int:
        return 1 if self.data.ndim == 3 else int(self.data.shape[3])


@dataclass
class ParticleSourceSpec:
    name: str
    pos: np.ndarray  # (N, 3) float32, local voxel coordinates
    attr: np.ndarray  # (N, C) float32
    radius: Optional[np.ndarray] = None  # (N,) float32 or None
    constant_radius: float = 0.5
    chain: str = "idem"
    tf_points: List[Tuple[int, RGBA]] = field(default_factory=list)
    weight: float = 2.0

    @property
    def feature_dim(self) -> int:
        return int(self.attr.shape[1])


@dataclass
class Camera:
    look_at: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    rotation: np.ndarray = field(default_factory=lambda: np.eye(3))
    distance: float = -4.5


@dataclass
class SceneSpec:
    name: str
    global_size: Tuple[int, int, int]
    local_size: Tuple[int, int, int]
    particle_grid: Tuple[int, int, int]
    position: Tuple[int, int, int] = (0, 0, 0)
    scale: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    framebuffer: Tuple[int, int] = (512, 512)
    viewport: Optional[Tuple[int, int, int, int]] = None
    transfer_size: int = 1024
    volumes: List[VolumeSourceSpec] = field(default_factory=list)
    particles: List[ParticleSourceSpec] = field(default_factory=list)
    # field sources (float3 vector fields with advection) and the advection steering
    # (step factor, history weight, seeding period, seeding duration, seed points)
    fields: List[VolumeSourceSpec] = field(default_factory=list)
    advection: Optional[Tuple[float, float, int, int, int]] = None
    step: float = 0.5
    interpolation: bool = True
    dither_mode: int = 1
    ao_enabled: bool = False
    ao_weight: float = 0.5
    render_mode: int = 0
    clipping: List[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = field(default_factory=list)
    camera: Camera = field(default_factory=Camera)
    background: RGBA = (0.0, 0.0, 0.0, 0.0)  # as IceT hands it to the draw callback

    def vp(self) -> Tuple[int, int, int, int]:
        if self.viewport is not None:
            return self.viewport
        return (0, 0, self.framebuffer[0], self.framebuffer[1])

    def with_camera(self, cam: Camera) -> "SceneSpec":
        return dataclasses.replace(self, camera=cam)


# --------------------------------------------------------------------------- fields
def _axis_centres(n: int) -> np.ndarray:
    return np.arange(n, dtype=np.float64) + 0.5


def gaussian_mixture(
    shape: Tuple[int, int, int],
    seed: int,
    count: int = 8,
    sigma: Tuple[float, float] = (4.0, 12.0),
    amp: Tuple[float, float] = (0.2, 1.0),
) -> np.ndarray:
    """Sum of `count` axis-separable 3D Gaussians; centres U[0,n)^3, sigma and
    amplitude uniform in the given ranges.  Returns float32 (Z, Y, X)."""
    Z, Y, X = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(0.0, 1.0, size=(count, 3)) * np.array([X, Y, Z], dtype=np.float64)
    s = rng.uniform(sigma[0], sigma[1], size=count)
    a = rng.uniform(amp[0], amp[1], size=count)
    xs, ys, zs = _axis_centres(X), _axis_centres(Y), _axis_centres(Z)
    out = np.zeros((Z, Y, X), dtype=np.float64)
    for k in range(count):
        gx = np.exp(-((xs - c[k, 0]) ** 2) / (2.0 * s[k] ** 2))
        gy = np.exp(-((ys - c[k, 1]) ** 2) / (2.0 * s[k] ** 2))
        gz = np.exp(-((zs - c[k, 2]) ** 2) / (2.0 * s[k] ** 2))
        out += a[k] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return out.astype(np.float32)


def _noise_weights(n: int, freq: int) -> np.ndarray:
    """(n, freq+1) tensor-product weights of smoothstep value noise."""
    u = _axis_centres(n) / n * freq
    i0 = np.minimum(np.floor(u).astype(np.int64), freq - 1)
    t = u - i0
    w = t * t * (3.0 - 2.0 * t)
    W = np.zeros((n, freq + 1), dtype=np.float64)
    W[np.arange(n), i0] = 1.0 - w
    W[np.arange(n), i0 + 1] += w
    return W


def value_noise(shape: Tuple[int, int, int], seed: int, freq: int = 4, octaves: int = 3) -> np.ndarray:
    """Smoothstep value noise, `octaves` octaves (frequency x2, amplitude /2),
    normalised to [0, 1].  float64 (Z, Y, X)."""
    Z, Y, X = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros((Z, Y, X), dtype=np.float64)
    amp_sum = 0.0
    for o in range(octaves):
        f = freq * (2 ** o)
        lat = rng.uniform(0.0, 1.0, size=(f + 1, f + 1, f + 1))
        Wx, Wy, Wz = _noise_weights(X, f), _noise_weights(Y, f), _noise_weights(Z, f)
        t = np.einsum("kji,xi->kjx", lat, Wx)
        t = np.einsum("kjx,yj->kyx", t, Wy)
        t = np.einsum("kyx,zk->zyx", t, Wz)
        a = 0.5 ** o
        out += a * t
        amp_sum += a
    out /= amp_sum
    lo, hi = out.min(), out.max()
    return (out - lo) / max(hi - lo, 1e-12)


def curl_noise(shape: Tuple[int, int, int], seed: int, freq: int = 4) -> np.ndarray:
    """Curl of a 3-component value-noise potential, scaled so max |v| = 1.
    float32 (Z, Y, X, 3)."""
    ax = value_noise(shape, seed * 3 + 0, freq, 2)
    ay = value_noise(shape, seed * 3 + 1, freq, 2)
    az = value_noise(shape, seed * 3 + 2, freq, 2)
    # np.gradient returns derivatives along axes (z, y, x)
    daz_dz, daz_dy, daz_dx = np.gradient(az)
    day_dz, day_dy, day_dx = np.gradient(ay)
    dax_dz, dax_dy, dax_dx = np.gradient(ax)
    del daz_dz, day_dy, dax_dx
    v = np.stack([daz_dy - day_dz, dax_dz - daz_dx, day_dx - dax_dy], axis=-1)
    m = np.sqrt((v * v).sum(axis=-1)).max()
    return (v / max(m, 1e-12)).astype(np.float32)


def uniform_particles(count: int, extent: Tuple[int, int, int], seed: int) -> Tuple[np.ndarray, np.ndarray]:
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(0.0, 1.0, size=(count, 3)) * np.array(extent, dtype=np.float64)
    pos = np.minimum(pos.astype(np.float32), np.nextafter(np.array(extent, np.float32), 0))
    w = rng.uniform(0.0, 1.0, size=(count, 1)).astype(np.float32)
    return pos.astype(np.float32), w


# --------------------------------------------------------------------------- transfer functions
SMALL_TF: List[Tuple[int, RGBA]] = [
    (0, (0.0, 0.0, 0.4, 0.0)),
    (256, (0.0, 0.5, 1.0, 0.05)),
    (614, (1.0, 0.85, 0.1, 0.35)),
    (1024, (1.0, 0.15, 0.1, 0.9)),
]

DENSITY_TF: List[Tuple[int, RGBA]] = [
    (0, (0.05, 0.05, 0.3, 0.0)),
    (307, (0.1, 0.4, 0.9, 0.01)),
    (512, (0.95, 0.8, 0.2, 0.15)),
    (666, (1.0, 0.45, 0.1, 0.6)),
    (1024, (1.0, 0.2, 0.1, 0.9)),
]

VECTOR_TF: List[Tuple[int, RGBA]] = [
    (0, (0.2, 1.0, 0.2, 0.0)),
    (512, (0.2, 0.8, 0.5, 0.01)),
    (1024, (0.9, 1.0, 0.9, 0.15)),
]

# particles become visible (alpha >= 0.5) only for weights >= 0.995
PARTICLE_TF: List[Tuple[int, RGBA]] = [
    (0, (1.0, 1.0, 1.0, 0.0)),
    (1014, (1.0, 1.0, 0.8, 0.0)),
    (1024, (1.0, 0.9, 0.3, 1.0)),
]


# --------------------------------------------------------------------------- cameras
def rotation_axis(axis: Sequence[float], deg: float) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    th = math.radians(deg)
    c, s = math.cos(th), math.sin(th)
    x, y, z = a
    C = 1.0 - c
    return np.array(
        [
            [c + x * x * C, x * y * C - z * s, x * z * C + y * s],
            [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
            [z * x * C - y * s, z * y * C + x * s, c + z * z * C],
        ]
    )


DEFAULT_POSE = Camera()
FILL_POSE = Camera(rotation=rotation_axis((1.0, 1.0, 0.0), 30.0), distance=-2.5)


def random_pose(rng: np.random.Generator) -> Camera:
    axis = rng.normal(size=3)
    deg = float(rng.uniform(0.0, 360.0))
    dist = float(rng.uniform(-4.5, -2.5))
    return Camera(rotation=rotation_axis(axis, deg), distance=dist)


# --------------------------------------------------------------------------- workloads
def small_scene(n: int = 64, fb: Tuple[int, int] = (512, 512), seed: int = 1) -> SceneSpec:
    """64^3 float32 Gaussian mixture (8 Gaussians, sigma U[4,12], amplitude
    U[0.2,1]), 512x512, trilinear, step 0.5, 4-point TF, no iso, no SSAO."""
    s = n / 64.0
    dens = gaussian_mixture((n, n, n), seed, 8, (4.0 * s, 12.0 * s))
    vol = VolumeSourceSpec("gaussians", dens, "idem", list(SMALL_TF), 2.0, 0.0)
    return SceneSpec(
        name=f"small{n}",
        global_size=(n, n, n),
        local_size=(n, n, n),
        particle_grid=(max(1, n // 4),) * 3,
        framebuffer=fb,
        volumes=[vol],
        step=0.5,
        interpolation=True,
    )


def precision_scene(n: int = 512, fb: Tuple[int, int] = (3840, 2160), step: float = 0.25, seed: int = 1) -> SceneSpec:
    """The small scene's single Gaussian-mixture source at n^3 (sigma scaled with
    n), 3840x2160, trilinear, step 0.25: the workload of the FP16/BF16 brick study."""
    spec = small_scene(n=n, fb=fb, seed=seed)
    return dataclasses.replace(spec, name=f"precision{n}", step=step)


def prod_scene(
    n: int = 256,
    fb: Tuple[int, int] = (1920, 1080),
    n_particles: int = 2_000_000,
    cells: Optional[int] = None,
    pose: Camera = DEFAULT_POSE,
) -> SceneSpec:
    """256^3 x 3 sources, iso + clipping plane + SSAO at 1080p."""
    s = n / 64.0
    g = gaussian_mixture((n, n, n), 2, 8, (4.0 * s, 12.0 * s))
    g = g / max(float(g.max()), 1e-6)
    noise = value_noise((n, n, n), 2, 4, 3)
    dens = np.clip(0.75 * g + 0.3 * noise, 0.0, 1.0).astype(np.float32)
    vec = curl_noise((n, n, n), 3, 4)
    cells = cells if cells is not None else max(1, n // 4)
    pos, w = uniform_particles(n_particles, (n, n, n), 4)
    return SceneSpec(
        name=f"prod{n}",
        global_size=(n, n, n),
        local_size=(n, n, n),
        particle_grid=(cells, cells, cells),
        framebuffer=fb,
        volumes=[
            VolumeSourceSpec("density", dens, "idem", list(DENSITY_TF), 2.0, 0.5),
            VolumeSourceSpec("vector", vec, "length", list(VECTOR_TF), 2.0, 0.0),
        ],
        particles=[ParticleSourceSpec("particles", pos, w, None, 0.5, "idem", list(PARTICLE_TF), 2.0)],
        step=0.5,
        interpolation=True,
        ao_enabled=True,
        ao_weight=0.5,
        clipping=[((0.0, 0.0, 0.0), (1.0, 0.0, 0.0))],
        camera=pose,
    )
