"""Seeded synthetic scenes for the benchmark workloads.

A self-contained xoshiro256** generator seeded through splitmix64 (doubles
formed as (x >> 11) * 2^-53), so a seed gives the same scenes on every host.

Workloads (room 10 x 8 x 3 m unless stated):
  c1  the reference's examples/basic_room.yaml at 25x20x12 (upwind, CG)
  c2  256x204x80, 2 fans, 4 boxes, QUICK + buoyancy (beta 3.4e-3, alpha
      2.1e-5, T0 26 C, fan 0 blowing 15 C air), dt 0.005, 200 steps
  c3  the c2 generator at 512x408x160, dt 0.0025, 100 steps
  c4  64 rooms, grids log-uniform per axis in [32,160]x[32,128]x[16,64],
      1-3 fans, 0-5 boxes, dt 0.005, 300 steps
  c5  Poisson arrivals (50/s) of rooms with grids in [24,256]x[24,204]x[24,80],
      1-4 fans, 0-6 boxes, <= 200 steps or until velocity_change < 1e-4
"""
from __future__ import annotations

import math
from typing import Iterator, List, Optional, Tuple

from .config import FanConfig, ObstacleConfig, OpeningConfig, RoomConfig, SceneConfig, SolverConfig

MASK64 = (1 << 64) - 1

# the six axis-aligned (azimuth, elevation) blowing directions
AXIS_DIRECTIONS = [(0.0, 0.0), (180.0, 0.0), (90.0, 0.0), (270.0, 0.0), (0.0, 90.0), (0.0, -90.0)]


def splitmix64(x: int) -> Tuple[int, int]:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return x, z ^ (z >> 31)


class Rng:
    """xoshiro256** (Blackman & Vigna), state seeded by splitmix64."""

    def __init__(self, seed: int):
        x = seed & MASK64
        self.s = []
        for _ in range(4):
            x, z = splitmix64(x)
            self.s.append(z)

    def next_u64(self) -> int:
        s = self.s
        r = (((s[1] * 5) & MASK64) << 7 | ((s[1] * 5) & MASK64) >> 57) & MASK64
        r = (r * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & MASK64
        return r

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def log_uniform_int(self, lo: int, hi: int) -> int:
        return max(lo, min(hi, int(round(math.exp(self.uniform(math.log(lo), math.log(hi)))))))

    def exponential(self, rate: float) -> float:
        return -math.log(1.0 - self.uniform()) / rate


def random_scene(seed: int, grid=(64, 48, 24), size=(10.0, 8.0, 3.0), n_fans: int = 2,
                 n_obstacles: int = 4, openings: bool = False, rng: Optional[Rng] = None,
                 **solver) -> SceneConfig:
    """Fans at uniform interior positions with axis-snapped directions and
    2-5 m/s; boxes with edges in [0.3, 1.5] m anywhere in the room."""
    r = rng if rng is not None else Rng(seed)
    L = size
    fans = []
    for n in range(n_fans):
        pos = tuple(r.uniform(0.5, L[a] - 0.5) for a in range(3))
        d = AXIS_DIRECTIONS[r.integer(0, 5)]
        fans.append(FanConfig(position=pos, direction=d, velocity=r.uniform(2.0, 5.0),
                              radius=0.15, name=f"fan{n}"))
    obs = []
    for n in range(n_obstacles):
        edge = [r.uniform(0.3, 1.5) for _ in range(3)]
        lo = [r.uniform(0.0, L[a] - edge[a]) for a in range(3)]
        obs.append(ObstacleConfig(bbox_min=tuple(lo), bbox_max=tuple(lo[a] + edge[a] for a in range(3)),
                                  name=f"box{n}"))
    ops = []
    if openings:
        ops = [OpeningConfig(wall="north", center=(L[0] * 0.5, 1.2), size=(1.0, 1.2), name="window"),
               OpeningConfig(wall="south", center=(L[0] * 0.3, 1.0), size=(0.8, 2.0), name="door")]
    return SceneConfig(room=RoomConfig(size=tuple(size), grid=tuple(grid)), fans=fans,
                       obstacles=obs, openings=ops, solver=SolverConfig(**solver),
                       metadata={"generator_seed": seed})


def _thermal_solver(dt: float, steps: int, **kw) -> dict:
    s = dict(dt=dt, max_steps=steps, convergence_tol=0.0, advection="quick", buoyancy=True,
             beta=3.4e-3, alpha=2.1e-5, T_initial=26.0)
    s.update(kw)
    return s


def workload_c2(grid=(256, 204, 80), seed: int = 2, dt: float = 0.005, steps: int = 200,
                **kw) -> SceneConfig:
    cfg = random_scene(seed, grid=grid, n_fans=2, n_obstacles=4, **_thermal_solver(dt, steps, **kw))
    cfg.fans[0].T_inlet = 15.0
    return cfg


def workload_c3(seed: int = 3, **kw) -> SceneConfig:
    return workload_c2(grid=(512, 408, 160), seed=seed, dt=0.0025, steps=100, **kw)


def workload_c4(n_scenes: int = 64, seed: int = 4, steps: int = 300, **kw) -> List[SceneConfig]:
    out = []
    for i in range(n_scenes):
        r = Rng((seed << 32) + i)
        grid = (r.log_uniform_int(32, 160), r.log_uniform_int(32, 128), r.log_uniform_int(16, 64))
        out.append(random_scene((seed << 32) + i, grid=grid, n_fans=r.integer(1, 3),
                                n_obstacles=r.integer(0, 5), rng=r,
                                **_thermal_solver(0.005, steps, **kw)))
    return out


def workload_c5(n_requests: int = 2000, rate: float = 50.0, seed: int = 5, steps: int = 200,
                **kw) -> Iterator[Tuple[float, SceneConfig]]:
    """Yields (arrival time [s], scene) pairs of a Poisson stream."""
    r = Rng(seed)
    t = 0.0
    for i in range(n_requests):
        t += r.exponential(rate)
        grid = (r.log_uniform_int(24, 256), r.log_uniform_int(24, 204), r.log_uniform_int(24, 80))
        s = _thermal_solver(0.005, steps, **kw)
        s["convergence_tol"] = 1e-4
        yield t, random_scene(i, grid=grid, n_fans=r.integer(1, 4), n_obstacles=r.integer(0, 6),
                              rng=r, **s)
