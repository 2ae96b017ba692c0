"""Kick-drift-kick leapfrog with the tree refit every step and rebuilt on a
schedule, each step captured in a HIP graph (libbhut_hip's bh_sim_*).

bhut itself only offers forces; its galaxy example integrates with scipy's
solve_ivp and builds a new Tree for every right-hand side
(examples/galaxy/galaxy_simulation.py:277-284).  Here positions and
velocities stay in device memory (torch tensors as carriers) for the whole
run; refit follows bhut/tree/refit.py:17-109 and rebuild Tree.rebuild
(bhut/api.py:564-606).
"""

from __future__ import annotations

import numpy as np

from . import _native


class Leapfrog:
    """KDK leapfrog of one system on the GPU.

    precision "fp64": forces from the plain walk (bit-faithful to bhut);
    "fp32": the grouped FP32 form.
    """

    def __init__(self, positions, velocities, masses, *, theta: float = 0.5, softening: float = 0.01,
                 G: float = 1.0, leaf_size: int = 32, precision: str = "fp32",
                 opts: _native.F32Opts | None = None, device: int = 0):
        import torch

        if precision not in ("fp32", "fp64"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', got '{precision}'")
        dev = torch.device("cuda", device)
        as_t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64) if not torch.is_tensor(a) else a,
                                         dtype=torch.float64, device=dev).contiguous().clone()
        self.pos = as_t(positions)
        self.vel = as_t(velocities)
        self.mass = as_t(masses)
        n = self.pos.shape[0]
        if self.pos.ndim != 2 or self.pos.shape[1] not in (2, 3) or self.vel.shape != self.pos.shape \
                or self.mass.shape != (n,):
            raise ValueError("positions and velocities (n, dim) and masses (n,) required")
        self.theta, self.softening, self.G = float(theta), float(softening), float(G)
        self.precision = precision
        torch.cuda.synchronize(dev)
        self.sim = _native.Simulation(self.pos.data_ptr(), self.vel.data_ptr(), self.mass.data_ptr(), n,
                                      self.pos.shape[1], leaf_size=leaf_size, theta=theta, softening=softening,
                                      G=G, precision=32 if precision == "fp32" else 64, opts=opts)

    def run(self, steps: int, dt: float, *, rebuild_every: int = 10, use_graph: bool = True) -> "Leapfrog":
        self.sim.run(dt, steps, rebuild_every, use_graph)
        return self

    def accelerations(self):
        import torch
        out = torch.empty_like(self.pos)
        self.sim.accelerations_dev(out.data_ptr())
        return out

    def kinetic_energy(self) -> float:
        return float(0.5 * (self.mass * (self.vel * self.vel).sum(dim=1)).sum())

    def potential_energy(self, chunk: int = 0, progress=None) -> float:
        """FP64 direct sum of the softened (Plummer) pair potential,
        W = 1/2 sum_i m_i phi_i, optionally over chunks of targets (progress
        is called with the number of targets done after each chunk)."""
        import torch
        n, d = self.pos.shape
        phi = torch.empty(n, dtype=torch.float64, device=self.pos.device)
        torch.cuda.synchronize(self.pos.device)
        step = chunk or n
        for a in range(0, n, step):
            b = min(n, a + step)
            if a == 0 and b == n:
                _native.direct_dev(self.pos.data_ptr(), self.mass.data_ptr(), n, self.pos.data_ptr(), n, d, True,
                                   self.softening, self.G, None, phi.data_ptr())
            else:
                # targets a..b are particles a..b: the self pair is skipped by
                # zero distance contributing nothing only when softened, so
                # remove it explicitly below
                _native.direct_dev(self.pos.data_ptr(), self.mass.data_ptr(), n, self.pos[a:b].data_ptr(), b - a,
                                   d, False, self.softening, self.G, None, phi[a:b].data_ptr())
            _native.ctx().synchronize()
            if progress:
                progress(b)
        if chunk and chunk < n:
            if self.softening > 0:
                phi += self.G * self.mass / self.softening  # the self term -G m_i / eps
            else:
                raise ValueError("chunked potential needs softening > 0")
        return float(0.5 * (self.mass * phi).sum())

    def energy(self, chunk: int = 0, progress=None) -> float:
        return self.kinetic_energy() + self.potential_energy(chunk, progress)

    def info(self) -> dict:
        return self.sim.info()
