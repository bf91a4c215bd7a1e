"""Grid container with the reference's shape/spacing contract.

Mirrors ``CartesianGrid2D`` (reference ``grid/cartesian.py:56-107``):
total arrays are ``[Ny, Nx]`` with one ghost ring; ``from_interior(nx, ny,
Lx, Ly)`` gives ``Nx = nx + 2`` and ``dx = Lx / nx``.
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class CartesianGrid2D:
    Nx: int
    Ny: int
    Lx: float
    Ly: float
    dx: float
    dy: float

    @classmethod
    def from_interior(cls, nx_interior: int, ny_interior: int, Lx: float, Ly: float):
        """``grid/cartesian.py:82-107``."""
        if nx_interior < 1 or ny_interior < 1:
            raise ValueError("interior sizes must be positive")
        return cls(Nx=nx_interior + 2, Ny=ny_interior + 2, Lx=float(Lx), Ly=float(Ly),
                   dx=Lx / nx_interior, dy=Ly / ny_interior)

    @property
    def shape(self):
        return (self.Ny, self.Nx)


ArakawaCGrid2D = CartesianGrid2D
