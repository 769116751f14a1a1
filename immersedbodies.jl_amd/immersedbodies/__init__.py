"""Host-side mirror of ImmersedBodies.jl's Problem / solver API over libibpm, the
MI355X implementation of the IBPM time step."""
from .problem import (CNAB, Bodies, CartesianGrid, Fluid, GridIndices, GridVelocity,
                      MovingGrid, MovingRigidBody, MultiDomainGrid, Panels, Problem,
                      RigidBodyTransform, State, StaticBody, StaticGrid, any_fsi,
                      boundary_total_force, coords, default_gridstep, default_scheme, eachbody,
                      fluid_grid, gridstep, n_panels, panel_range, subdomain, timestep)
from .solver import (AbstractSolver, HipCnabSolver, advance, create_solver, getstate, solve,
                     solve_)
from . import curves
from .batch import Batch

__all__ = [
    "CNAB", "Bodies", "CartesianGrid", "Fluid", "GridIndices", "GridVelocity", "MovingGrid",
    "MovingRigidBody",
    "MultiDomainGrid", "Panels", "Problem", "RigidBodyTransform", "State", "StaticBody",
    "StaticGrid", "boundary_total_force", "coords", "default_gridstep", "default_scheme",
    "fluid_grid", "gridstep", "n_panels", "subdomain", "timestep", "AbstractSolver",
    "HipCnabSolver", "advance", "create_solver", "getstate", "solve", "solve_", "curves",
    "Batch", "any_fsi", "eachbody", "panel_range",
]
