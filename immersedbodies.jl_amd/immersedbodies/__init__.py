"""Host-side mirror of ImmersedBodies.jl's Problem / solver API over libibpm, the
MI355X implementation of the IBPM time step."""
from .problem import (CNAB, AbstractBody, AbstractScheme, FsiBody, GridMotion, GridU, GridV,
                      GridVertices, GridΓ, MultiDomainExtents, PanelSection, PanelState,
                      PresetBody, boundary_ds, boundary_ds_, panel_section, Bodies, CartesianGrid, Fluid, GridIndices, GridVelocity,
                      MovingGrid, MovingRigidBody, MultiDomainGrid, Panels, Problem,
                      RigidBodyTransform, State, StaticBody, StaticGrid, any_fsi, arraysize,
                      boundary_force, boundary_force_, boundary_len, boundary_len_,
                      boundary_pos, boundary_pos_, boundary_total_force, boundary_total_force_,
                      boundary_vel, boundary_vel_, coords, default_gridstep, default_scheme,
                      eachbody, extents, fluid_grid, gridsize, gridstep, n_panels,
                      panel_range, subdomain, timestep, vorticity, vorticity_, x_velocity,
                      x_velocity_, y_velocity, y_velocity_)
from .solver import (AbstractSolver, CnabSolver, HipCnabSolver, advance, create_solver, getstate, solve,
                     solve_)
from . import curves
from .batch import Batch
from .saving import (ArrayValue, MultiDomainValue, SaveHDF5, SaveNothing, SaveNpz,
                     SolutionSaver, SolutionValue, SolutionValues, Timesteps, TimestepRange,
                     body_group, fluid_group)
from .io import load_state, load_state_, save_state

__all__ = [
    "CNAB", "Bodies", "CartesianGrid", "Fluid", "GridIndices", "GridVelocity", "MovingGrid",
    "MovingRigidBody",
    "MultiDomainGrid", "Panels", "Problem", "RigidBodyTransform", "State", "StaticBody",
    "StaticGrid", "boundary_total_force", "coords", "default_gridstep", "default_scheme",
    "fluid_grid", "gridstep", "n_panels", "subdomain", "timestep", "AbstractSolver",
    "HipCnabSolver", "advance", "create_solver", "getstate", "solve", "solve_", "curves",
    "Batch", "any_fsi", "eachbody", "panel_range", "arraysize", "boundary_force",
    "boundary_force_", "boundary_len", "boundary_len_", "boundary_pos", "boundary_pos_",
    "boundary_total_force_", "boundary_vel", "boundary_vel_", "extents", "gridsize",
    "vorticity", "vorticity_", "x_velocity", "x_velocity_", "y_velocity", "y_velocity_",
    "ArrayValue", "MultiDomainValue", "SaveHDF5", "SaveNothing", "SaveNpz", "SolutionSaver",
    "SolutionValues", "TimestepRange", "body_group", "fluid_group", "AbstractBody",
    "AbstractScheme", "FsiBody", "GridMotion", "GridU", "GridV", "GridVertices", "GridΓ",
    "MultiDomainExtents", "PanelSection", "PanelState", "PresetBody", "boundary_ds",
    "boundary_ds_", "panel_section", "CnabSolver", "SolutionValue", "Timesteps",
    "load_state", "load_state_", "save_state",
]
