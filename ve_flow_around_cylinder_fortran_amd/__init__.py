"""MI355X assembler for the 2D square-root-conformation viscoelastic
flow-past-a-cylinder solver: element residuals, forward-difference element
Jacobians and their deterministic CSR scatter, in HIP behind a C ABI
(include/vef.h).  Python side: mesh generation, synthetic states and a ctypes
mirror of the reference's assembly call."""
from .assembly import Assembler, Params, VefError, load_library  # noqa: F401
from .mesh import Mesh, cylinder_channel, workload_mesh, read_dta, write_dta  # noqa: F401
from .state import smooth_state, rest_state  # noqa: F401
