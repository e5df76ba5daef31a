"""vef_arclength_noalloc (no counterpart in the reference; for callers that
cannot allocate): the arclength row through the C ABI with the library's
file-scope scratch array, against vef_arclength (the handle's own scratch),
bit for bit.  Two calls in a row share the array: first the 40,000-element
mesh (40 node chunks of partial sums), then the 1,600-element mesh (2 chunks),
each for the pseudo-arclength row with its Jacobian row and for the pinned
row."""
import ctypes

import numpy as np
import pytest

from ve_flow_around_cylinder_fortran_amd import Assembler, Params, cylinder_channel
from ve_flow_around_cylinder_fortran_amd.mesh import workload_mesh

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

VEF_RES_JAC = 1
PIN, PSEUDO = 0, 1


def rows(asm, seed, fn, mode):
    """b and the values of row N-1 after `fn` (vef_arclength or its variant)."""
    N, nnz = asm.N, asm.nnz
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    x = torch.randn(N, dtype=torch.float64, device="cuda:0", generator=g)
    xo = x + 1e-3 * torch.randn(N, dtype=torch.float64, device="cuda:0", generator=g)
    t = torch.randn(N, dtype=torch.float64, device="cuda:0", generator=g)
    t = t / torch.linalg.vector_norm(t)
    b = torch.full((N,), 7.0, dtype=torch.float64, device="cuda:0")
    v = torch.full((nnz,), 3.0, dtype=torch.float64, device="cuda:0")
    p = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(asm._h, mode, p(x), p(xo), p(t), 3 if mode == PSEUDO else 2, 0.1, 0.02, 0.75, VEF_RES_JAC, p(b), p(v),
            st)
    torch.cuda.synchronize()
    assert rc == 0, asm.lib.vef_last_error(asm._h).decode()
    rowptr, _ = asm.pattern()
    return b.cpu().numpy(), v[int(rowptr[N - 1]):].cpu().numpy()


def test_noalloc_arclength_matches_vef_arclength():
    pset = dict(ReN=0.0, BvN=0.59, VeN=0.25, DfN=1e-3, WiN0=1.0, model="L-PTT")
    for name, mesh in (("40k", workload_mesh(40000)), ("1.6k", cylinder_channel(16, 12, 20, 32))):
        print(f"[gpu noalloc] {name}: {mesh.n_nodes} nodes", flush=True)
        asm = Assembler(mesh, Params(**pset))
        lib = asm.lib
        vp, i32, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        for f in (lib.vef_arclength, lib.vef_arclength_noalloc):
            f.argtypes = [vp, i32, vp, vp, vp, i32, d, d, d, i32, vp, vp, vp]
            f.restype = ctypes.c_int
        for mode in (PSEUDO, PIN):
            b0, r0 = rows(asm, 11, lib.vef_arclength, mode)
            b1, r1 = rows(asm, 11, lib.vef_arclength_noalloc, mode)
            assert np.array_equal(b0, b1) and np.array_equal(r0, r1), (name, mode)
            assert np.isfinite(b1[-1]) and b1[-1] != 7.0
        assert lib.vef_arclength_noalloc(None, PSEUDO, None, None, None, 3, 0.1, 0.02, 0.75, VEF_RES_JAC, None, None,
                                         None) != 0
