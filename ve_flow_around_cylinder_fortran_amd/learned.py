"""Learned predictor for the arclength continuation: a per-node MLP
(include/vef_mlp.h; BF16 MFMA, FP32 accumulation, Adam) trained online on the
converged states the continuation produces, proposing the next step's state
in place of the tangent predictor x + t*dS (updateSolutionArclength,
src/Continuation_Arclength_mod.F90:339-377).

The network learns the tangent predictor's error per node (inputs: the node's
coordinates, its converged state, its predicted increment, the predicted
increment of EU and dS — 16 inputs, scaled per field by the first values
seen; outputs: the node's 6 unknowns of the error, scaled likewise).  After
every converged step the error of that step's tangent prediction joins a short
buffer and the network takes a number of Adam steps on the buffer.  The
continuation evaluates the step's residual at the tangent prediction and at
the learned proposal and keeps the tangent predictor whenever the learned
state's residual is not smaller, so the FP64 Newton iteration judges every
proposal and the tangent predictor is the fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .assembly import VefError
from .mesh import Mesh

_lib = None


def library_path() -> str:
    return os.environ.get("VEF_MLP_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "libvef_mlp.so")


def load_library():
    """Loads the MLP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise VefError(f"{path} not found: build it with __graft_entry__.build()")
    L = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.vef_mlp_create.argtypes = [i32, i32, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.vef_mlp_num_params.argtypes = [vp]
    L.vef_mlp_num_params.restype = i64
    L.vef_mlp_get_params.argtypes = [vp, vp]
    L.vef_mlp_set_params.argtypes = [vp, vp]
    L.vef_mlp_forward.argtypes = [vp, i64, vp, vp, vp]
    L.vef_mlp_gradient.argtypes = [vp, i64, vp, vp, vp, vp, vp]
    L.vef_mlp_adam_step.argtypes = [vp, i64, vp, vp, ctypes.c_float, vp, vp]
    L.vef_mlp_last_error.argtypes = [vp]
    L.vef_mlp_last_error.restype = ctypes.c_char_p
    L.vef_mlp_destroy.argtypes = [vp]
    L.vef_mlp_destroy.restype = None
    for f in ("vef_mlp_create", "vef_mlp_get_params", "vef_mlp_set_params", "vef_mlp_forward", "vef_mlp_gradient",
              "vef_mlp_adam_step"):
        getattr(L, f).restype = ctypes.c_int
    _lib = L
    return L


class NodeMLP:
    """The network of include/vef_mlp.h on torch device tensors (float32)."""

    def __init__(self, n_in: int, n_out: int, seed: int = 0, device: str = "cuda:0"):
        import torch

        self.torch = torch
        self.lib = load_library()
        self.device = device
        self.n_in, self.n_out = n_in, n_out
        h = ctypes.c_void_p()
        if self.lib.vef_mlp_create(n_in, n_out, seed, ctypes.byref(h)) != 0:
            raise VefError("vef_mlp_create failed: " + self.lib.vef_mlp_last_error(None).decode())
        self._h = h
        self.n_params = int(self.lib.vef_mlp_num_params(h))

    def _check(self, rc, what):
        if rc != 0:
            raise VefError(f"{what} failed: {self.lib.vef_mlp_last_error(self._h).decode()}")

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def _f32(self, a):
        torch = self.torch
        return (a if torch.is_tensor(a) else torch.from_numpy(np.asarray(a))).to(
            device=self.device, dtype=torch.float32).contiguous()

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.vef_mlp_get_params(self._h, ctypes.c_void_p(out.ctypes.data)), "vef_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise VefError("wrong parameter count")
        self._check(self.lib.vef_mlp_set_params(self._h, ctypes.c_void_p(p.ctypes.data)), "vef_mlp_set_params")

    def forward(self, X):
        X = self._f32(X)
        Y = self.torch.empty((X.shape[0], self.n_out), dtype=self.torch.float32, device=self.device)
        self._check(self.lib.vef_mlp_forward(self._h, X.shape[0], ctypes.c_void_p(X.data_ptr()),
                                             ctypes.c_void_p(Y.data_ptr()), self._stream()), "vef_mlp_forward")
        return Y

    def gradient(self, X, T):
        X, T = self._f32(X), self._f32(T)
        g = self.torch.empty(self.n_params, dtype=self.torch.float32, device=self.device)
        loss = self.torch.empty(1, dtype=self.torch.float32, device=self.device)
        self._check(self.lib.vef_mlp_gradient(self._h, X.shape[0], ctypes.c_void_p(X.data_ptr()),
                                              ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                                              ctypes.c_void_p(loss.data_ptr()), self._stream()), "vef_mlp_gradient")
        return g, loss

    def adam_step(self, X, T, lr: float):
        X, T = self._f32(X), self._f32(T)
        loss = self.torch.empty(1, dtype=self.torch.float32, device=self.device)
        self._check(self.lib.vef_mlp_adam_step(self._h, X.shape[0], ctypes.c_void_p(X.data_ptr()),
                                               ctypes.c_void_p(T.data_ptr()), float(lr),
                                               ctypes.c_void_p(loss.data_ptr()), self._stream()), "vef_mlp_adam_step")
        return loss

    def close(self):
        if getattr(self, "_h", None):
            # work on any stream may still use the handle's buffers
            self.torch.cuda.synchronize(self.device)
            self.lib.vef_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


N_FEATURES = 16


class LearnedPredictor:
    """Online-trained per-node correction of the tangent predictor.

    For a step from the converged x_k with tangent prediction
    x_t = x_k + t*dS, the inputs of node i are its coordinates, its state in
    x_k, its predicted increment (x_t - x_k)_i, the predicted increment of EU
    and dS; the output is the node's part of the tangent predictor's error
    x_{k+1} - x_t.  Every converged step adds the error its tangent
    prediction made to a short buffer and trains the network on it; the
    proposal is x_t plus the predicted error with its component along t
    removed, so the pseudo-arclength row t.(x - x_k) = dS still holds.

    overlap=True trains on a second stream (default priority) while the next
    step's Newton solve runs on the main one: the training round for an
    observation is launched after the proposal that follows it, so each
    proposal uses the network trained up to the observation before (the
    handle is never used by both streams at once: training waits for the
    proposal, the next proposal waits for the training)."""

    def __init__(self, mesh: Mesh, seed: int = 0, epochs: int = 200, lr: float = 2e-3, buffer: int = 4,
                 device: str = "cuda:0", overlap: bool = False, sample_nodes: int = 0):
        import torch

        self.torch = torch
        # sample_nodes > 0: each Adam step takes that many rows of the buffer,
        # drawn uniformly (a seeded device generator), instead of all of them
        self.sample_nodes = int(sample_nodes)
        self._gen = torch.Generator(device=device)
        self._gen.manual_seed(seed)
        self.mesh = mesh
        self.net = NodeMLP(N_FEATURES, 6, seed=seed, device=device)
        self.device = device
        self.epochs, self.lr, self.buffer = epochs, lr, buffer
        g = mesh.geometry
        xy = mesh.coords.astype(np.float32)
        self.xy = torch.from_numpy(np.stack([xy[:, 0] / max(g["L_up"], g["L_down"]), xy[:, 1] / g["H"]], 1)).to(device)
        self.data = []          # (X, T) pairs
        self.in_scales = None   # state (6), increment (6), dEU, dS
        self.err_scale = None   # per field
        self.losses = []
        self.overlap = overlap
        self._side = torch.cuda.Stream(device=device) if overlap else None
        self._trained = None    # event of the training round in flight (overlap)
        self._loss = None
        self._rounds = 0        # training rounds launched
        self._untrained = False  # data added since the last round (overlap)

    @staticmethod
    def _nodal(x):
        return x[:-1].view(-1, 6)

    def features(self, x_cur, x_t):
        torch = self.torch
        u = self._nodal(x_cur).float()
        du = (self._nodal(x_t) - self._nodal(x_cur)).float()
        dEU = float(x_t[-1] - x_cur[-1])
        dS = float(torch.linalg.vector_norm(x_t - x_cur))
        if self.in_scales is None:
            self.in_scales = (u.std(dim=0).clamp_min(1e-12), du.abs().max(dim=0).values.clamp_min(1e-12),
                              abs(dEU) if dEU != 0 else 1.0, dS if dS > 0 else 1.0)
        s_u, s_du, s_eu, s_ds = self.in_scales
        n = u.shape[0]
        col = lambda v: torch.full((n, 1), float(v), dtype=torch.float32, device=self.device)  # noqa: E731
        return torch.cat([self.xy, u / s_u, du / s_du, col(dEU / s_eu), col(dS / s_ds)], dim=1).contiguous()

    def observe(self, X, err):
        """Adds (features of a tangent prediction, its error x_{k+1} - x_t) and trains."""
        torch = self.torch
        e = self._nodal(err).float()
        if self.err_scale is None:
            self.err_scale = e.abs().max(dim=0).values.clamp_min(1e-12)
        self.data.append((X, (e / self.err_scale).contiguous()))
        self.data = self.data[-self.buffer:]
        if self.overlap:
            self._untrained = True
            return
        loss = self._train_round()
        self._rounds += 1
        self.losses.append(float(loss.item()))

    def _train_round(self):
        """`epochs` Adam steps on the buffer (or on row samples of it)."""
        torch = self.torch
        Xa = torch.cat([d[0] for d in self.data]).contiguous()
        Ta = torch.cat([d[1] for d in self.data]).contiguous()
        loss = None
        n = Xa.shape[0]
        for _ in range(self.epochs):
            if 0 < self.sample_nodes < n:
                idx = torch.randint(0, n, (self.sample_nodes,), generator=self._gen, device=self.device)
                loss = self.net.adam_step(Xa[idx].contiguous(), Ta[idx].contiguous(), self.lr)
            else:
                loss = self.net.adam_step(Xa, Ta, self.lr)
        return loss

    def ready(self) -> bool:
        return self._rounds >= 1

    def _wait_training(self):
        if self._trained is not None:
            self.torch.cuda.current_stream(self.device).wait_event(self._trained)
            self.losses.append(float(self._loss.item()))
            self._trained = None

    def train_pending(self):
        """overlap=True: launches the training round for the observations not yet
        trained on, on the second stream, after the main stream's work so far."""
        if not (self.overlap and self._untrained):
            return
        torch = self.torch
        self._wait_training()
        main = torch.cuda.current_stream(self.device)
        self._side.wait_stream(main)
        with torch.cuda.stream(self._side):
            for X, T in self.data:
                X.record_stream(self._side)
                T.record_stream(self._side)
            self._loss = self._train_round()
            self._trained = torch.cuda.Event()
            self._trained.record(self._side)
        self._rounds += 1
        self._untrained = False

    def propose(self, X, x_t, tangent):
        """x_t plus the predicted error, projected orthogonally to the tangent."""
        torch = self.torch
        self._wait_training()
        c = torch.zeros_like(x_t)
        c[:-1] = (self.net.forward(X).double() * self.err_scale.double()).reshape(-1)
        c = c - torch.dot(c, tangent) * tangent
        return x_t + c
