"""Multifrontal LU in nested-dissection order: the symbolic analysis (host)
and the GPU factorisation and solves (libvef_mf.so, include/vef_mf.h).

This is the role MUMPS plays for the reference: solveLinearSystem hands each
Newton matrix to PETSc's KSPPREONLY + PCLU with MATSOLVERMUMPS
(src/Petsc_mod.F90:127-141, 807-871), a multifrontal sparse LU.  The node
unknowns (6 per mesh node) are ordered by nested dissection of the mesh's
elements; the dense arclength row and column stay a border (Schur complement
of one scalar), as in linsolve.BandLU.

Symbolic analysis (once per mesh, here):

* Nested dissection on element sets: a set larger than `leaf` elements is
  split at the median of its element centroids along its longer extent; the
  separator is the set of still-unnumbered nodes the two halves share (the
  element interface: corner and mid-side nodes of the shared edges).  Each
  half recurses; a small set's remaining nodes form a leaf.  Nodes are
  numbered in postorder (left subtree, right subtree, separator), so a
  subtree's nodes are consecutive.
* Front of tree node t: its own nodes (the pivots, s = 6 per node) and the
  update nodes U(t) — the nodes of t's elements numbered after t's subtree,
  all of which lie in ancestors' separators (u = 6 per node).  Every element
  touching a pivot node of t belongs to t's element set, so each stored
  entry A(i, j) of the node block lands in the front of whichever of its two
  nodes is numbered first.
* Fronts are processed by height (leaves 0; a parent after all its
  children).  Within a height, fronts are grouped into buckets of similar
  size and padded to the bucket's largest pivot count s_pad and update count
  u_pad (pivot padding is an identity block), so that each bucket is one
  strided batch: the bucket's fronts are consecutive column-major
  (s_pad + u_pad)^2 blocks.

Numerical factorisation (GPU, per height): the children's update blocks
(Schur complements) are added into their parents' fronts (extend-add, the
first children of all parents, then the second ones: the order is fixed);
then per bucket F11 = P L U (rocSOLVER getrf, partial pivoting within the
pivot block), X = F11^-1 F12 (getrs) in place of F12, and F22 -= F21 X
(rocBLAS gemm) — F22 is then the front's update block for its parent.
Solves: forward by height, w = F11^-1 b_s and b_u -= F21 w (the update rows'
contributions summed in a fixed order per unknown); backward in reverse,
x_s = w - X x_u.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import sys
from typing import List

import numpy as np


# ---------------------------------------------------------------------------
# Symbolic analysis
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Front:
    nodes: np.ndarray        # pivot nodes (mesh node ids), in elimination order
    update: np.ndarray       # update nodes (mesh node ids), in elimination order
    children: List[int]
    elements: np.ndarray
    first: int = 0           # position of the subtree's first node
    last: int = -1           # position of the front's last pivot node
    height: int = 0
    parent: int = -1
    slot: int = 0            # 0: first child of its parent, 1: second


def nested_dissection(coords: np.ndarray, conn: np.ndarray, leaf: int = 4):
    """Element-interface nested dissection.  Returns (fronts, order, pos):
    the front tree (root = fronts[0]), the node order (position -> node) and
    its inverse (node -> position)."""
    conn = np.asarray(conn, dtype=np.int64)
    n_nodes = int(coords.shape[0])
    cen = np.asarray(coords, dtype=np.float64)[conn].mean(axis=1)
    placed = np.zeros(n_nodes, dtype=bool)
    fronts: List[Front] = []
    stack = [(np.arange(conn.shape[0], dtype=np.int64), -1, 0)]
    while stack:  # preorder: a separator is placed before its halves are split
        els, parent, slot = stack.pop()
        idx = len(fronts)
        if len(els) <= leaf:
            nodes = np.unique(conn[els])
            own = nodes[~placed[nodes]]
            placed[own] = True
            fronts.append(Front(own, None, [], els, parent=parent, slot=slot))
        else:
            c = cen[els]
            ax = int(np.argmax(c.max(axis=0) - c.min(axis=0)))
            key = c[:, ax]
            med = np.median(key)
            left, right = els[key <= med], els[key > med]
            if len(left) == 0 or len(right) == 0:
                o = np.argsort(key, kind="stable")
                h = len(els) // 2
                left, right = els[o[:h]], els[o[h:]]
            sep = np.intersect1d(np.unique(conn[left]), np.unique(conn[right]))
            sep = sep[~placed[sep]]
            placed[sep] = True
            fronts.append(Front(sep, None, [], els, parent=parent, slot=slot))
            # pushed right first so that the left half is numbered first
            stack.append((right, idx, 1))
            stack.append((left, idx, 0))
        if parent >= 0:
            fronts[parent].children.append(idx)
    if not placed.all():
        raise ValueError("nested dissection left nodes unnumbered (nodes outside every element?)")
    # postorder numbering: children's subtrees, then the front's own nodes
    order = np.empty(n_nodes, dtype=np.int64)
    nxt = 0
    post: List[int] = []
    visit = [(0, False)]
    while visit:
        t, done = visit.pop()
        if not done:
            fronts[t].first = nxt
            visit.append((t, True))
            for c in reversed(fronts[t].children):
                visit.append((c, False))
            continue
        own = fronts[t].nodes
        order[nxt:nxt + len(own)] = own
        nxt += len(own)
        fronts[t].last = nxt - 1
        post.append(t)
    pos = np.empty(n_nodes, dtype=np.int64)
    pos[order] = np.arange(n_nodes)
    for t in post:  # children before parents
        f = fronts[t]
        f.nodes = order[f.last - len(f.nodes) + 1:f.last + 1]
        nodes = np.unique(conn[f.elements])
        upd = nodes[pos[nodes] > f.last]
        f.update = upd[np.argsort(pos[upd])]
        f.height = 0 if not f.children else 1 + max(fronts[c].height for c in f.children)
    return fronts, order, pos


def _front_flops(s, u):
    s = np.asarray(s, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    return 2.0 / 3.0 * s ** 3 + 2.0 * s * s * u + 2.0 * s * u * u


def buckets_by_size(s: np.ndarray, u: np.ndarray, slack: float = 1.3, min_flops: float = 2.0e6):
    """Groups fronts of one height (pivot and update counts s, u) into
    buckets padded to their largest s and u: fronts sorted by size, a bucket
    grows while its padded flops stay within `slack` times its own (buckets
    whose padded work is below min_flops take any size)."""
    order = np.lexsort((u, s))[::-1]
    out = []
    cur: List[int] = []
    s_max = u_max = 0
    real = 0.0
    for f in order.tolist():
        s2, u2 = max(s_max, int(s[f])), max(u_max, int(u[f]))
        real2 = real + float(_front_flops(s[f], u[f]))
        padded = (len(cur) + 1) * float(_front_flops(s2, u2))
        if cur and padded > slack * real2 and padded > min_flops:
            out.append((cur, s_max, u_max))
            cur, s_max, u_max, real = [], 0, 0, 0.0
            s2, u2, real2 = int(s[f]), int(u[f]), float(_front_flops(s[f], u[f]))
        cur.append(f)
        s_max, u_max, real = s2, u2, real2
    if cur:
        out.append((cur, s_max, u_max))
    return out


@dataclasses.dataclass
class MultifrontalPlan:
    """Everything the GPU library needs, as flat arrays (fronts numbered
    bucket by bucket, buckets by height)."""
    n: int                      # node unknowns (6 per node), ND order
    pos: np.ndarray             # node -> ND position
    unknown_pos: np.ndarray     # node unknown (6 g + q) -> ND unknown index (6 pos[g] + q)
    bucket_height: np.ndarray
    bucket_count: np.ndarray
    bucket_s: np.ndarray        # padded pivot rows
    bucket_u: np.ndarray        # padded update rows
    front_s: np.ndarray         # pivot rows (6 per node)
    front_u: np.ndarray         # update rows
    rows_off: np.ndarray        # n_fronts + 1
    rows: np.ndarray            # ND unknown index per front row: s pivots, then u updates
    parent: np.ndarray          # parent front or -1
    slot: np.ndarray            # child slot in the parent (0, 1)
    ea_off: np.ndarray          # n_fronts + 1, offsets of ea_map
    ea_map: np.ndarray          # per update row: its physical row in the parent front
    front_base: np.ndarray      # element offset of each front in the front storage
    total: int                  # elements of the front storage
    flops: float                # factorisation flops (real sizes)
    flops_padded: float

    @property
    def n_fronts(self) -> int:
        return len(self.front_s)

    @property
    def m_pad(self) -> np.ndarray:
        """Padded order of each front."""
        b = np.repeat(np.arange(len(self.bucket_count)), self.bucket_count)
        return (self.bucket_s + self.bucket_u)[b]

    @property
    def s_pad(self) -> np.ndarray:
        b = np.repeat(np.arange(len(self.bucket_count)), self.bucket_count)
        return self.bucket_s[b]


def build_plan(coords, conn, leaf: int = 4, slack: float = 1.3) -> MultifrontalPlan:
    fronts, order, pos = nested_dissection(coords, conn, leaf)
    n_nodes = len(order)
    H = max(f.height for f in fronts)
    # buckets by height, fronts renumbered bucket by bucket
    new_id = np.empty(len(fronts), dtype=np.int64)
    bh, bc, bs, bu = [], [], [], []
    seq: List[int] = []
    for h in range(H + 1):
        ids = np.array([t for t, f in enumerate(fronts) if f.height == h], dtype=np.int64)
        s = np.array([6 * len(fronts[t].nodes) for t in ids], dtype=np.int64)
        u = np.array([6 * len(fronts[t].update) for t in ids], dtype=np.int64)
        for members, s_max, u_max in buckets_by_size(s, u, slack):
            bh.append(h)
            bc.append(len(members))
            bs.append(max(s_max, 0))
            bu.append(u_max)
            for k in members:
                new_id[ids[k]] = len(seq)
                seq.append(int(ids[k]))
    nf = len(seq)
    bucket_s = np.array(bs, dtype=np.int64)
    bucket_u = np.array(bu, dtype=np.int64)
    bucket_count = np.array(bc, dtype=np.int64)
    m_b = bucket_s + bucket_u
    bucket_base = np.concatenate([[0], np.cumsum(bucket_count * m_b * m_b)])
    front_base = np.empty(nf, dtype=np.int64)
    s_pad = np.empty(nf, dtype=np.int64)
    m_pad = np.empty(nf, dtype=np.int64)
    k = 0
    for b in range(len(bucket_count)):
        for j in range(bucket_count[b]):
            front_base[k] = bucket_base[b] + j * m_b[b] * m_b[b]
            s_pad[k] = bucket_s[b]
            m_pad[k] = m_b[b]
            k += 1
    front_s = np.array([6 * len(fronts[t].nodes) for t in seq], dtype=np.int64)
    front_u = np.array([6 * len(fronts[t].update) for t in seq], dtype=np.int64)
    parent = np.array([new_id[fronts[t].parent] if fronts[t].parent >= 0 else -1 for t in seq], dtype=np.int64)
    slot = np.array([fronts[t].slot for t in seq], dtype=np.int64)
    q6 = np.arange(6, dtype=np.int64)
    rows_l, ea_l = [], []
    for k, t in enumerate(seq):
        f = fronts[t]
        nodes = np.concatenate([f.nodes, f.update])
        rows_l.append((6 * pos[nodes][:, None] + q6[None, :]).ravel())
    for k, t in enumerate(seq):
        f = fronts[t]
        p = parent[k]
        if p < 0:
            ea_l.append(np.zeros(0, dtype=np.int64))
            continue
        pf = fronts[seq[p]]
        # physical row of each update node in the parent: pivot nodes first, then update nodes at s_pad
        first_piv = pf.last - len(pf.nodes) + 1
        pp = pos[f.update]
        in_piv = pp <= pf.last
        j = np.searchsorted(pos[pf.update], pp)
        node_row = np.where(in_piv, 6 * (pp - first_piv), s_pad[p] + 6 * j)
        if len(pf.update):
            found = pos[pf.update][np.minimum(j, len(pf.update) - 1)] == pp
        else:
            found = np.zeros(len(pp), dtype=bool)
        if not np.all(np.where(in_piv, pp >= first_piv, found)):
            raise AssertionError("update node missing from its parent's front")
        ea_l.append((node_row[:, None] + q6[None, :]).ravel())
    rows_off = np.concatenate([[0], np.cumsum([len(r) for r in rows_l])]).astype(np.int64)
    ea_off = np.concatenate([[0], np.cumsum([len(e) for e in ea_l])]).astype(np.int64)
    unknown_pos = (6 * pos[:, None] + q6[None, :]).ravel()
    fl = float(_front_flops(front_s, front_u).sum())
    flp = float((bucket_count * _front_flops(bucket_s, bucket_u)).sum())
    plan = MultifrontalPlan(
        n=6 * n_nodes, pos=pos, unknown_pos=unknown_pos, bucket_height=np.array(bh, dtype=np.int64),
        bucket_count=bucket_count, bucket_s=bucket_s, bucket_u=bucket_u, front_s=front_s, front_u=front_u,
        rows_off=rows_off, rows=np.concatenate(rows_l).astype(np.int64), parent=parent, slot=slot, ea_off=ea_off,
        ea_map=np.concatenate(ea_l).astype(np.int64), front_base=front_base, total=int(bucket_base[-1]), flops=fl,
        flops_padded=flp)
    plan._fronts = fronts       # tree in analysis numbering (for the value map)
    plan._new_id = new_id
    return plan


def value_map(plan: MultifrontalPlan, rowptr: np.ndarray, col: np.ndarray) -> np.ndarray:
    """dst[k] = position in the front storage of stored entry k of the CSR
    pattern (node block), -1 for the arclength row and column (a pattern of
    plan.n rows has none).  The pattern must hold full 6x6 node blocks
    (preAllocatePetsc's)."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    n = len(rowptr) - 1
    n1 = n - 1 if n == plan.n + 1 else n   # with or without the arclength border
    nn = n1 // 6
    if n1 != plan.n:
        raise ValueError("pattern and plan sizes differ")
    fronts = plan._fronts
    new_id = plan._new_id
    pos = plan.pos
    # node pairs from the rows of the first unknown of each node
    r0 = 6 * np.arange(nn, dtype=np.int64)
    lens = rowptr[r0 + 1] - rowptr[r0]
    a = np.repeat(np.arange(nn, dtype=np.int64), lens)
    slot0 = np.repeat(rowptr[r0], lens) + (np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens))
    c0 = col[slot0]
    keep = (c0 < n1) & (c0 % 6 == 0)
    a, slot0, b = a[keep], slot0[keep], c0[keep] // 6
    # full node blocks: every unknown row of a node has the first row's columns
    for q in range(1, 6):
        if not np.array_equal(rowptr[r0 + q + 1] - rowptr[r0 + q], lens):
            raise ValueError("the pattern's node blocks are not full 6x6 blocks")
    # front and local node rows
    owner = np.where(pos[a] <= pos[b], a, b)
    other = np.where(pos[a] <= pos[b], b, a)
    front_of_node = np.empty(nn, dtype=np.int64)
    first_of = np.empty(nn, dtype=np.int64)
    for t, f in enumerate(fronts):
        front_of_node[f.nodes] = new_id[t]
        first_of[f.nodes] = f.last - len(f.nodes) + 1
    F = front_of_node[owner]
    s_pad = plan.s_pad
    m_pad = plan.m_pad
    row_owner = 6 * (pos[owner] - first_of[owner])
    # the other node: a pivot of the same front, or an update node (index in the update list)
    same = front_of_node[other] == F
    upd_pos = np.full(len(a), -1, dtype=np.int64)
    # keys (front, position) of every update node, sorted
    ukey_f = np.concatenate([np.full(len(f.update), new_id[t], dtype=np.int64) for t, f in enumerate(fronts)])
    ukey_p = np.concatenate([pos[f.update] for f in fronts])
    ukey_j = np.concatenate([np.arange(len(f.update), dtype=np.int64) for f in fronts])
    o = np.lexsort((ukey_p, ukey_f))
    keys = ukey_f[o] * nn + ukey_p[o]
    want = F[~same] * nn + pos[other[~same]]
    at = np.searchsorted(keys, want)
    if np.any(at >= len(keys)) or np.any(keys[np.minimum(at, len(keys) - 1)] != want):
        raise AssertionError("an entry's later node is not in its front")
    upd_pos[~same] = ukey_j[o][at]
    row_other = np.where(same, 6 * (pos[other] - first_of[other]), s_pad[F] + 6 * upd_pos)
    row_a = np.where(pos[a] <= pos[b], row_owner, row_other)
    row_b = np.where(pos[a] <= pos[b], row_other, row_owner)
    base = plan.front_base[F] + row_a + row_b * m_pad[F]   # (6a, 6b) entry: row of node a, column of node b
    dst = np.full(len(col), -1, dtype=np.int64)
    off = slot0 - rowptr[6 * a]                               # column offset of block b within a's rows
    for q in range(6):
        rs = rowptr[6 * a + q] + off
        for qq in range(6):
            dst[rs + qq] = base + q + qq * m_pad[F]
    return dst


# ---------------------------------------------------------------------------
# Reference execution of a plan in numpy (used by the tests as the checker of
# the index maps the GPU library consumes)
# ---------------------------------------------------------------------------
def numpy_factor_solve(plan: MultifrontalPlan, dst: np.ndarray, values: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Solves the node block A11 x = b (b in the original unknown order) by
    executing the plan's scatter, extend-add, partial factorisations and
    solves with numpy/scipy on the host."""
    import scipy.linalg as sl

    buf = np.zeros(plan.total)
    m_pad, s_pad = plan.m_pad, plan.s_pad
    nf = plan.n_fronts
    for f in range(nf):
        base, m, s = plan.front_base[f], m_pad[f], s_pad[f]
        for i in range(plan.front_s[f], s):
            buf[base + i + i * m] = 1.0
    k = dst >= 0
    buf[dst[k]] = values[k]
    F = [buf[plan.front_base[f]:plan.front_base[f] + m_pad[f] ** 2].reshape(m_pad[f], m_pad[f]).T for f in range(nf)]
    heights = plan.bucket_height[np.repeat(np.arange(len(plan.bucket_count)), plan.bucket_count)]
    lu = [None] * nf
    for h in range(int(heights.max()) + 1):
        for sl_ in (0, 1):
            for f in np.nonzero((heights < h))[0]:
                p = plan.parent[f]
                if p >= 0 and heights[p] == h and plan.slot[f] == sl_:
                    s, u = s_pad[f], plan.front_u[f]
                    mp = plan.ea_map[plan.ea_off[f]:plan.ea_off[f + 1]]
                    F[p][np.ix_(mp, mp)] += F[f][s:s + u, s:s + u]
        for f in np.nonzero(heights == h)[0]:
            s = s_pad[f]
            if s == 0:
                continue
            lu[f] = sl.lu_factor(F[f][:s, :s])
            F[f][:s, s:] = sl.lu_solve(lu[f], F[f][:s, s:])
            F[f][s:, s:] -= F[f][s:, :s] @ F[f][:s, s:]
    r = np.zeros(plan.n)
    r[plan.unknown_pos] = b
    W = [None] * nf
    for h in range(int(heights.max()) + 1):
        for f in np.nonzero(heights == h)[0]:
            s, fs, u = s_pad[f], plan.front_s[f], plan.front_u[f]
            rows = plan.rows[plan.rows_off[f]:plan.rows_off[f + 1]]
            w = np.zeros(s)
            w[:fs] = r[rows[:fs]]
            if s:
                w = sl.lu_solve(lu[f], w)
            W[f] = w
            r[rows[fs:]] -= F[f][s:s + u, :s] @ w
    x = np.zeros(plan.n)
    for h in range(int(heights.max()), -1, -1):
        for f in np.nonzero(heights == h)[0]:
            s, fs, u = s_pad[f], plan.front_s[f], plan.front_u[f]
            rows = plan.rows[plan.rows_off[f]:plan.rows_off[f + 1]]
            w = W[f] - F[f][:s, s:s + u] @ x[rows[fs:]]
            x[rows[:fs]] = w[:fs]
    return x[plan.unknown_pos]


# ---------------------------------------------------------------------------
# GPU factorisation (libvef_mf.so)
# ---------------------------------------------------------------------------
class _Plan(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("n_buckets", ctypes.c_int32), ("bucket_height", ctypes.c_void_p),
                ("bucket_count", ctypes.c_void_p), ("bucket_s", ctypes.c_void_p), ("bucket_u", ctypes.c_void_p),
                ("n_fronts", ctypes.c_int32), ("front_s", ctypes.c_void_p), ("front_u", ctypes.c_void_p),
                ("rows_off", ctypes.c_void_p), ("rows", ctypes.c_void_p), ("parent", ctypes.c_void_p),
                ("slot", ctypes.c_void_p), ("ea_off", ctypes.c_void_p), ("ea_map", ctypes.c_void_p)]


_mf = None


def _mf_lib():
    """libvef_mf.so (multifrontal LU, include/vef_mf.h)."""
    global _mf
    if _mf is None:
        from .linsolve import RfLUError

        path = os.environ.get("VEF_MF_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                "libvef_mf.so")
        if not os.path.exists(path):
            raise RfLUError(f"{path} not found: build it with __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.vef_mf_create.argtypes = [ctypes.POINTER(_Plan), i32, ctypes.POINTER(vp)]
        L.vef_mf_layout.argtypes = [vp, vp, ctypes.POINTER(i64)]
        L.vef_mf_device_bytes.argtypes = [vp]
        L.vef_mf_device_bytes.restype = i64
        L.vef_mf_set.argtypes = [vp, i64, vp, vp, vp]
        L.vef_mf_factor.argtypes = [vp, vp]
        L.vef_mf_profile.argtypes = [vp, i32]
        L.vef_mf_heights.argtypes = [vp]
        L.vef_mf_heights.restype = i32
        L.vef_mf_profile_read.argtypes = [vp, vp]
        L.vef_mf_set_pivoting.argtypes = [vp, i32]
        L.vef_mf_zero_pivots.argtypes = [vp, ctypes.POINTER(i32)]
        L.vef_mf_solve.argtypes = [vp, vp, vp, vp]
        L.vef_mf_csr_mv.argtypes = [i64, vp, vp, vp, vp, vp, i32, vp, vp, vp]
        L.vef_mf_last_error.argtypes = [vp]
        L.vef_mf_last_error.restype = ctypes.c_char_p
        L.vef_mf_destroy.argtypes = [vp]
        _mf = L
    return _mf


def _border_base():
    from .linsolve import _BorderedLU

    return _BorderedLU


class MultifrontalLU(_border_base()):
    """Multifrontal LU of the Newton matrix on the GPU in nested-dissection
    order (the MUMPS role), the arclength row and column as a border.  Same
    interface as linsolve.BandLU (factor(values), solve(b), precision "f64"
    or "f32" with FP64 residuals and GMRES refinement); it needs the mesh
    (coordinates and connectivity) for the dissection.

    pivoting=False (the default) factorises each front's pivot block without
    row interchanges: the nested-dissection order keeps each node's unknowns
    together with its velocities before its pressure, so the pressure pivots
    are Schur complements of the velocity block, and on the 40,000-element
    Newton matrix the factorisation takes 88 ms in FP32 and 111 ms in FP64
    against 162 and 190 ms with getrf's partial pivoting, its solves
    more accurate (relative residual 3e-6 and 6e-15 against 2e-4 and 3e-13).
    Every solve is checked as with pivoting; one that fails (a zero or tiny
    pivot: non-finite, or above the tolerance after refinement) refactorises
    that matrix with pivoting and solves again (pivot_fallbacks counts them).
    pivoting=True always pivots."""

    def __init__(self, rowptr, col, coords, conn, device="cuda:0", check_tol: float = None, precision: str = "f64",
                 refine_tol: float = 1.0e-3, leaf: int = 4, plan: MultifrontalPlan = None, pivoting: bool = False):
        import torch

        from .linsolve import RfLUError

        self.plan = plan if plan is not None else build_plan(coords, conn, leaf)
        self._border_setup(rowptr, col, device, precision, check_tol, refine_tol, self.plan.unknown_pos)
        if self.n1 != self.plan.n:
            raise RfLUError("MultifrontalLU: the pattern and the mesh differ")
        P = self.plan
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)  # noqa: E731
        keep = [i32(P.bucket_height), i32(P.bucket_count), i32(P.bucket_s), i32(P.bucket_u), i32(P.front_s),
                i32(P.front_u), i64(P.rows_off), i32(P.rows), i32(P.parent), i32(P.slot), i64(P.ea_off),
                i32(P.ea_map)]
        ptr = [ctypes.c_void_p(a.ctypes.data) for a in keep]
        cp = _Plan(P.n, len(P.bucket_count), ptr[0], ptr[1], ptr[2], ptr[3], P.n_fronts, ptr[4], ptr[5], ptr[6],
                   ptr[7], ptr[8], ptr[9], ptr[10], ptr[11])
        L = _mf_lib()
        self.h = ctypes.c_void_p()
        rc = L.vef_mf_create(ctypes.byref(cp), 64 if precision == "f64" else 32, ctypes.byref(self.h))
        if rc != 0:
            msg = L.vef_mf_last_error(self.h).decode() if self.h else "bad arguments"
            self.close()
            raise RfLUError(f"vef_mf_create: {msg}")
        base = np.zeros(P.n_fronts, dtype=np.int64)
        total = ctypes.c_int64()
        self._err(L.vef_mf_layout(self.h, ctypes.c_void_p(base.ctypes.data), ctypes.byref(total)), "vef_mf_layout")
        if total.value != P.total or not np.array_equal(base, P.front_base):
            raise RfLUError("vef_mf_layout differs from the plan's front layout")
        self.dst = torch.from_numpy(value_map(P, rowptr, col)).to(device)
        self.device_bytes = int(L.vef_mf_device_bytes(self.h))
        self.pivoting = bool(pivoting)
        self.pivot_fallbacks = 0
        self._pivoted = self.pivoting  # the last factorisation's
        self._err(L.vef_mf_set_pivoting(self.h, 1 if self.pivoting else 0), "vef_mf_set_pivoting")

    def close(self):
        if getattr(self, "h", None):
            _mf_lib().vef_mf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc, what):
        if rc != 0:
            from .linsolve import RfLUError

            raise RfLUError(f"{what}: {_mf_lib().vef_mf_last_error(self.h).decode()}")

    def _factor11(self, values):
        L = _mf_lib()
        st = self._stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        self._err(L.vef_mf_set(self.h, int(self.dst.numel()), p(self.dst), p(values), st), "vef_mf_set")
        self._err(L.vef_mf_factor(self.h, st), "vef_mf_factor")

    def factor(self, values):
        """Factorises the matrix with the given device values (same pattern),
        with the configured pivoting."""
        if self._pivoted != self.pivoting:
            self._err(_mf_lib().vef_mf_set_pivoting(self.h, 1 if self.pivoting else 0), "vef_mf_set_pivoting")
            self._pivoted = self.pivoting
        self._unpivoted = not self._pivoted
        super().factor(values)

    def _refactor_pivoted(self):
        """The last matrix once more, with row interchanges (a solve on the
        factorisation without them failed its check)."""
        self._err(_mf_lib().vef_mf_set_pivoting(self.h, 1), "vef_mf_set_pivoting")
        self._pivoted = True
        self._unpivoted = False
        self.pivot_fallbacks += 1
        super().factor(self._A)

    def _solve11(self, y):
        """A11^-1 y in the nested-dissection order (y is consumed)."""
        x = self.torch.empty_like(y)
        self._err(_mf_lib().vef_mf_solve(self.h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                         self._stream()), "vef_mf_solve")
        return x

    def zero_pivots(self) -> int:
        """Fronts whose last factorisation met a zero pivot (synchronises)."""
        c = ctypes.c_int32()
        self._err(_mf_lib().vef_mf_zero_pivots(self.h, ctypes.byref(c)), "vef_mf_zero_pivots")
        return int(c.value)

    def profile(self, enable: bool = True):
        """Per-height timing of the following factorisations (HIP events)."""
        self._err(_mf_lib().vef_mf_profile(self.h, 1 if enable else 0), "vef_mf_profile")

    def profile_read(self) -> np.ndarray:
        """[height][extend-add ms, buckets ms] of the last factorisation."""
        L = _mf_lib()
        ms = np.zeros(2 * int(L.vef_mf_heights(self.h)), dtype=np.float32)
        self._err(L.vef_mf_profile_read(self.h, ctypes.c_void_p(ms.ctypes.data)), "vef_mf_profile_read")
        return ms.reshape(-1, 2)


class NodeBlockLU:
    """Multifrontal LU (FP64) of a square matrix on the node unknowns alone
    (no arclength border), e.g. the shifted operator A - sigma B of the
    linear stability problem.  factor(values), solve(b) with device vectors
    in the pattern's order; the first solve after a factorisation is checked
    by its normwise backward error."""

    def __init__(self, rowptr, col, coords, conn, device="cuda:0", leaf: int = 4, plan: MultifrontalPlan = None):
        import torch

        from .linsolve import RfLUError

        self.torch = torch
        self.device = device
        self.plan = plan if plan is not None else build_plan(coords, conn, leaf)
        P = self.plan
        self.n = len(rowptr) - 1
        if self.n != P.n:
            raise RfLUError("NodeBlockLU: the pattern and the mesh differ")
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)  # noqa: E731
        keep = [i32(P.bucket_height), i32(P.bucket_count), i32(P.bucket_s), i32(P.bucket_u), i32(P.front_s),
                i32(P.front_u), i64(P.rows_off), i32(P.rows), i32(P.parent), i32(P.slot), i64(P.ea_off),
                i32(P.ea_map)]
        ptr = [ctypes.c_void_p(a.ctypes.data) for a in keep]
        cp = _Plan(P.n, len(P.bucket_count), ptr[0], ptr[1], ptr[2], ptr[3], P.n_fronts, ptr[4], ptr[5], ptr[6],
                   ptr[7], ptr[8], ptr[9], ptr[10], ptr[11])
        L = _mf_lib()
        self.h = ctypes.c_void_p()
        if L.vef_mf_create(ctypes.byref(cp), 64, ctypes.byref(self.h)) != 0:
            msg = L.vef_mf_last_error(self.h).decode() if self.h else "bad arguments"
            self.close()
            raise RfLUError(f"vef_mf_create: {msg}")
        self.dst = torch.from_numpy(value_map(P, rowptr, col)).to(device)
        self.pu = torch.from_numpy(P.unknown_pos).to(device)
        self.ptrA = torch.from_numpy(np.asarray(rowptr, dtype=np.int64)).to(device)
        self.indA = torch.from_numpy(np.asarray(col, dtype=np.int64)).to(device)
        self._A = None
        self.backward_errors = []

    close = MultifrontalLU.close
    __del__ = MultifrontalLU.__del__
    _err = MultifrontalLU._err
    zero_pivots = MultifrontalLU.zero_pivots

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def factor(self, values):
        L = _mf_lib()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        st = self._stream()
        self._err(L.vef_mf_set(self.h, int(self.dst.numel()), p(self.dst), p(values), st), "vef_mf_set")
        self._err(L.vef_mf_factor(self.h, st), "vef_mf_factor")
        self._A = values
        self._Anorm = float(self.torch.linalg.vector_norm(values))
        self._checked = False

    def solve(self, b):
        torch = self.torch
        y = torch.empty_like(b)
        y[self.pu] = b
        x = torch.empty_like(b)
        self._err(_mf_lib().vef_mf_solve(self.h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                         self._stream()), "vef_mf_solve")
        x = x[self.pu]
        if not self._checked:
            A = torch.sparse_csr_tensor(self.ptrA, self.indA, self._A, size=(self.n, self.n))
            rn = float(torch.linalg.vector_norm(torch.mv(A, x) - b))
            scale = self._Anorm * float(torch.linalg.vector_norm(x)) + float(torch.linalg.vector_norm(b))
            self.backward_errors.append(rn / max(scale, 1e-300))
            self._checked = True
        return x
