"""CPU restatement (numpy) of the learned preconditioner's network, its loss and its gradient.

TEST INFRASTRUCTURE ONLY: tests/ load it as the checker of the GPU kernels in
ffea_amd/csrc/ffh_mpnn.h; nothing in the product path imports it.

The network has no counterpart in the reference (its massless solver is Jacobi-PCG only,
src/NoMassCGSolver.cpp:165-197); what it must agree with is the definition in
include/ffea_hip.h / ffh_mpnn.h.  Two modes:
- exact=True:  float64 throughout -- the mathematical definition; `loss_grad`'s backward pass
  is pinned against central finite differences of `loss` in this mode (tests).
- exact=False: the GPU's rounding points -- fp32 embedding/output arithmetic in the kernels'
  order, bf16 storage of h and g, bf16 weights and operands of the dense layers (accumulated in
  float64 here, fp32 on the MFMA: the residual difference is within the tests' tolerance).

Parameter layout (6499 floats): We[32][4], be[32]; per layer l = 0..2: W1[32][32], W2[32][32],
b[32]; Wo[3][32], bo[3].
"""
from __future__ import annotations

import numpy as np

HID, NIN, NOUT, NLAY = 32, 4, 3, 3
OFF_WE = 0
OFF_BE = OFF_WE + HID * NIN
OFF_LAY = OFF_BE + HID
LAY_STRIDE = 2 * HID * HID + HID
OFF_WO = OFF_LAY + NLAY * LAY_STRIDE
OFF_BO = OFF_WO + NOUT * HID
NUM_PARAMS = OFF_BO + NOUT


def unpack(w):
    w = np.asarray(w)
    p = {"We": w[OFF_WE:OFF_BE].reshape(HID, NIN), "be": w[OFF_BE:OFF_LAY],
         "Wo": w[OFF_WO:OFF_BO].reshape(NOUT, HID), "bo": w[OFF_BO:NUM_PARAMS]}
    for l in range(NLAY):
        o = OFF_LAY + l * LAY_STRIDE
        p[f"W1_{l}"] = w[o:o + HID * HID].reshape(HID, HID)
        p[f"W2_{l}"] = w[o + HID * HID:o + 2 * HID * HID].reshape(HID, HID)
        p[f"b_{l}"] = w[o + 2 * HID * HID:o + LAY_STRIDE]
    return p


def pack(p):
    w = np.zeros(NUM_PARAMS)
    w[OFF_WE:OFF_BE] = p["We"].ravel()
    w[OFF_BE:OFF_LAY] = p["be"]
    for l in range(NLAY):
        o = OFF_LAY + l * LAY_STRIDE
        w[o:o + HID * HID] = p[f"W1_{l}"].ravel()
        w[o + HID * HID:o + 2 * HID * HID] = p[f"W2_{l}"].ravel()
        w[o + 2 * HID * HID:o + LAY_STRIDE] = p[f"b_{l}"]
    w[OFF_WO:OFF_BO] = p["Wo"].ravel()
    w[OFF_BO:] = p["bo"]
    return w


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


class Graph:
    """The linear-node blocks of a matrix given as scalar CSR over 3*N_t rows (the reference's
    layout): neighbours per node in ascending column order, A_ij = tr K_ij / sqrt(tr K_ii tr K_jj)
    (tr K_ii as sum_c 1/(1/K_ii,cc), the GPU's expression), D^-1 and the drag share c_i."""

    def __init__(self, key, col, val, lin_ids, drag, blob_of, pinned):
        lin_ids = np.asarray(lin_ids)
        pos = {int(o): k for k, o in enumerate(lin_ids)}
        n = lin_ids.size
        self.n = n
        self.dinv = np.zeros(3 * n)
        nbr = []
        for k, i in enumerate(lin_ids):
            acc = {}
            for c in range(3):
                r = 3 * int(i) + c
                for p in range(key[r], key[r + 1]):
                    j, cj = divmod(int(col[p]), 3)
                    if j in pos:
                        if cj == c:
                            acc[pos[j]] = acc.get(pos[j], 0.0) + float(val[p])
                        else:
                            acc.setdefault(pos[j], 0.0)
                        if j == int(i) and cj == c:
                            self.dinv[3 * k + c] = 1.0 / float(val[p])
            nbr.append(sorted(acc.items()))
        tri = np.array([sum(1.0 / self.dinv[3 * k + c] for c in range(3)) for k in range(n)])
        self.cols = [np.array([j for j, _ in row], np.int64) for row in nbr]
        self.aw = [np.array([t / np.sqrt(tri[k] * tri[j]) for j, t in row], np.float64).astype(np.float32)
                   for k, row in enumerate(nbr)]
        self.cf = (3.0 * np.asarray(drag, np.float64) / tri).astype(np.float32)
        self.blob_of = np.asarray(blob_of)
        self.nblobs = int(self.blob_of.max()) + 1
        self.nreal = np.bincount(self.blob_of, minlength=self.nblobs)
        self.live = ~np.asarray(pinned, bool)

    def apply(self, h, exact):
        """g_i = sum_j A_ij h_j (fp32 in ascending-column order unless exact)."""
        out = np.zeros_like(h, dtype=np.float64 if exact else np.float32)
        for i in range(self.n):
            if exact:
                out[i] = (self.aw[i].astype(np.float64)[:, None] * h[self.cols[i]]).sum(axis=0)
            else:
                acc = np.zeros(h.shape[1], np.float32)
                for w, j in zip(self.aw[i], self.cols[i]):
                    acc = acc + np.float32(w) * h[j].astype(np.float32)
                out[i] = acc
        return out


def _scale(g, u):
    s = np.zeros(g.nblobs)
    uu = (u.astype(np.float64) ** 2).sum(axis=1)
    for b in range(g.nblobs):
        if g.nreal[b] > 0:
            s[b] = np.sqrt(uu[g.blob_of == b].sum() / g.nreal[b])
    return s


def forward(w, g: Graph, u, exact=False):
    """u = D^-1 r, [n, 3] float64 (CG path).  Returns dict with h0..h3, g1..g3, x, o, s and
    z = u + s_b o (0 on pinned rows)."""
    p = unpack(np.asarray(w, np.float64 if exact else np.float32))
    u = np.asarray(u, np.float64)
    s = _scale(g, u)
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)[g.blob_of]
    if exact:
        x = np.concatenate([u * inv[:, None], g.cf.astype(np.float64)[:, None]], axis=1)
        h = np.maximum(x @ p["We"].T + p["be"], 0.0)
    else:
        x = np.concatenate([(u * inv[:, None]).astype(np.float32), g.cf[:, None]], axis=1)
        acc = np.tile(p["be"], (g.n, 1)).astype(np.float32)
        for c in range(NIN):
            acc = acc + p["We"][:, c][None, :] * x[:, c][:, None]
        h = bf16(np.maximum(acc, np.float32(0)))
    out = {"x": x, "h0": h, "s": s}
    for l in range(NLAY):
        gl = g.apply(h, exact)
        if exact:
            pre = h @ p[f"W1_{l}"].T + gl @ p[f"W2_{l}"].T + p[f"b_{l}"]
            h = np.maximum(pre, 0.0)
        else:
            gl = bf16(gl)
            pre = (h.astype(np.float64) @ bf16(p[f"W1_{l}"]).T.astype(np.float64)
                   + gl.astype(np.float64) @ bf16(p[f"W2_{l}"]).T.astype(np.float64)).astype(np.float32) + p[f"b_{l}"]
            h = bf16(np.maximum(pre, np.float32(0)))
        out[f"g{l + 1}"] = gl
        out[f"h{l + 1}"] = h
    if exact:
        o = h @ p["Wo"].T + p["bo"]
    else:
        o = np.tile(p["bo"], (g.n, 1)).astype(np.float32)
        for j in range(HID):
            o = o + p["Wo"][:, j][None, :] * h[:, j][:, None]
    out["o"] = o
    z = u + s[g.blob_of][:, None] * o.astype(np.float64)
    z[~g.live] = 0.0
    z[s[g.blob_of] == 0] = 0.0
    out["z"] = z
    return out


def loss(w, g: Graph, r, e, exact=True):
    return loss_grad(w, g, r, e, exact, want_grad=False)[0]


def loss_grad(w, g: Graph, r, e, exact=True, want_grad=True):
    """Training loss of one snapshot (residual r, correction e = x* - x_k; [n, 3] float32 on the
    GPU) and its gradient over the parameters."""
    p = unpack(np.asarray(w, np.float64 if exact else np.float32))
    r = np.asarray(r, np.float64)
    e = np.asarray(e, np.float64)
    if exact:
        u = r * g.dinv.reshape(-1, 3)
    else:
        r = r.astype(np.float32).astype(np.float64)
        e = e.astype(np.float32).astype(np.float64)
        u = (r * g.dinv.reshape(-1, 3)).astype(np.float32).astype(np.float64)
    f = forward(w, g, u, exact)
    s = f["s"]
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)[g.blob_of][:, None]
    t = e * inv - u * inv
    if not exact:
        t = t.astype(np.float32).astype(np.float64)
    den = np.zeros(g.nblobs)
    for b in range(g.nblobs):
        den[b] = ((e * inv)[g.blob_of == b] ** 2).sum()
    live = g.live & (den[g.blob_of] > 0)
    o = f["o"].astype(np.float64)
    d = o - t
    L = 0.0
    for b in range(g.nblobs):
        m = live & (g.blob_of == b)
        if den[b] > 0:
            L += (d[m] ** 2).sum() / den[b]
    L /= g.nblobs
    if not want_grad:
        return L, None
    dO = np.where(live[:, None], 2.0 * d / np.where(den > 0, den, 1.0)[g.blob_of][:, None] / g.nblobs, 0.0)
    grad = {"Wo": dO.T @ f["h3"].astype(np.float64), "bo": dO.sum(axis=0)}
    dh = dO @ p["Wo"].astype(np.float64)
    for l in reversed(range(NLAY)):
        hl, hprev, gl = f[f"h{l + 1}"], f[f"h{l}"], f[f"g{l + 1}"]
        dP = dh * (hl > 0)
        dPo = dP if exact else bf16(dP.astype(np.float32)).astype(np.float64)
        grad[f"W1_{l}"] = dPo.T @ hprev.astype(np.float64)
        grad[f"W2_{l}"] = dPo.T @ gl.astype(np.float64)
        grad[f"b_{l}"] = dP.sum(axis=0)
        W1 = p[f"W1_{l}"].astype(np.float64) if exact else bf16(p[f"W1_{l}"]).astype(np.float64)
        W2 = p[f"W2_{l}"].astype(np.float64) if exact else bf16(p[f"W2_{l}"]).astype(np.float64)
        dX1 = dPo @ W1
        dX2 = dPo @ W2
        dh = dX1 + g.apply(dX2, True)  # A is symmetric
        dh = dh * (hprev > 0)
    grad["We"] = dh.T @ f["x"].astype(np.float64)
    grad["be"] = dh.sum(axis=0)
    return L, pack(grad)


def graph_from_oracle(blob, num_nodes, topo, pinned, blob_of_node=None):
    """Graph of an oracle blob (pyoracle.Blob): its assembled CSR, linear nodes in ascending
    original index (the GPU's row order), drag from the blob."""
    key, col, val = blob.csr()
    lin = np.unique(np.asarray(topo)[:, :4])
    drag = blob.get("stokes_drag")[lin]
    if blob_of_node is None:
        blob_of_node = np.zeros(num_nodes, np.int64)
    pin = np.zeros(num_nodes, bool)
    pin[np.asarray(pinned, np.int64)] = True
    return Graph(key, col, val, lin, drag, np.asarray(blob_of_node)[lin], pin[lin]), lin
