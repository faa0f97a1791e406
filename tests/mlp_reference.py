"""CPU reference of the tile classifier's arithmetic (test infrastructure, the
checker of the HIP kernels in baby_shark_amd/csrc/surrogate.hip).

The network is 3 -> 64 -> 64 -> 64 -> 1 with ReLU, parameters in the order of
include/bs_voxel.h (W1[64][3], b1, W2[64][64], b2, W3[64][64], b3, W4[64],
b4; W[out][in]).  `forward` / `gradient` use the kernels' rounding points:
weights and activations rounded to BF16 (round to nearest even), the input
layer summed in float32 in the kernels' order, the 64-wide products summed
exactly here (float64) where the kernels accumulate in FP32 on the matrix
cores, and the backward pass's output-layer and hidden-layer error terms
rounded to BF16 before they enter a matrix product.  With bf16=False every
step is plain float64 (for the finite-difference check of the formulas)."""
import numpy as np

N_PARAMS = 8641
_LAYOUT = [("W1", (64, 3)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
           ("W4", (64,)), ("b4", ())]


def bf16(a):
    """float32 values rounded to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    out, off = {}, 0
    for name, shape in _LAYOUT:
        n = int(np.prod(shape)) if shape else 1
        out[name] = p[off:off + n].reshape(shape) if shape else p[off]
        off += n
    assert off == N_PARAMS
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in _LAYOUT])


def forward(p, x, shift=(0.0, 0.0, 0.0), scale=1.0, rounded=True):
    """Network outputs for the (n, 3) inputs and the intermediate values."""
    P = unpack(np.asarray(p, np.float32 if rounded else np.float64))
    if rounded:
        r = bf16
        xn = (np.asarray(x, np.float32) - np.asarray(shift, np.float32)) * np.float32(scale)
        w1 = r(P["W1"])
        z1 = np.broadcast_to(P["b1"].astype(np.float32), (xn.shape[0], 64)).copy()
        for c in range(3):  # the kernels' order: ((b + w0 x0) + w1 x1) + w2 x2, float32
            z1 = (z1 + w1[:, c][None, :] * xn[:, c:c + 1]).astype(np.float32)
    else:
        def r(a):
            return np.asarray(a, np.float64)
        xn = (np.asarray(x, np.float64) - np.asarray(shift, np.float64)) * float(scale)
        z1 = P["b1"][None, :] + xn @ P["W1"].T
    h1 = r(np.maximum(z1, 0))
    z2 = (h1.astype(np.float64) @ r(P["W2"]).astype(np.float64).T + P["b2"]).astype(np.float32 if rounded else np.float64)
    h2 = r(np.maximum(z2, 0))
    z3 = (h2.astype(np.float64) @ r(P["W3"]).astype(np.float64).T + P["b3"]).astype(np.float32 if rounded else np.float64)
    h3 = r(np.maximum(z3, 0))
    out = h3.astype(np.float64) @ r(P["W4"]).astype(np.float64) + float(P["b4"])
    return out, dict(xn=xn, h1=h1, h2=h2, h3=h3, P=P, r=r)


def gradient(p, x, y, shift=(0.0, 0.0, 0.0), scale=1.0, rounded=True):
    """(gradient of mean((net(x) - y)^2) in parameter order, that loss)."""
    out, c = forward(p, x, shift, scale, rounded)
    P, r = c["P"], c["r"]
    h1, h2, h3 = (c[k].astype(np.float64) for k in ("h1", "h2", "h3"))
    xn = c["xn"].astype(np.float64)
    e = out - np.asarray(y, np.float64)
    n = e.shape[0]
    g = 2.0 * e / n
    d = {"W4": g @ h3, "b4": g.sum()}
    dz3 = (h3 > 0) * (g[:, None] * r(P["W4"]).astype(np.float64)[None, :])
    d["b3"] = dz3.sum(0)
    dz3b = r(dz3.astype(np.float32) if rounded else dz3).astype(np.float64)
    d["W3"] = dz3b.T @ h2
    dz2 = (h2 > 0) * (dz3b @ r(P["W3"]).astype(np.float64))
    d["b2"] = dz2.sum(0)
    dz2b = r(dz2.astype(np.float32) if rounded else dz2).astype(np.float64)
    d["W2"] = dz2b.T @ h1
    dz1 = (h1 > 0) * (dz2b @ r(P["W2"]).astype(np.float64))
    d["b1"] = dz1.sum(0)
    d["W1"] = dz1.T @ xn
    return pack(d), float(np.mean(e ** 2))
