"""FP64 CPU restatement of the warm-start MLP (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker for the BF16 MFMA kernels in
dg-bgk_amd/csrc/mlp.hip.  With bf16=True it rounds weights, inputs and
stored activations to BF16 (round to nearest even) exactly where the device
kernels do, so the two differ only by FP32-vs-FP64 accumulation.

Parameter packing (as the device): for each layer l, W_l [dout][din] then
b_l [dout]; layers: fin->width (SiLU), (hidden-1) x width->width (SiLU),
width->fout.  Loss: mean over rows and outputs of (y - t)^2.
"""
import numpy as np


def bf16_round(x):
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def shapes(fin, width, hidden, fout):
    out = []
    for l in range(hidden + 1):
        di = fin if l == 0 else width
        do = fout if l == hidden else width
        out.append((do, di))
    return out


def unpack(p, fin, width, hidden, fout):
    ws, o = [], 0
    for do, di in shapes(fin, width, hidden, fout):
        W = p[o:o + do * di].reshape(do, di)
        o += do * di
        b = p[o:o + do]
        o += do
        ws.append((W, b))
    assert o == len(p)
    return ws


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss_grad(p, x, t, fin, width, hidden, fout, bf16=False):
    r = (lambda a: bf16_round(a)) if bf16 else (lambda a: np.asarray(a, np.float64))
    ws = unpack(np.asarray(p, np.float64), fin, width, hidden, fout)
    Wr = [(r(W), np.asarray(b, np.float64)) for W, b in ws]
    a = r(x)
    acts, zs = [a], []
    for l in range(hidden):
        W, b = Wr[l]
        z = a @ W.T + b
        zs.append(r(z))
        a = r(z * sig(z))
        acts.append(a)
    W, b = Wr[hidden]
    y = a @ W.T + b
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / n
    grads = [None] * (hidden + 1)
    grads[hidden] = (dy.T @ acts[hidden], dy.sum(0))
    da = dy @ W
    for l in range(hidden - 1, -1, -1):
        z = zs[l]
        s = sig(z)
        dz = da * (s * (1.0 + z * (1.0 - s)))
        if bf16:
            dz = bf16_round(dz)
        grads[l] = (dz.T @ acts[l], dz.sum(0))
        if l > 0:
            da = dz @ Wr[l][0]
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, g, y


def forward(p, x, fin, width, hidden, fout, bf16=False):
    return loss_grad(p, x, np.zeros((len(x), fout)), fin, width, hidden, fout, bf16)[2]


def init_params(fin, width, hidden, fout, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for do, di in shapes(fin, width, hidden, fout):
        out.append(rng.standard_normal(do * di) * np.sqrt(2.0 / di))
        out.append(np.zeros(do))
    return np.concatenate(out)
