"""CPU restatement of the surrogate MLP (35 -> 256 -> 256 -> 256 -> 4, ReLU) in float64.

TEST INFRASTRUCTURE ONLY (tests/ import it to check the HIP training/inference kernels).
No reference counterpart exists: the surrogate is a build-only component.

bf16=True reproduces the GPU arithmetic model -- BF16-rounded inputs, weights,
activations and output gradients, exact products, wide accumulation -- so the GPU
gradients can be compared tightly; bf16=False is the plain FP64 model whose gradients
are checked against finite differences.
"""
import numpy as np

NIN, K0, HID, NOUT, NOP = 35, 64, 256, 4, 32
LK = [K0, HID, HID, HID]
LN = [HID, HID, HID, NOP]


def param_slices():
    out, o = [], 0
    for L in range(4):
        w = (o, o + LK[L] * LN[L])
        o = w[1]
        b = (o, o + LN[L])
        o = b[1]
        out.append((w, b))
    return out, o


def bf16(x):
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    sl, n = param_slices()
    assert len(params) == n
    Ws = [np.asarray(params[a:b], np.float64).reshape(LK[L], LN[L]) for L, ((a, b), _) in enumerate(sl)]
    bs = [np.asarray(params[a:b], np.float64) for (_, (a, b)) in sl]
    return Ws, bs


def embed(genomes, lo, hi):
    g = np.asarray(genomes, np.float32).reshape(-1, NIN)
    x = np.zeros((len(g), K0), np.float32)
    x[:, :NIN] = (g - lo) * (np.float32(1.0) / (hi - lo)).astype(np.float32)
    return x.astype(np.float64)


def forward(params, genomes, lo, hi, bf=True):
    Ws, bs = unpack(params)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    X = r(embed(genomes, lo, hi))
    acts = [X]
    h = X
    for L in range(3):
        h = r(np.maximum(h @ r(Ws[L]) + bs[L], 0.0))
        acts.append(h)
    out = h @ r(Ws[3]) + bs[3]
    return out, acts


def loss_grad(params, genomes, targets, lo, hi, bf=True):
    """Mean loss over the samples and its gradient (parameter layout of the HIP library)."""
    Ws, bs = unpack(params)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    out, acts = forward(params, genomes, lo, hi, bf)
    t = np.asarray(targets, np.float64).reshape(-1, NOUT)
    n = len(t)
    over = t[:, 3]
    live = over < 0.5
    d = np.zeros_like(out)
    e = out[:, :3] - 0.1 * t[:, :3]
    loss = np.sum((e ** 2)[live])
    d[:, :3] = np.where(live[:, None], 2.0 * e / n, 0.0)
    z = out[:, 3]
    p = 1.0 / (1.0 + np.exp(-z))
    loss += np.sum(np.maximum(z, 0) - z * over + np.log1p(np.exp(-np.abs(z))))
    d[:, 3] = (p - over) / n
    loss /= n
    grads_W, grads_b = [None] * 4, [None] * 4
    dY = r(d)
    for L in range(3, -1, -1):
        Xin = acts[L]
        grads_W[L] = Xin.T @ dY
        grads_b[L] = dY.sum(0)
        if L > 0:
            dY = r((dY @ r(Ws[L]).T) * (Xin > 0))
    grads_W[0][NIN:, :] = 0.0
    grads_W[3][:, NOUT:] = 0.0
    grads_b[3][NOUT:] = 0.0
    flat = np.concatenate([np.concatenate([grads_W[L].ravel(), grads_b[L]]) for L in range(4)])
    return loss, flat


def fitness_targets(samples):
    """(genomes x seeds) flight distances -> [mean, CVaR of the worst quarter, worst,
    oversize flag] with the reference's float accumulation order (sim.cu:340-357)."""
    s = np.asarray(samples, np.float32)
    n, k = s.shape
    mean = np.zeros(n, np.float32)
    for j in range(k):
        mean = (mean + s[:, j] / np.float32(k)).astype(np.float32)
    srt = np.sort(s, axis=1)
    tail = (k + 3) // 4
    cv = np.zeros(n, np.float32)
    for j in range(tail):
        cv = (cv + srt[:, j] / np.float32(tail)).astype(np.float32)
    over = np.all(s == -100.0, axis=1).astype(np.float32)
    return np.stack([mean, cv, srt[:, 0], over], 1).astype(np.float32)
