"""CPU oracle of the learned WENO weights: the 5 -> 32 -> 32 -> 3 perceptron,
its training loss and gradient, and the learned reconstruction / operator.

TEST INFRASTRUCTURE ONLY (imported by ``tests/`` and nothing in the product
path). The reference has no learned weights; this restates the semantics the
HIP kernels implement, in FP64, optionally emulating the kernels' BF16
rounding points (inputs, weights and hidden activations rounded to BF16,
round-to-nearest-even, everything else in FP64 where the kernels use FP32
accumulators):

* inputs of a stencil ``v``: ``mean = (((v0 + v4) + (v1 + v3)) + v2) * 0.2``,
  ``s = max |v - mean|``, ``x = (v - mean) * (1 / s)`` (0 if ``s == 0``), in the
  state dtype, then FP32;
* ``h1 = relu(W1 x + b1)``, ``h2 = relu(W2 h1 + b2)``,
  ``z = W3 h2 + b3 + ln d``, ``omega = softmax(z)``;
* the value is ``h1 + omega0 (h0 - h1) + omega2 (h2 - h1)`` of the three
  candidates ``h`` (the weighted sum, insensitive to the FP32 rounding of
  ``sum omega``);
* a cell with ``max beta > tau (min beta + eps)`` keeps WENO-JS, one with
  ``max beta <= sigma min beta`` takes the ideal weights ``d``.

Pins (``tests/test_oracle_mlp.py``): the unrounded gradient against central
finite differences of the unrounded loss; softmax invariants (weights sum to
1, ``z = 0`` gives the ideal weights); mirror symmetry of the inputs; flagged
cells equal to the WENO-JS reconstruction of ``weno_np``; with the weights
of the ideal-weight network (all parameters 0) the learned reconstruction is
the linear 5th-order interpolant.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from oracle import weno_np

NPARAM = 1347
NIN, NHID, NOUT = 5, 32, 3
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 160, 192, 1216, 1248, 1344
RECORD = 12
IDEAL = np.array([0.1, 0.6, 0.3])
LN_D32 = np.log(IDEAL).astype(np.float32)


def unpack(p: np.ndarray) -> dict[str, np.ndarray]:
    p = np.asarray(p)
    if p.dtype != np.float64:
        p = p.astype(np.float32)
    assert p.shape == (NPARAM,)
    return {
        "W1": p[OFF_W1:OFF_B1].reshape(NHID, NIN),
        "b1": p[OFF_B1:OFF_W2],
        "W2": p[OFF_W2:OFF_B2].reshape(NHID, NHID),
        "b2": p[OFF_B2:OFF_W3],
        "W3": p[OFF_W3:OFF_B3].reshape(NOUT, NHID),
        "b3": p[OFF_B3:NPARAM],
    }


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]
                          ).astype(np.float32)


def bf16(x: Any) -> np.ndarray:
    """Round to BF16 (nearest, ties to even) through FP32; returns FP32."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).reshape(f.shape)


def features(v: np.ndarray) -> np.ndarray:
    """Inputs of stencils ``v`` (..., 5) in their own dtype, returned as FP32."""
    v = np.asarray(v)
    dt = v.dtype.type
    mean = (((v[..., 0] + v[..., 4]) + (v[..., 1] + v[..., 3])) + v[..., 2]) * dt(0.2)
    d = v - mean[..., None]
    s = np.max(np.abs(d), axis=-1)
    inv = dt(1) / np.where(s > 0, s, dt(1))
    x = np.where(s[..., None] > 0, d * inv[..., None], dt(0))
    return x.astype(np.float32)


def _softmax(z: np.ndarray) -> np.ndarray:
    z = z - np.max(z, axis=-1, keepdims=True)
    e = np.exp(z)
    return e / np.sum(e, axis=-1, keepdims=True)


def forward(p: np.ndarray, x: np.ndarray, *, emulate: bool = True) -> dict[str, np.ndarray]:
    """Logits and weights of inputs ``x`` (n, 5); with ``emulate`` the kernels'
    BF16 rounding points are applied."""
    P = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float32)
    xt = r(x).astype(np.float64)
    W1, W2, W3 = (r(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    a1 = xt @ W1.T + P["b1"].astype(np.float64)
    h1 = r(np.maximum(a1, 0.0)).astype(np.float64)
    a2 = h1 @ W2.T + P["b2"].astype(np.float64)
    h2 = r(np.maximum(a2, 0.0)).astype(np.float64)
    if emulate:
        c3 = (P["b3"] + LN_D32).astype(np.float64)
    else:
        c3 = P["b3"].astype(np.float64) + np.log(IDEAL)
    z = h2 @ W3.T + c3
    return {"x": xt, "h1": h1, "h2": h2, "z": z, "omega": _softmax(z),
            "W1": W1, "W2": W2, "W3": W3}


def loss_and_grad(p: np.ndarray, rec: np.ndarray, *, lambda_js: float, delta: float,
                  emulate: bool = True) -> tuple[float, np.ndarray]:
    """Mean loss over sample records ``rec`` (n, 12) and its gradient with
    respect to the packed parameters (``emulate`` rounds the operands of the
    backward products through the hidden layers to BF16, as the kernel does)."""
    rec = np.asarray(rec, dtype=np.float32)
    n = rec.shape[0]
    x, e, wjs, c = rec[:, :5], rec[:, 5:8].astype(np.float64), \
        rec[:, 8:11].astype(np.float64), rec[:, 11].astype(np.float64)
    f = forward(p, x, emulate=emulate)
    w = f["omega"]
    r = np.sum(w * e, axis=1)
    den = np.mean(e * e, axis=1) + delta
    inv = np.where(den > 0, 1.0 / np.where(den > 0, den, 1.0), 0.0)
    dj = w - wjs
    loss = np.sum(c * (r * r * inv + lambda_js * np.sum(dj * dj, axis=1))) / n
    g = (2.0 * c * r * inv)[:, None] * e + (2.0 * c * lambda_js)[:, None] * dj
    dz = w * (g - np.sum(w * g, axis=1, keepdims=True))

    rb = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    dh2 = rb(dz).astype(np.float64) @ f["W3"]
    dh2 = np.where(f["h2"] > 0, dh2, 0.0)
    dh1 = rb(dh2).astype(np.float64) @ f["W2"]
    dh1 = np.where(f["h1"] > 0, dh1, 0.0)

    grad = np.concatenate([np.ravel(dh1.T @ f["x"]), dh1.sum(0), np.ravel(dh2.T @ f["h1"]),
                           dh2.sum(0), np.ravel(dz.T @ f["h2"]), dz.sum(0)])
    return float(loss), grad / n


def stencils(u: np.ndarray) -> np.ndarray:
    """Right-side stencils ``u[i-2 .. i+2]`` of every cell with zero padding
    outside ``[0, m)`` (the 'same' convolution of ``convolve.py:113-114``)."""
    up = np.pad(u, 2)
    m = u.size
    return np.stack([up[k: k + m] for k in range(5)], axis=1)


def regimes(v: np.ndarray, eps: float, tau: float, sigma: float) -> np.ndarray:
    """0 (shock: WENO-JS) where ``max beta > tau (min beta + eps)``, 1
    (smooth: ideal weights) where ``max beta <= sigma min beta``, 2 (learned)
    otherwise, over the right-side indicators."""
    b = _betas(v)
    dt = v.dtype.type
    bmax = np.max(b, axis=1)
    bmin = np.min(b, axis=1)
    return np.where(bmax > dt(tau) * (bmin + dt(eps)), 0,
                    np.where(bmax <= dt(sigma) * bmin, 1, 2))


def _betas(v: np.ndarray) -> np.ndarray:
    dt = v.dtype.type
    a0, a1 = dt(13.0 / 12.0), dt(1.0 / 4.0)
    um2, um1, u0, up1, up2 = (v[:, k] for k in range(5))
    t00 = (u0 - dt(2) * um1) + um2
    t01 = (dt(3) * u0 - dt(4) * um1) + um2
    t10 = (up1 - dt(2) * u0) + um1
    t11 = up1 - um1
    t20 = (up2 - dt(2) * up1) + u0
    t21 = (up2 - dt(4) * up1) + dt(3) * u0
    return np.stack([a0 * t00 * t00 + a1 * t01 * t01, a0 * t10 * t10 + a1 * t11 * t11,
                     a0 * t20 * t20 + a1 * t21 * t21], axis=1)


def shock_flags(v: np.ndarray, eps: float, tau: float) -> np.ndarray:
    """``max beta > tau (min beta + eps)`` from the right-side indicators."""
    dt = v.dtype.type
    a0, a1 = dt(13.0 / 12.0), dt(1.0 / 4.0)
    um2, um1, u0, up1, up2 = (v[:, k] for k in range(5))
    t00 = (u0 - dt(2) * um1) + um2
    t01 = (dt(3) * u0 - dt(4) * um1) + um2
    t10 = (up1 - dt(2) * u0) + um1
    t11 = up1 - um1
    t20 = (up2 - dt(2) * up1) + u0
    t21 = (up2 - dt(4) * up1) + dt(3) * u0
    b = np.stack([a0 * t00 * t00 + a1 * t01 * t01, a0 * t10 * t10 + a1 * t11 * t11,
                  a0 * t20 * t20 + a1 * t21 * t21], axis=1)
    return np.max(b, axis=1) > dt(tau) * (np.min(b, axis=1) + dt(eps))


def candidates(v: np.ndarray) -> np.ndarray:
    """The three candidate values of the right reconstruction (weno.py:143-157)."""
    um2, um1, u0, up1, up2 = (v[:, k] for k in range(5))
    return np.stack([11 / 6 * u0 - 7 / 6 * um1 + 2 / 6 * um2,
                     2 / 6 * up1 + 5 / 6 * u0 - 1 / 6 * um1,
                     -1 / 6 * up2 + 5 / 6 * up1 + 2 / 6 * u0], axis=1)


def _combine(w: np.ndarray, h: np.ndarray) -> np.ndarray:
    """``sum_k w_k h_k`` as ``h1 + w0 (h0 - h1) + w2 (h2 - h1)``: equal when
    the weights sum to 1, and insensitive to their rounding otherwise."""
    return (h[:, 1] + w[:, 0] * (h[:, 0] - h[:, 1])) + w[:, 2] * (h[:, 2] - h[:, 1])


def learned_reconstruct(p: np.ndarray, tau: float, u: np.ndarray, eps: float, *,
                        sigma: float = 0.0, emulate: bool = True
                        ) -> tuple[np.ndarray, np.ndarray]:
    """``(ul, ur)`` of every cell: WENO-JS in the shock regime, the ideal
    weights in the smooth regime, the network's weights otherwise."""
    u = np.asarray(u)
    v = stencils(u)
    vm = v[:, ::-1]
    x = features(v)
    wr = forward(p, x, emulate=emulate)["omega"]
    wl = forward(p, x[:, ::-1], emulate=emulate)["omega"]
    cr, cl = candidates(v.astype(np.float64)), candidates(vm.astype(np.float64))
    ideal = np.broadcast_to(IDEAL, wr.shape)
    reg = regimes(v, eps, tau, sigma)
    ur = np.where(reg == 1, _combine(ideal, cr), _combine(wr, cr))
    ul = np.where(reg == 1, _combine(ideal, cl), _combine(wl, cl))
    jl, jr = weno_np.reconstruct_wenojs(u.astype(np.float64), eps)
    return np.where(reg == 0, jl, ul), np.where(reg == 0, jr, ur)


def learned_apply_operator(p: np.ndarray, tau: float, kind: str, grid: weno_np.Grid,
                           u: np.ndarray, eps: float, *, sigma: float = 0.0,
                           emulate: bool = True) -> np.ndarray:
    """The Burgers operator (``schemes.py:466-473``) with the learned
    reconstruction in place of WENO-JS: Rusanov (alpha = 1), Godunov,
    Lax-Friedrichs (the global speed max|u| of the ghost-filled state,
    ``scalar.py:258-278``) or Engquist-Osher (omega = 0,
    ``scalar.py:287-322``)."""
    u = weno_np.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    ul, ur = learned_reconstruct(p, tau, u, eps, sigma=sigma, emulate=emulate)
    fl, fr = ul**2 / 2, ur**2 / 2
    if kind in ("rusanov", "lf"):
        if kind == "rusanov":
            a = np.abs(u)
            a = np.maximum(a[1:], a[:-1])
        else:
            a = np.max(np.abs(u))
        fnum = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * (ul[1:] - ur[:-1])
    elif kind == "godunov":
        aavg = (ur[:-1] + ul[1:]) / 2
        fnum = np.where(aavg > 0, fr[:-1], fl[1:])
    elif kind == "eo":
        fnum = weno_np.engquist_osher_faces(ur[:-1], ul[1:])
    else:
        raise NotImplementedError(kind)
    f = np.pad(fnum, 1)
    return -(f[1:] - f[:-1]) / grid.dx


def ideal_params() -> np.ndarray:
    """All-zero parameters: z = ln d, so omega = d everywhere."""
    return np.zeros(NPARAM, dtype=np.float32)
