"""Complex-step derivatives of the oracle's SSP-RK3 step.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``). The step of
``oracle/weno_np.py`` is re-evaluated on complex arrays with the two
non-analytic operations made analytic by branching on the real part
(``|z| -> z sign(Re z)``, ``max(|a|, |b|)`` picks the operand with the larger
real modulus, the upwind switch compares real parts). Then
``Im f(u + i h e_j) / h`` with ``h = 1e-30`` is the exact derivative
column up to rounding, which gives a dense Jacobian of one step for small
grids -- the quantity the reference's ``adjoint_step`` builds with
``jax.jacfwd`` (``timestepping.py:174``) -- without finite-difference error.
Ties of ``max`` and zeros of the switch are measure-zero and avoided by the
tests' random data.
"""

from __future__ import annotations

import numpy as np

from oracle import weno_np as W


def _abs_cs(z: np.ndarray) -> np.ndarray:
    return z * np.sign(z.real)


def _reconstruct_cs(f: np.ndarray, eps: float) -> tuple[np.ndarray, np.ndarray]:
    # the reference's reconstruction is analytic (convolutions, powers,
    # divisions), so the oracle's own functions apply to complex input
    return W.reconstruct_wenojs(f, eps)


def apply_operator_cs(kind: str, grid: W.Grid, u: np.ndarray, *, eps: float,
                      velocity: float = 1.0) -> np.ndarray:
    """``apply_operator`` of ``oracle.weno_np`` on complex input."""
    u = W.apply_boundary_periodic(grid, u)
    ul, ur = _reconstruct_cs(u, eps)
    if kind == "rusanov":
        fl = ul**2 / 2
        fr = ur**2 / 2
        a = _abs_cs(u)
        a = np.where(a[1:].real >= a[:-1].real, a[1:], a[:-1])
        fnum = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * (ul[1:] - ur[:-1])
    elif kind == "godunov":
        fl = ul**2 / 2
        fr = ur**2 / 2
        aavg = (ur[:-1] + ul[1:]) / 2
        fnum = np.where(aavg.real > 0, fr[:-1], fl[1:])
    elif kind == "lf":
        # the global speed max|u| over the ghost-filled array: the entry of
        # the largest real modulus (scalar.py:276-277)
        fl = ul**2 / 2
        fr = ur**2 / 2
        au = _abs_cs(u)
        a = au[int(np.argmax(au.real))]
        fnum = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * (ul[1:] - ur[:-1])
    elif kind == "eo":
        # f(max(ur, 0)) + f(min(ul, 0)), the switches on the real parts
        p = np.where(ur[:-1].real > 0, ur[:-1], 0)
        q = np.where(ul[1:].real < 0, ul[1:], 0)
        fnum = p**2 / 2 + q**2 / 2
    elif kind == "advection":
        fnum = ur[:-1] if velocity > 0 else ul[1:]
    else:
        raise NotImplementedError(kind)
    F = np.pad(fnum, 1)
    L = -(F[1:] - F[:-1]) / grid.dx
    if kind == "advection":
        L = velocity * L
    return L


def step_cs(kind: str, grid: W.Grid, u: np.ndarray, dt: float, *, eps: float,
            velocity: float = 1.0) -> np.ndarray:
    """One SSP-RK3 step (``timestepping.py:312-320``) on complex input."""

    def fn(t, x):
        return apply_operator_cs(kind, grid, x, eps=eps, velocity=velocity)

    return W.ssprk33_advance(fn, dt, 0.0, u)


def step_jacobian(kind: str, grid: W.Grid, u: np.ndarray, dt: float, *,
                  eps: float, velocity: float = 1.0, h: float = 1.0e-30) -> np.ndarray:
    """Dense Jacobian of one step restricted to the interior cells:
    ``J[i, j] = d u+_i / d u_j`` for interior ``i, j`` (the periodic ghost
    fill maps every ghost input onto an interior cell)."""
    g = grid.nghosts
    n = grid.n - 2 * g
    J = np.empty((n, n))
    base = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    for j in range(n):
        z = base.astype(np.complex128)
        z[g + j] += 1j * h
        z = W.apply_boundary_periodic(grid, z)
        out = step_cs(kind, grid, z, dt, eps=eps, velocity=velocity)
        J[:, j] = out[g:g + n].imag / h
    return J


def operator_jacobian(kind: str, grid: W.Grid, u: np.ndarray, *, eps: float,
                      velocity: float = 1.0, h: float = 1.0e-30) -> np.ndarray:
    """Dense interior Jacobian of ``apply_operator`` by complex step."""
    g = grid.nghosts
    n = grid.n - 2 * g
    J = np.empty((n, n))
    base = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    for j in range(n):
        z = base.astype(np.complex128)
        z[g + j] += 1j * h
        z = W.apply_boundary_periodic(grid, z)
        out = apply_operator_cs(kind, grid, z, eps=eps, velocity=velocity)
        J[:, j] = out[g:g + n].imag / h
    return J


# {{{ the learned reconstruction's surrogate (complex step)


def _learned_features_cs(v: np.ndarray) -> np.ndarray:
    """``x = (v - mean) / max|v - mean|`` on complex stencils (n, 5), the
    maximum and its sign taken on the real parts (``oracle.mlp_np.features``
    without the FP32 rounding)."""
    mean = (((v[:, 0] + v[:, 4]) + (v[:, 1] + v[:, 3])) + v[:, 2]) * 0.2
    d = v - mean[:, None]
    km = np.argmax(np.abs(d.real), axis=1)
    dm = d[np.arange(d.shape[0]), km]
    s = dm * np.sign(dm.real)
    ok = s.real > 0
    return np.where(ok[:, None], d / np.where(ok, s, 1.0)[:, None], 0.0)


def _learned_weights_cs(p: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Softmax weights of the network on complex inputs: W1, W2, W3 rounded
    to BF16 as the kernels' operands, b3 + ln d summed in FP32 as their
    accumulator start, inputs and activations unrounded, ReLU on the real
    parts."""
    from oracle import mlp_np as M

    P = M.unpack(p)
    W1, W2, W3 = (M.bf16(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    c3 = (P["b3"] + M.LN_D32).astype(np.float32).astype(np.float64)
    a1 = x @ W1.T + P["b1"].astype(np.float64)
    h1 = np.where(a1.real > 0, a1, 0.0)
    a2 = h1 @ W2.T + P["b2"].astype(np.float64)
    h2 = np.where(a2.real > 0, a2, 0.0)
    z = h2 @ W3.T + c3
    z = z - np.max(z.real, axis=1, keepdims=True)
    e = np.exp(z)
    return e / np.sum(e, axis=1, keepdims=True)


def learned_reconstruct_cs(p: np.ndarray, tau: float, sigma: float, u: np.ndarray,
                           eps: float) -> tuple[np.ndarray, np.ndarray]:
    """``(ul, ur)`` of the learned reconstruction's surrogate on complex
    input: WENO-JS where the real stencil is in the shock regime, the ideal
    weights in the smooth regime, the network (``_learned_weights_cs``)
    otherwise; regimes from ``oracle.mlp_np.regimes`` of the real parts."""
    from oracle import mlp_np as M

    v = M.stencils(u)
    reg = M.regimes(v.real, eps, tau, sigma)
    vm = v[:, ::-1]
    wr = _learned_weights_cs(p, _learned_features_cs(v))
    wl = _learned_weights_cs(p, _learned_features_cs(vm))
    cr, cl = M.candidates(v), M.candidates(vm)
    ideal = np.broadcast_to(M.IDEAL, cr.shape)
    jl, jr = W.reconstruct_wenojs(u, eps)
    ur = np.where(reg == 0, jr, np.where(reg == 1, M._combine(ideal, cr), M._combine(wr, cr)))
    ul = np.where(reg == 0, jl, np.where(reg == 1, M._combine(ideal, cl), M._combine(wl, cl)))
    return ul, ur


def learned_operator_cs(p: np.ndarray, tau: float, sigma: float, kind: str,
                        grid: W.Grid, x: np.ndarray, *, eps: float) -> np.ndarray:
    """The Burgers operator (rusanov, godunov or eo) of the learned surrogate
    on complex input (``apply_operator_cs`` with ``learned_reconstruct_cs``)."""
    x = W.apply_boundary_periodic(grid, x)
    ul, ur = learned_reconstruct_cs(p, tau, sigma, x, eps)
    fl, fr = ul**2 / 2, ur**2 / 2
    if kind == "rusanov":
        a = _abs_cs(x)
        a = np.where(a[1:].real >= a[:-1].real, a[1:], a[:-1])
        fnum = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * (ul[1:] - ur[:-1])
    elif kind == "godunov":
        aavg = (ur[:-1] + ul[1:]) / 2
        fnum = np.where(aavg.real > 0, fr[:-1], fl[1:])
    elif kind == "eo":
        # f(max(ur, 0)) + f(min(ul, 0)), the switches on the real parts
        pp = np.where(ur[:-1].real > 0, ur[:-1], 0)
        qq = np.where(ul[1:].real < 0, ul[1:], 0)
        fnum = pp**2 / 2 + qq**2 / 2
    else:
        raise NotImplementedError(kind)
    F = np.pad(fnum, 1)
    return -(F[1:] - F[:-1]) / grid.dx


def learned_operator_jacobian(p: np.ndarray, tau: float, sigma: float, kind: str,
                              grid: W.Grid, u: np.ndarray, *, eps: float,
                              h: float = 1.0e-30) -> np.ndarray:
    """Dense interior Jacobian of the learned surrogate's operator by complex
    step."""
    g = grid.nghosts
    n = grid.n - 2 * g
    J = np.empty((n, n))
    base = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    for j in range(n):
        z = base.astype(np.complex128)
        z[g + j] += 1j * h
        z = W.apply_boundary_periodic(grid, z)
        J[:, j] = learned_operator_cs(p, tau, sigma, kind, grid, z, eps=eps)[g:g + n].imag / h
    return J


def learned_step_jacobian(p: np.ndarray, tau: float, sigma: float, kind: str,
                          grid: W.Grid, u: np.ndarray, dt: float, *, eps: float,
                          h: float = 1.0e-30) -> np.ndarray:
    """Dense interior Jacobian of one SSP-RK3 step of the learned surrogate
    (Burgers, ``kind`` rusanov or godunov) by complex step."""

    def operator(t, x):
        return learned_operator_cs(p, tau, sigma, kind, grid, x, eps=eps)

    g = grid.nghosts
    n = grid.n - 2 * g
    J = np.empty((n, n))
    base = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    for j in range(n):
        z = base.astype(np.complex128)
        z[g + j] += 1j * h
        z = W.apply_boundary_periodic(grid, z)
        out = W.ssprk33_advance(operator, dt, 0.0, z)
        J[:, j] = out[g:g + n].imag / h
    return J


# }}}
