"""The unrolled TGV iteration restated in torch binary64 (CPU), as the reference the GPU's
forward and hand-written adjoint are checked against: the same iteration from the same
start state and problem, differentiated by autograd.

The operators follow the library's (and the reference's) differences:
  F f(k) = f(k+1) - f(k), 0 at the last index;
  B f(k) = f(0) at k = 0, -f(n-2) at k = n-1, f(k) - f(k-1) between, 0 where n = 1;
  div p = B_x p_x + B_y p_y;  div_2 r = (F_x r_xx + F_y r_xy, F_x r_xy + F_y r_yy);
  grad x = (F_x x, F_y x);  E w = (B_x w_x, B_y w_y, (B_y w_x + B_x w_y) / 2);
  the 8x8 DCT with the orthonormal cosine basis, coefficient 8u + v for row frequency u.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def basis() -> torch.Tensor:
    c = torch.zeros(8, 8, dtype=torch.float64)
    for k in range(8):
        scale = math.sqrt(1 / 8) if k == 0 else 0.5
        for n in range(8):
            c[k, n] = scale * math.cos(math.pi * (2 * n + 1) * k / 16)
    return c


def _forward_x(f):
    return torch.cat([f[..., 1:] - f[..., :-1], torch.zeros_like(f[..., :1])], dim=-1)


def _forward_y(f):
    return torch.cat([f[..., 1:, :] - f[..., :-1, :], torch.zeros_like(f[..., :1, :])], dim=-2)


def _backward_x(f):
    if f.shape[-1] == 1:
        return torch.zeros_like(f)
    return torch.cat([f[..., :1], f[..., 1:-1] - f[..., :-2], -f[..., -2:-1]], dim=-1)


def _backward_y(f):
    if f.shape[-2] == 1:
        return torch.zeros_like(f)
    return torch.cat([f[..., :1, :], f[..., 1:-1, :] - f[..., :-2, :], -f[..., -2:-1, :]], dim=-2)


def _blocks(plane):
    b, h, w = plane.shape
    return plane.reshape(b, h // 8, 8, w // 8, 8).permute(0, 1, 3, 2, 4)


def _plane(blocks):
    b, rows, columns, _, _ = blocks.shape
    return blocks.permute(0, 1, 3, 2, 4).reshape(b, rows * 8, columns * 8)


def _dct(plane, c):
    blocks = _blocks(plane)
    return torch.matmul(torch.matmul(c, blocks), c.t().contiguous())


def _idct(coefficients, c):
    return _plane(torch.matmul(torch.matmul(c.t().contiguous(), coefficients), c))


def _ball(values, square, radius):
    norm = torch.sqrt(torch.clamp(square, min=1e-300))
    scale = torch.where(norm > radius, radius / norm, torch.ones_like(norm))
    return values * scale


def unrolled(start: np.ndarray, problems: list[np.ndarray], constants: dict, alpha0, alpha1, weight_map,
             iterations: int) -> torch.Tensor:
    """The last x of every picture, (count, H, W) binary64, differentiable in alpha0, alpha1
    and weight_map (binary64 tensors)."""
    state = torch.as_tensor(start, dtype=torch.float64)
    x, w, p, r = state[:, 0], state[:, 1:3], state[:, 3:5], state[:, 5:8]
    problem = torch.as_tensor(np.stack(problems), dtype=torch.float64)  # (count, rows, columns, 64, 4)
    count, rows, columns = problem.shape[:3]
    field = problem.reshape(count, rows, columns, 8, 8, 4)
    lower, upper, centre, weight = (field[..., k] for k in range(4))
    tau, sigma, gamma, rho = constants["tau"], constants["sigma"], constants["gamma"], constants["rho"]
    descent, ascent = tau * gamma, sigma * gamma
    t = tau * weight  # one sample a cell
    r_scale = 1.0 / (1.0 + t)
    c = basis()
    radius = alpha1 * weight_map
    for _ in range(iterations):
        v = x + descent * (_backward_y(p[:, 1]) + _backward_x(p[:, 0]))
        z = (_dct(v, c) + t * centre) * r_scale
        zeta = torch.minimum(torch.maximum(z, lower), upper)
        tilde = _idct(zeta, c)
        bar_x = 2 * tilde - x
        x_next = rho * tilde + (1 - rho) * x
        dx = _forward_x(r[:, 0]) + _forward_y(r[:, 2])
        dy = _forward_x(r[:, 2]) + _forward_y(r[:, 1])
        w_tilde = w + descent * (p + torch.stack([dx, dy], dim=1))
        w_bar = 2 * w_tilde - w
        p_hat = p + ascent * (torch.stack([_forward_x(bar_x), _forward_y(bar_x)], dim=1) - w_bar)
        p_tilde = _ball(p_hat, (p_hat**2).sum(dim=1, keepdim=True), radius)
        e = torch.stack([_backward_x(w_bar[:, 0]), _backward_y(w_bar[:, 1]),
                         0.5 * (_backward_y(w_bar[:, 0]) + _backward_x(w_bar[:, 1]))], dim=1)
        r_hat = r + ascent * e
        square = r_hat[:, 0] ** 2 + r_hat[:, 1] ** 2 + 2 * r_hat[:, 2] ** 2
        r_tilde = _ball(r_hat, square.unsqueeze(1), alpha0)
        x = x_next
        w = rho * w_tilde + (1 - rho) * w
        p = rho * p_tilde + (1 - rho) * p
        r = rho * r_tilde + (1 - rho) * r
    return x
