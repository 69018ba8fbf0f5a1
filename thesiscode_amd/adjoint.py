"""Gradient of a final-time functional with respect to the initial panel
vorticity (and positions), through libbve's adjoint.

The functional of the benchmark is J = sum_i A_i zeta_i(T) g(x_i(T)) over
the active panels with the Gaussian weight g(x) = exp(-(1 - x.x_g) / s^2)
("width" s, centre x_g); its cotangent is

    dJ/dzeta_i = A_i g(x_i),    dJ/dx_i = A_i zeta_i g(x_i) x_g / s^2.
"""
import math

import numpy as np

from .initial_conditions import latlon_to_xyz


class GaussianWeightFunctional:
    def __init__(self, lat=np.deg2rad(30.0), lon=np.deg2rad(90.0), width=0.2):
        self.centre = latlon_to_xyz(lat, lon)
        self.width = width

    def weight(self, x):
        return np.exp(-(1.0 - x @ self.centre) / (self.width * self.width))

    def value(self, x, zeta, area):
        return float(np.sum(area * zeta * self.weight(x)))

    def difference(self, x, zeta, dx, dzeta, area):
        """J(x + dx, zeta + dzeta) - J(x, zeta) from the differences: g(x + dx)
        = g(x) exp(dx.c / w^2), so g(x + dx) - g(x) = g(x) expm1(dx.c / w^2),
        and no term is the difference of two nearly equal numbers."""
        g = self.weight(x)
        dg = g * np.expm1((dx @ self.centre) / (self.width * self.width))
        return math.fsum((area * (dzeta * (g + dg) + zeta * dg)).tolist())

    def cotangent(self, x, zeta, area):
        g = self.weight(x)
        lam_zeta = area * g
        lam_x = (area * zeta * g / (self.width * self.width))[:, None] * self.centre[None, :]
        return lam_x, lam_zeta


def central_difference(ctx, functional, area, dt, n_steps, j, eps):
    """(J(zeta0 + eps e_j) - J(zeta0 - eps e_j)) / (2 eps) for the resident
    active state of ctx, each J difference to the unperturbed trajectory
    propagated in difference form (bve_difference_forward): the quotient
    carries rounding relative to itself, so a small eps is not swamped by
    cancellation."""
    e = np.zeros(ctx.n_active)
    out = []
    for sgn in (1.0, -1.0):
        e[j] = sgn * eps
        xT, zT, dxT, dzT = ctx.difference_forward(dt, n_steps, e)
        out.append(functional.difference(xT, zT, dxT, dzT, area))
    return (out[0] - out[1]) / (2.0 * eps)


def gradient(ctx, functional, area, dt, n_steps, checkpoint_every):
    """dJ/dzeta_0 and dJ/dx_0 of the resident active state (libbve context
    with panels loaded); returns (J, grad_zeta0, grad_x0)."""
    xT, zT = ctx.adjoint_forward(dt, n_steps, checkpoint_every)
    lam_x, lam_zeta = functional.cotangent(xT, zT, area)
    gx, gz = ctx.adjoint_backward(lam_x, lam_zeta)
    return functional.value(xT, zT, area), gz, gx
