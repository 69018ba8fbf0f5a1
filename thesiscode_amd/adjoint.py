"""Gradient of a final-time functional with respect to the initial panel
vorticity (and positions), through libbve's adjoint.

The functional of the benchmark is J = sum_i A_i zeta_i(T) g(x_i(T)) over
the active panels with the Gaussian weight g(x) = exp(-(1 - x.x_g) / s^2)
("width" s, centre x_g); its cotangent is

    dJ/dzeta_i = A_i g(x_i),    dJ/dx_i = A_i zeta_i g(x_i) x_g / s^2.
"""
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

    def cotangent(self, x, zeta, area):
        g = self.weight(x)
        lam_zeta = area * g
        lam_x = (area * zeta * g / (self.width * self.width))[:, None] * self.centre[None, :]
        return lam_x, lam_zeta


def gradient(ctx, functional, area, dt, n_steps, checkpoint_every):
    """dJ/dzeta_0 and dJ/dx_0 of the resident active state (libbve context
    with panels loaded); returns (J, grad_zeta0, grad_x0)."""
    xT, zT = ctx.adjoint_forward(dt, n_steps, checkpoint_every)
    lam_x, lam_zeta = functional.cotangent(xT, zT, area)
    gx, gz = ctx.adjoint_backward(lam_x, lam_zeta)
    return functional.value(xT, zT, area), gz, gx
