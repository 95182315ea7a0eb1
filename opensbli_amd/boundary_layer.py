"""Laminar compressible boundary-layer profile for flat-plate initial states.

The counterpart of the reference's Boundary_layer_profile / Initialise_Katzer
(opensbli/utilities/katzer_init.py:19-180, 240-265): the similarity form of
the compressible boundary-layer equations with a Sutherland viscosity law
(White, Viscous Fluid Flow, eqs. 7.32),

    (C f'')' + f f'' = 0
    (C g')' + Pr f g' + Pr (gamma - 1) M^2 C f''^2 = 0,   C = sqrt(g)(1 + s)/(g + s)

with f'(0) = 0, f(0) = 0, f'(inf) = 1, g(inf) = 1 and either g'(0) = 0
(adiabatic wall) or g(0) = Tw, solved by shooting on (f''(0), g(0) or g'(0))
with Newton iterations.  Here the ODE system is integrated with scipy's
DOP853; the profile is then mapped from the similarity variable to the
physical wall distance by the Howarth-Dorodnitsyn stretch dy = g d(eta) and
scaled so that the displacement thickness is 1, as the reference does
(katzer_init.py:147-167).  The wall-normal mass flux follows from continuity
for a profile that grows like sqrt(x) (katzer_init.py:435-467), here in
closed form rather than by a finite difference in x.

Host-side set-up code, run once before the time loop.
"""
import numpy as np
from scipy.integrate import solve_ivp
from scipy.interpolate import CubicSpline


class SimilarityProfile:
    """u(y), T(y), rho*v(y) of the laminar boundary layer at Re_delta*, with
    y in units of the displacement thickness and free-stream values 1."""

    def __init__(self, Minf, Re, Pr=0.72, gama=1.4, Tw=-1.0, Tinf=288.0, suth=110.4, eta_max=10.0, npts=1001):
        self.Minf, self.Re, self.Pr, self.gama = Minf, Re, Pr, gama
        self.s = suth / Tinf
        self.adiabatic = Tw < 0
        self.Tw_given = Tw
        self.eta = np.linspace(0.0, eta_max, npts)
        self._solve()
        self._physical()

    # -- similarity equations --------------------------------------------
    def _rhs(self, _eta, v):
        s, pr, g1m2 = self.s, self.Pr, (self.gama - 1.0) * self.Minf ** 2
        f, fp, fpp, g, gp = v
        sg = np.sqrt(g)
        c = sg * (1.0 + s) / (g + s)
        dcdg = (0.5 / sg - sg / (g + s)) * (1.0 + s) / (g + s)
        cp = dcdg * gp
        return [fp, fpp, -fpp * (cp + f) / c, gp, -gp * (cp + pr * f) / c - pr * g1m2 * fpp * fpp]

    def _shoot(self, a, b):
        """Integrate from the wall with f''(0) = a and g(0) = b (adiabatic) or
        g'(0) = b (fixed wall temperature); returns the end state."""
        v0 = [0.0, 0.0, a, b, 0.0] if self.adiabatic else [0.0, 0.0, a, self.Tw_given, b]
        sol = solve_ivp(self._rhs, (self.eta[0], self.eta[-1]), v0, method="DOP853", t_eval=self.eta,
                        rtol=1e-11, atol=1e-12)
        if not sol.success:
            raise RuntimeError("boundary-layer integration failed: " + sol.message)
        return sol.y

    def _solve(self):
        g1m2 = (self.gama - 1.0) * self.Minf ** 2
        if self.adiabatic:  # recovery-temperature guess
            b = 1.0 + 0.5 * 0.84 * g1m2
            a = 0.47 * b ** 0.21
        else:
            a, b = 0.47, 0.0
        for _ in range(30):
            y = self._shoot(a, b)
            res = np.array([y[1, -1] - 1.0, y[3, -1] - 1.0])
            if np.abs(res).sum() < 1e-11:
                break
            da, db = 1e-6 * max(abs(a), 1e-3), 1e-6 * max(abs(b), 1e-3)
            ya, yb = self._shoot(a + da, b), self._shoot(a, b + db)
            J = np.array([[(ya[1, -1] - y[1, -1]) / da, (yb[1, -1] - y[1, -1]) / db],
                          [(ya[3, -1] - y[3, -1]) / da, (yb[3, -1] - y[3, -1]) / db]])
            step = np.linalg.solve(J, -res)
            a, b = a + step[0], b + step[1]
        else:
            raise RuntimeError("boundary-layer shooting did not converge")
        self.soln = self._shoot(a, b)
        self.fpp0 = a
        self.Twall = self.soln[3, 0]

    def _physical(self):
        f, fp, g = self.soln[0], self.soln[1], self.soln[3]
        eta = self.eta
        de = np.diff(eta)
        # Howarth-Dorodnitsyn: y ~ int g d(eta); displacement thickness ~ int (g - f') d(eta)
        z = np.concatenate([[0.0], np.cumsum(0.5 * de * (g[1:] + g[:-1]))])
        scale = np.sum(0.5 * de * ((g - fp)[1:] + (g - fp)[:-1]))
        self.scale = scale
        self.y = z / scale
        self.u = fp
        self.T = g
        self._u = CubicSpline(self.y, self.u)
        self._T = CubicSpline(self.y, self.T)
        ru = self.u / self.T
        self._ru = CubicSpline(self.y, ru)
        self._int_ru = self._ru.antiderivative()
        # station of the profile: delta*(x) = sqrt(x / x0) with delta*(x0) = 1 at Re_delta* = Re
        self.x0 = 0.5 * self.Re / scale ** 2

    def evaluate(self, yq):
        """rho, u, rho*v, T at wall distances yq (free stream above the profile)."""
        yq = np.asarray(yq, dtype=np.float64)
        top = self.y[-1]
        inside = yq <= top
        yc = np.clip(yq, 0.0, top)
        u = np.where(inside, self._u(yc), 1.0)
        T = np.where(inside, self._T(yc), 1.0)
        rho = 1.0 / T
        # continuity for rho*u = F(y / sqrt(x/x0)):  d(rho v)/dy = F'(y) y / (2 x0)
        #   rho*v(y) = (y F(y) - int_0^y F) / (2 x0)
        F = np.where(inside, self._ru(yc), 1.0)
        intF = np.where(inside, self._int_ru(yc), self._int_ru(top) + (yq - top))
        rv = (yq * F - intF) / (2.0 * self.x0)
        return rho, u, rv, T
