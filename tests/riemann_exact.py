"""The exact solution of the 1D Riemann problem for an ideal gas of index gamma (Toro, "Riemann Solvers and Numerical
Methods for Fluid Dynamics", 3rd ed., sections 4.2-4.5), written from the textbook as an independent grader of the
oracle's hydro (the reference grades its scheme the same way: crates/sr-physics/src/reference/riemann.rs:16-210).

The star pressure p* is the root of f_L(p) + f_R(p) + (u_R - u_L) = 0 (Newton-Raphson, then bisection as a
safeguard); u* = (u_L + u_R)/2 + (f_R - f_L)/2; sample(xi) evaluates the self-similar solution at xi = x/t.
"""
import math


class Solution:
    def __init__(self, left, right, gamma):
        self.g = gamma
        self.L, self.R = left, right  # (rho, u, p)
        self.p_star, self.u_star = self._solve()

    def _c(self, s):
        return math.sqrt(self.g * s[2] / s[0])

    def _f(self, p, s):
        rho, _, pk = s
        g = self.g
        if p > pk:  # shock: Rankine-Hugoniot
            a = 2.0 / ((g + 1.0) * rho)
            b = (g - 1.0) / (g + 1.0) * pk
            q = math.sqrt(a / (p + b))
            return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + b))
        c = self._c(s)  # rarefaction: isentrope and Riemann invariant
        e = (g - 1.0) / (2.0 * g)
        f = 2.0 * c / (g - 1.0) * ((p / pk) ** e - 1.0)
        df = 1.0 / (rho * c) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
        return f, df

    def _solve(self):
        L, R, g = self.L, self.R, self.g
        du = R[1] - L[1]
        if 2.0 * (self._c(L) + self._c(R)) / (g - 1.0) <= du:
            raise ValueError("vacuum opens between the two states")
        fun = lambda p: self._f(p, L)[0] + self._f(p, R)[0] + du
        p = max(1e-12, 0.5 * (L[2] + R[2]))
        for _ in range(100):
            fl, dl = self._f(p, L)
            fr, dr = self._f(p, R)
            step = (fl + fr + du) / (dl + dr)
            pn = max(1e-14, p - step)
            if abs(pn - p) <= 1e-15 * (pn + p):
                p = pn
                break
            p = pn
        lo, hi = 1e-14, max(L[2], R[2]) * 10.0 + 1.0
        if abs(fun(p)) > 1e-12 * (L[2] + R[2]):  # Newton stalled: bisection in a bracket
            while fun(hi) < 0.0:
                hi *= 2.0
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if fun(mid) > 0.0:
                    hi = mid
                else:
                    lo = mid
            p = 0.5 * (lo + hi)
        u = 0.5 * (L[1] + R[1]) + 0.5 * (self._f(p, R)[0] - self._f(p, L)[0])
        return p, u

    def star_densities(self):
        g, p = self.g, self.p_star
        out = []
        for s in (self.L, self.R):
            r = p / s[2]
            if p > s[2]:
                gm = (g - 1.0) / (g + 1.0)
                out.append(s[0] * (r + gm) / (gm * r + 1.0))
            else:
                out.append(s[0] * r ** (1.0 / g))
        return tuple(out)

    def right_shock_speed(self):
        g, rho, u, p = self.g, *self.R
        return u + self._c(self.R) * math.sqrt((g + 1.0) / (2.0 * g) * self.p_star / p + (g - 1.0) / (2.0 * g))

    def sample(self, xi):
        """(rho, u, p) at xi = x / t."""
        g, ps, us = self.g, self.p_star, self.u_star
        rl, rr = self.star_densities()
        if xi <= us:
            rho, u, p = self.L
            c = self._c(self.L)
            if ps > p:  # left shock
                s = u - c * math.sqrt((g + 1.0) / (2.0 * g) * ps / p + (g - 1.0) / (2.0 * g))
                return (rho, u, p) if xi <= s else (rl, us, ps)
            cs = c * (ps / p) ** ((g - 1.0) / (2.0 * g))
            if xi <= u - c:
                return rho, u, p
            if xi >= us - cs:
                return rl, us, ps
            uu = 2.0 / (g + 1.0) * (c + 0.5 * (g - 1.0) * u + xi)
            cc = 2.0 / (g + 1.0) * (c + 0.5 * (g - 1.0) * (u - xi))
            return rho * (cc / c) ** (2.0 / (g - 1.0)), uu, p * (cc / c) ** (2.0 * g / (g - 1.0))
        rho, u, p = self.R
        c = self._c(self.R)
        if ps > p:  # right shock
            s = self.right_shock_speed()
            return (rr, us, ps) if xi <= s else (rho, u, p)
        cs = c * (ps / p) ** ((g - 1.0) / (2.0 * g))
        if xi >= u + c:
            return rho, u, p
        if xi <= us + cs:
            return rr, us, ps
        uu = 2.0 / (g + 1.0) * (-c + 0.5 * (g - 1.0) * u + xi)
        cc = 2.0 / (g + 1.0) * (c - 0.5 * (g - 1.0) * (u - xi))
        return rho * (cc / c) ** (2.0 / (g - 1.0)), uu, p * (cc / c) ** (2.0 * g / (g - 1.0))
