"""FP64 numpy reference of the discrete adjoint.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
adjoint.  The product path never imports this module.

The reference has no adjoint.  This module is the reverse of the oracle's
own stage map (oracle.c pre_step_derivatives / solve_eqn_set, which restate
CSolver_AWS.cpp:829-1504 on periodic lines without sponges) and of its RK4
and TVDRK3 steps with the filter (RK4.hpp:35-118, TVDRK3.hpp:38-96,
CSolver_AWS.cpp:1682-1893), written independently of the GPU kernels:

- every compact operator is the dense matrix of the oracle's own line
  operator (or_line_op, column by column), and its adjoint is that matrix
  transposed; nothing assumes the periodic D1 skew-adjoint or D2 and the
  filter self-adjoint, so the GPU's use of those identities is checked,
  not shared;
- the pointwise right-hand side is held as lists of monomials in named
  intermediates (each reference statement expanded term by term), and the
  reverse sweep differentiates every monomial factor by factor;
- nothing is taped: the reverse pass recomputes the forward intermediates.

Pinned by tests/test_oracle_adjoint.py: the numpy forward stage and the RK4
and TVDRK3 steps against the C oracle (1e-12), the dot-product identity
against central differences of the forward stage, and the kinetic-energy
gradient through either scheme against central differences of whole
C-oracle runs with the dt held fixed.
"""
import numpy as np

import pyoracle as po

C43, C23, C13 = 4.0 / 3.0, 2.0 / 3.0, 1.0 / 3.0


def _monomials_rhs():
    """dt*R_c = dt * sum(coef * prod(factors)) per conserved variable c.

    Factor names: prims U V W, mu, MuX/MuY/MuZ (= amu * T_j), first
    derivatives Ux..Tz, second Uxx..Tzz, mixed Uxy Uxz Vxy Vyz Wxz Wyz, and
    the flux / continuity derivatives (contX, mxX, ..., eZ)."""
    eq = [[] for _ in range(5)]
    # continuity (CSolver_AWS.cpp:1269)
    for d in "XYZ":
        eq[0].append((-1.0, ("cont" + d,)))
    # x momentum (:1311-1316)
    eq[1] += [(C43, ("mu", "Uxx")), (1.0, ("mu", "Uyy")), (1.0, ("mu", "Uzz")),
              (C13, ("mu", "Vxy")), (C13, ("mu", "Wxz")),
              (C43, ("MuX", "Ux")), (-C23, ("MuX", "Vy")), (-C23, ("MuX", "Wz")),
              (1.0, ("MuY", "Uy")), (1.0, ("MuY", "Vx")),
              (1.0, ("MuZ", "Wx")), (1.0, ("MuZ", "Uz"))]
    # y momentum (:1360-1365)
    eq[2] += [(C43, ("mu", "Vyy")), (1.0, ("mu", "Vxx")), (1.0, ("mu", "Vzz")),
              (C13, ("mu", "Uxy")), (C13, ("mu", "Wyz")),
              (C43, ("MuY", "Vy")), (-C23, ("MuY", "Ux")), (-C23, ("MuY", "Wz")),
              (1.0, ("MuX", "Uy")), (1.0, ("MuX", "Vx")),
              (1.0, ("MuZ", "Wy")), (1.0, ("MuZ", "Vz"))]
    # z momentum (:1408-1414)
    eq[3] += [(C43, ("mu", "Wzz")), (1.0, ("mu", "Wyy")), (1.0, ("mu", "Wxx")),
              (C13, ("mu", "Uxz")), (C13, ("mu", "Vyz")),
              (C43, ("MuZ", "Wz")), (-C23, ("MuZ", "Ux")), (-C23, ("MuZ", "Vy")),
              (1.0, ("MuX", "Wx")), (1.0, ("MuX", "Uz")),
              (1.0, ("MuY", "Wy")), (1.0, ("MuY", "Vz"))]
    for c, f in ((1, "mx"), (2, "my"), (3, "mz")):
        for d in "XYZ":
            eq[c].append((-1.0, (f + d,)))
    # energy (:1458-1495): qtemp
    e = eq[4]
    for j in "xyz":
        e.append(("cpPr", ("Mu" + j.upper(), "T" + j)))
        e.append(("cpPr", ("mu", "T" + j + j)))
    # vt1 = mu * (...)
    lap = {"U": ((C43, "Uxx"), (1.0, "Uyy"), (1.0, "Uzz")),
           "V": ((1.0, "Vxx"), (C43, "Vyy"), (1.0, "Vzz")),
           "W": ((1.0, "Wxx"), (1.0, "Wyy"), (C43, "Wzz"))}
    for u, terms in lap.items():
        for cf, dd in terms:
            e.append((cf, ("mu", u, dd)))
    for a in ("Ux", "Vy", "Wz"):
        e.append((C43, ("mu", a, a)))
    for a in ("Uy", "Uz", "Vx", "Vz", "Wx", "Wy"):
        e.append((1.0, ("mu", a, a)))
    for a, b in (("Ux", "Vy"), ("Ux", "Wz"), ("Vy", "Wz")):
        e.append((-C43, ("mu", a, b)))
    for a, b in (("Uy", "Vx"), ("Uz", "Wx"), ("Vz", "Wy")):
        e.append((2.0, ("mu", a, b)))
    for u, m in (("U", "Vxy"), ("U", "Wxz"), ("V", "Uxy"), ("V", "Wyz"), ("W", "Uxz"),
                 ("W", "Vyz")):
        e.append((C13, ("mu", u, m)))
    # vt2
    e += [(C43, ("U", "MuX", "Ux")), (C43, ("V", "MuY", "Vy")), (C43, ("W", "MuZ", "Wz")),
          (-C23, ("U", "MuX", "Vy")), (-C23, ("U", "MuX", "Wz")),
          (-C23, ("V", "MuY", "Ux")), (-C23, ("V", "MuY", "Wz")),
          (-C23, ("W", "MuZ", "Ux")), (-C23, ("W", "MuZ", "Vy"))]
    for u, M, a, b in (("U", "MuY", "Uy", "Vx"), ("U", "MuZ", "Uz", "Wx"),
                       ("V", "MuX", "Uy", "Vx"), ("V", "MuZ", "Vz", "Wy"),
                       ("W", "MuX", "Uz", "Wx"), ("W", "MuY", "Vz", "Wy")):
        e.append((1.0, (u, M, a)))
        e.append((1.0, (u, M, b)))
    # Euler fluxes
    for d in "XYZ":
        e.append((-1.0, ("e" + d,)))
    return eq


RHS_TERMS = _monomials_rhs()
FIRST = [f + j for f in "UVWT" for j in "xyz"]
SECOND = [f + j + j for f in "UVWT" for j in "xyz"]
# mixed derivatives the right-hand side reads: name -> (inner first derivative, outer direction)
MIXED = {"Uxy": ("Ux", 1), "Vxy": ("Vx", 1), "Uxz": ("Ux", 2), "Vyz": ("Vy", 2),
         "Wxz": ("Wx", 2), "Wyz": ("Wy", 2)}
DIR = {"x": 0, "y": 1, "z": 2, "X": 0, "Y": 1, "Z": 2}


class Adjoint:
    """Dense-operator forward and reverse of the oracle's stage map and of
    its RK4 or TVDRK3 step (cfg.rk) on a periodic (Nz, Ny, Nx) grid.
    `cfg` is a pyoracle config."""

    def __init__(self, cfg):
        self.dims = (cfg.Nx, cfg.Ny, cfg.Nz)
        self.shape = (cfg.Nz, cfg.Ny, cfg.Nx)
        for d in range(3):
            if cfg.bc_type[d] != po.PERIODIC_SOLVE:
                raise ValueError("the numpy adjoint covers periodic lines only")
        lengths = (cfg.Lx, cfg.Ly, cfg.Lz)
        self.rk = cfg.rk
        self.gamma = cfg.gamma
        self.R = 1.0 / cfg.gamma
        self.cpPr = self.R * cfg.gamma / (cfg.gamma - 1.0) / cfg.Pr
        self.mu_ref = cfg.mu_ref
        self.M = {}
        for d in range(3):
            n = self.dims[d]
            dh = lengths[d] / (n - 1.0)
            for kind in (0, 1, 2):  # filter, D1, D2
                cols = [po.line_op(kind, po.PERIODIC_SOLVE, e, dh, cfg.alpha_f)
                        for e in np.eye(n)]
                self.M[kind, d] = np.array(cols).T

    # ---------------------------------------------------------------- operators
    def op(self, kind, d, f, transpose=False):
        M = self.M[kind, d].T if transpose else self.M[kind, d]
        axis = 2 - d
        return np.moveaxis(np.tensordot(M, f, axes=([1], [axis])), 0, axis)

    def filt(self, q, order, transpose=False):
        seq = order[::-1] if transpose else order
        out = []
        for f in q:
            for d in seq:
                f = self.op(0, d, f, transpose)
            out.append(f)
        return out

    # ---------------------------------------------------------------- forward
    def intermediates(self, q):
        g = self.gamma
        rho, rU, rV, rW, rE = q
        v = {"rho": rho, "rU": rU, "rV": rV, "rW": rW, "rE": rE}
        v["U"], v["V"], v["W"] = rU / rho, rV / rho, rW / rho
        v["p"] = (g - 1) * (rE - 0.5 * rho * (v["U"] ** 2 + v["V"] ** 2 + v["W"] ** 2))
        v["T"] = v["p"] / (rho * self.R)
        v["mu"] = self.mu_ref * v["T"] ** 0.76
        v["amu"] = 0.76 * self.mu_ref * v["T"] ** (0.76 - 1.0)
        for f in FIRST:
            v[f] = self.op(1, DIR[f[1]], v[f[0]])
        for f in SECOND:
            v[f] = self.op(2, DIR[f[1]], v[f[0]])
        for m, (inner, d) in MIXED.items():
            v[m] = self.op(1, d, v[inner])
        for j in "xyz":
            v["Mu" + j.upper()] = v["amu"] * v["T" + j]
        for d, u in (("X", "U"), ("Y", "V"), ("Z", "W")):
            vel = v[u]
            v["cont" + d] = self.op(1, DIR[d], {"X": rU, "Y": rV, "Z": rW}[d])
            fl = {"mx": rU * vel, "my": rV * vel, "mz": rW * vel, "e": rE * vel + vel * v["p"]}
            fl[{"X": "mx", "Y": "my", "Z": "mz"}[d]] += v["p"]
            for k, t in fl.items():
                v[k + d] = self.op(1, DIR[d], t)
        return v

    def _coef(self, c):
        return self.cpPr if c == "cpPr" else c

    def rhs(self, q, dt):
        """dt * RHS(q), the oracle's k2 for stage state q"""
        v = self.intermediates(q)
        out = []
        for terms in RHS_TERMS:
            acc = np.zeros(self.shape)
            for c, fac in terms:
                t = self._coef(c)
                for f in fac:
                    t = t * v[f]
                acc = acc + t
            out.append(dt * acc)
        return out

    # ---------------------------------------------------------------- reverse
    def rhs_vjp(self, q, lam, dt):
        """(d(dt R)/dq)^T lam"""
        g = self.gamma
        v = self.intermediates(q)
        b = {}

        def add(name, val):
            b[name] = b.get(name, 0.0) + val

        for c, terms in enumerate(RHS_TERMS):
            w = dt * lam[c]
            for coef, fac in terms:
                cf = self._coef(coef)
                for k, f in enumerate(fac):
                    t = cf * w
                    for kk, ff in enumerate(fac):
                        if kk != k:
                            t = t * v[ff]
                    add(f, t)
        zero = np.zeros(self.shape)
        for d, u in (("X", "U"), ("Y", "V"), ("Z", "W")):
            dd = DIR[d]
            add({"X": "rU", "Y": "rV", "Z": "rW"}[d], self.op(1, dd, b.get("cont" + d, zero), True))
            vel = v[u]
            for k, m in (("mx", "rU"), ("my", "rV"), ("mz", "rW")):
                tb = self.op(1, dd, b.get(k + d, zero), True)
                add(m, tb * vel)
                add(u, tb * v[m])
                if k == {"X": "mx", "Y": "my", "Z": "mz"}[d]:
                    add("p", tb)
            tb = self.op(1, dd, b.get("e" + d, zero), True)
            add("rE", tb * vel)
            add(u, tb * (v["rE"] + v["p"]))
            add("p", tb * vel)
        for j in "xyz":
            mb = b.get("Mu" + j.upper(), zero)
            add("amu", mb * v["T" + j])
            add("T" + j, mb * v["amu"])
        for m, (inner, d) in MIXED.items():
            add(inner, self.op(1, d, b.get(m, zero), True))
        for f in SECOND:
            add(f[0], self.op(2, DIR[f[1]], b.get(f, zero), True))
        for f in FIRST:
            add(f[0], self.op(1, DIR[f[1]], b.get(f, zero), True))
        T = v["T"]
        add("T", b.get("mu", zero) * 0.76 * v["mu"] / T)
        add("T", b.get("amu", zero) * (0.76 - 1.0) * v["amu"] / T)
        rho = v["rho"]
        Tb = b.get("T", zero)
        add("p", Tb / (rho * self.R))
        add("rho", -Tb * T / rho)
        pb = b.get("p", zero)
        add("rE", (g - 1) * pb)
        add("rho", -(g - 1) * 0.5 * (v["U"] ** 2 + v["V"] ** 2 + v["W"] ** 2) * pb)
        for u, m in (("U", "rU"), ("V", "rV"), ("W", "rW")):
            ub = b.get(u, zero) - (g - 1) * rho * v[u] * pb
            add(m, ub / rho)
            add("rho", -ub * v[u] / rho)
        return [b.get(k, zero) for k in ("rho", "rU", "rV", "rW", "rE")]

    # ---------------------------------------------------------------- RK4
    @staticmethod
    def filter_order(count):
        m = count % 3
        return (0, 1, 2) if m == 1 else (1, 2, 0) if m == 2 else (2, 0, 1)

    def rk4_step(self, q, dt, order):
        """one RK4 step (RK4.hpp:35-118) followed by the filter in `order`
        (None: no filter)"""
        k1 = self.rhs(q, dt)
        k2 = self.rhs([a + 0.5 * k for a, k in zip(q, k1)], dt)
        k3 = self.rhs([a + 0.5 * k for a, k in zip(q, k2)], dt)
        k4 = self.rhs([a + k for a, k in zip(q, k3)], dt)
        acc = [a + k1_ / 6 + k2_ / 3 + k3_ / 3 + k4_ / 6
               for a, k1_, k2_, k3_, k4_ in zip(q, k1, k2, k3, k4)]
        return self.filt(acc, order) if order is not None else acc

    def tvdrk3_step(self, q, dt, order):
        """one TVDRK3 step (TVDRK3.hpp:38-96: Q2 = q + k1, Q3 = 3/4 q +
        1/4 (Q2 + k2), q' = 1/3 q + 2/3 (Q3 + k3)) followed by the filter in
        `order` (None: no filter)"""
        k1 = self.rhs(q, dt)
        Q2 = [a + k for a, k in zip(q, k1)]
        k2 = self.rhs(Q2, dt)
        Q3 = [0.75 * a + 0.25 * b + 0.25 * k for a, b, k in zip(q, Q2, k2)]
        k3 = self.rhs(Q3, dt)
        acc = [a / 3 + 2.0 / 3.0 * b + 2.0 / 3.0 * k for a, b, k in zip(q, Q3, k3)]
        return self.filt(acc, order) if order is not None else acc

    def step(self, q, dt, order):
        return self.rk4_step(q, dt, order) if self.rk == 4 else self.tvdrk3_step(q, dt, order)

    def step_vjp(self, q, dt, order, lam):
        if self.rk == 4:
            return self.rk4_step_vjp(q, dt, order, lam)
        return self.tvdrk3_step_vjp(q, dt, order, lam)

    def conserved_from_ic(self, rho, U, V, W, p):
        g = self.gamma
        return [rho, rho * U, rho * V, rho * W, p / (g - 1) + 0.5 * rho * (U * U + V * V + W * W)]

    def ke_gradient(self, ic, nsteps, dt, filter_step=1):
        """J = <rho |u|^2 / 2> after nsteps steps of the configured scheme
        (RK4 or TVDRK3) with fixed dt and the filter every filter_step
        steps; returns J, dJ/dU0, dJ/dV0, dJ/dW0 at fixed initial density
        and pressure"""
        rho0, U0, V0, W0, p0 = [np.asarray(a, dtype=np.float64) for a in ic]
        q = self.conserved_from_ic(rho0, U0, V0, W0, p0)
        states, orders = [], []
        fcount = 0
        for n in range(1, nsteps + 1):
            states.append(q)
            order = None
            if n % filter_step == 0:
                fcount += 1
                order = self.filter_order(fcount)
            orders.append(order)
            q = self.step(q, dt, order)
        N = q[0].size
        rho = q[0]
        u = [q[i] / rho for i in (1, 2, 3)]
        J = float(np.sum(0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / N)
        lam = [-0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / N, u[0] / N, u[1] / N, u[2] / N,
               np.zeros(self.shape)]
        for n in range(nsteps - 1, -1, -1):
            lam = self.step_vjp(states[n], dt, orders[n], lam)
        # m_i = rho U_i, E = p/(g-1) + rho |U|^2 / 2 at fixed rho and p
        grads = [rho0 * (lam[i + 1] + (U0, V0, W0)[i] * lam[4]) for i in range(3)]
        return J, grads[0], grads[1], grads[2]

    def rk4_step_vjp(self, q, dt, order, lam):
        """cotangent of the step's start state from that of its end state;
        with Q2 = q + k1/2, Q3 = q + k2/2, Q4 = q + k3 and the filtered
        acc = q + k1/6 + k2/3 + k3/3 + k4/6"""
        k1 = self.rhs(q, dt)
        Q2 = [a + 0.5 * k for a, k in zip(q, k1)]
        k2 = self.rhs(Q2, dt)
        Q3 = [a + 0.5 * k for a, k in zip(q, k2)]
        k3 = self.rhs(Q3, dt)
        Q4 = [a + k for a, k in zip(q, k3)]
        la = self.filt(lam, order, transpose=True) if order is not None else list(lam)
        qb = list(la)
        Q4b = self.rhs_vjp(Q4, [x / 6 for x in la], dt)
        qb = [a + c for a, c in zip(qb, Q4b)]
        Q3b = self.rhs_vjp(Q3, [x / 3 + c for x, c in zip(la, Q4b)], dt)
        qb = [a + c for a, c in zip(qb, Q3b)]
        Q2b = self.rhs_vjp(Q2, [x / 3 + 0.5 * c for x, c in zip(la, Q3b)], dt)
        qb = [a + c for a, c in zip(qb, Q2b)]
        Q1b = self.rhs_vjp(q, [x / 6 + 0.5 * c for x, c in zip(la, Q2b)], dt)
        return [a + c for a, c in zip(qb, Q1b)]

    def tvdrk3_step_vjp(self, q, dt, order, lam):
        """cotangent of the step's start state from that of its end state
        for tvdrk3_step: with la the cotangent of the unfiltered result,
        Q3b = 2/3 la + J3^T (2/3 la), Q2b = 1/4 Q3b + J2^T (1/4 Q3b),
        qb = 1/3 la + 3/4 Q3b + Q2b + J1^T Q2b (J_s^T: rhs_vjp at Q_s)"""
        k1 = self.rhs(q, dt)
        Q2 = [a + k for a, k in zip(q, k1)]
        k2 = self.rhs(Q2, dt)
        Q3 = [0.75 * a + 0.25 * b + 0.25 * k for a, b, k in zip(q, Q2, k2)]
        la = self.filt(lam, order, transpose=True) if order is not None else list(lam)
        t3 = [2.0 / 3.0 * x for x in la]
        Q3b = [a + c for a, c in zip(t3, self.rhs_vjp(Q3, t3, dt))]
        t2 = [0.25 * x for x in Q3b]
        Q2b = [a + c for a, c in zip(t2, self.rhs_vjp(Q2, t2, dt))]
        J1 = self.rhs_vjp(q, Q2b, dt)
        return [x / 3 + 0.75 * a + b + c for x, a, b, c in zip(la, Q3b, Q2b, J1)]
