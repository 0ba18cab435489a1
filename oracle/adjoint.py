"""FP64 numpy reference of the discrete adjoint.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
adjoint.  The product path never imports this module.

The reference has no adjoint.  This module is the reverse of the oracle's
own stage map (oracle.c pre_step_derivatives / solve_eqn_set /
post_step_bc, which restate CSolver_AWS.cpp:829-1670), of its RK4 and
TVDRK3 steps with the filter (RK4.hpp:35-118, TVDRK3.hpp:38-96,
CSolver_AWS.cpp:1682-1893) and of the sponge update (updateSponge,
CSolver_AWS.cpp:1937-2031) and of the wall handling (preStepBCHandling,
:553-826; the wall rows of postStepBCHandling, :1506-1605), on periodic or
non-periodic lines, with sponge faces, walls or both, written
independently of the GPU kernels:

- every compact operator is the dense matrix of the oracle's own line
  operator (or_line_op, column by column), and its adjoint is that matrix
  transposed; nothing assumes the periodic D1 skew-adjoint or D2 and the
  filter self-adjoint, so the GPU's use of those identities is checked,
  not shared;
- the pointwise right-hand side is held as lists of monomials in named
  intermediates (each reference statement expanded term by term), and the
  reverse sweep differentiates every monomial factor by factor;
- nothing is taped: the reverse pass recomputes the forward intermediates;
- sponges: the source sigma (A - q) (twice in the energy equation, Q3),
  the zeroed right-hand side on the sponge face planes, the running
  average A with its energy blend and the reset of the face planes to A
  after each step, and the stale primitives: at the first stage after a
  sponge update the x derivatives, x fluxes and mu of the reference come
  from the primitive arrays of the filtered state before the reset (the
  "X set" below), so their cotangents flow to that state, not to the
  stage state;
- walls: before each stage the wall faces (adiabatic X0..Z1, then moving
  X0..Z1) overwrite the momenta with rho times the wall velocity and the
  energy with that of p = rho R T_w, T_w = calcNeumann of the six inward
  temperatures, in place (at the first stage the state the RK update
  builds on); the right-hand side on a wall plane keeps only the normal
  continuity flux term of the last wall face through the point;
- walls beside sponges: at the first stage after a sponge update the wall
  correction acts on the state and on the X set's state alike (the wall
  temperature from the X set's temperatures, which on the sponge planes
  are the stale ones; the face values, formed with the state's density,
  written to both).

Pinned by tests/test_oracle_adjoint.py: the numpy forward stage and the RK4
and TVDRK3 steps against the C oracle (1e-12; with sponges and non-periodic
lines over several steps, averages included), the dot-product identity
against central differences of the forward stage (in the stage state, the
stale state and the averages), and the kinetic-energy gradient through
either scheme against central differences of whole C-oracle runs with the
dt held fixed, periodic and with sponges.
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


# the X set: the primitives the reference's x block reads (its primitive
# arrays as they stand: those of the stage state, or at the first stage
# after a sponge update those of the filtered state before the reset)
XSET = {"U": "U_X", "V": "V_X", "W": "W_X", "p": "p_X", "T": "T_X"}


class Adjoint:
    """Dense-operator forward and reverse of the oracle's stage map and of
    its RK4 or TVDRK3 step (cfg.rk) on an (Nz, Ny, Nx) grid, periodic or
    non-periodic per direction, with or without sponge faces.  `cfg` is a
    pyoracle config."""

    def __init__(self, cfg):
        self.dims = (cfg.Nx, cfg.Ny, cfg.Nz)
        self.shape = (cfg.Nz, cfg.Ny, cfg.Nx)
        kinds = [cfg.bc_kind[f] for f in range(6)]
        is_wall = [k in (po.ADIABATIC_WALL, po.MOVING_ADIABATIC_WALL) for k in kinds]
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
                cols = [po.line_op(kind, cfg.bc_type[d], e, dh, cfg.alpha_f)
                        for e in np.eye(n)]
                self.M[kind, d] = np.array(cols).T
        # sponges (CSolver_AWS.cpp:1937-2031, SpongeBC.hpp): sigma from the
        # oracle, the face planes, the averaging constants
        self.sponge = any(k == po.SPONGE for k in kinds)
        self.sigma = None
        self.onplane = np.zeros(self.shape, dtype=bool)
        if self.sponge:
            o = po.Solver(cfg)
            self.sigma = o.sponge_sigma().copy()
            o.close()
            for f in range(6):
                if kinds[f] != po.SPONGE:
                    continue
                d, hi = f // 2, f % 2
                idx = [slice(None)] * 3
                idx[2 - d] = -1 if hi else 0
                self.onplane[tuple(idx)] = True
        self.avgT = cfg.sponge_avgT
        self.epsP = cfg.sponge_epsP
        self.spP = 1.0 / cfg.gamma
        self.zmask = np.where(self.onplane, 0.0, 1.0)
        # walls: the faces in the reference's order (adiabatic, then moving)
        # with their velocities, and for the right-hand side the normal of
        # the last wall face through each point (-1: none)
        self.walls = []
        for kind in (po.ADIABATIC_WALL, po.MOVING_ADIABATIC_WALL):
            for f in range(6):
                if kinds[f] != kind:
                    continue
                d = f // 2
                uw = [cfg.wall_vel[f][m] if (kind == po.MOVING_ADIABATIC_WALL and m != d) else 0.0
                      for m in range(3)]
                self.walls.append((d, f % 2, uw))
        self.rowdir = np.full(self.shape, -1, dtype=int)
        for f in range(6):
            if is_wall[f]:
                self.rowdir[self.plane(f // 2, f % 2)] = f // 2

    @staticmethod
    def plane(d, hi, m=0):
        """index of the plane m points inward from face (d, hi)"""
        idx = [slice(None)] * 3
        idx[2 - d] = -1 - m if hi else m
        return tuple(idx)

    # ---------------------------------------------------------------- walls
    CN = np.array([360.0, -450.0, 400.0, -225.0, 72.0, -10.0]) / 147.0  # calcNeumann

    def wall_pre(self, q, keep=None, qs=None, keeps=None):
        """preStepBCHandling on a state (a new list); keep (a list) receives
        the state before each face.  With qs (the X set's state at the first
        stage after a sponge update; equal to q off the sponge planes) the
        pair: the wall temperature from qs's temperatures, the face values
        (with q's density) written to both; returns (q, qs), keeps receiving
        qs before each face"""
        q = [x.copy() for x in q]
        if qs is not None:
            qs = [x.copy() for x in qs]
        g = self.gamma
        for d, hi, uw in self.walls:
            if keep is not None:
                keep.append([x.copy() for x in q])
            if keeps is not None:
                keeps.append([x.copy() for x in qs])
            T = self.prims(q if qs is None else qs)["T"]
            Tw = sum(c * T[self.plane(d, hi, m + 1)] for m, c in enumerate(self.CN))
            fc = self.plane(d, hi)
            rho = q[0][fc]
            for k in range(3):
                q[1 + k][fc] = rho * uw[k]
            q[4][fc] = rho * self.R * Tw / (g - 1) + 0.5 * rho * sum(u * u for u in uw)
            if qs is not None:
                for k in range(5):
                    qs[k][fc] = q[k][fc]
        return q if qs is None else (q, qs)

    def wall_pre_vjp(self, q, lam, qs=None, lqs=None):
        """the cotangent of the state before preStepBCHandling from that of
        the state after it; with qs, those of the pair (q, qs) from (lam,
        lqs)"""
        if not self.walls:
            return list(lam) if qs is None else (list(lam), list(lqs))
        keep, keeps = [], []
        self.wall_pre(q, keep, qs, keeps if qs is not None else None)
        lam = [x.copy() for x in lam]
        if qs is not None:
            lqs = [x.copy() for x in lqs]
        g = self.gamma
        for i in range(len(self.walls) - 1, -1, -1):
            d, hi, uw = self.walls[i]
            st = keep[i]
            sT = st if qs is None else keeps[i]   # the temperatures' state
            lT = lam if qs is None else lqs
            v = self.prims(sT)
            fc = self.plane(d, hi)
            rho = st[0][fc]
            Tw = sum(c * v["T"][self.plane(d, hi, m + 1)] for m, c in enumerate(self.CN))
            lf = [lam[k][fc].copy() for k in range(5)]
            if qs is not None:
                for k in range(5):
                    lf[k] = lf[k] + lqs[k][fc]
                    lqs[k][fc] = 0.0
            lE = lf[4]
            lam[0][fc] = lf[0] + sum(lf[1 + k] * uw[k] for k in range(3)) + lE * (
                self.R * Tw / (g - 1) + 0.5 * sum(u * u for u in uw))
            for k in range(1, 5):
                lam[k][fc] = 0.0
            gT = lE * rho * self.R / (g - 1)
            for m, c in enumerate(self.CN):
                ip = self.plane(d, hi, m + 1)
                r = sT[0][ip]
                u = [sT[k][ip] / r for k in (1, 2, 3)]
                Tb = gT * c
                lT[4][ip] += Tb * (g - 1) / (r * self.R)
                for k in range(3):
                    lT[1 + k][ip] -= Tb * (g - 1) * u[k] / (r * self.R)
                lT[0][ip] += Tb * ((g - 1) * 0.5 * sum(x * x for x in u) / (r * self.R)
                                   - v["T"][ip] / r)
        return lam if qs is None else (lam, lqs)

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
    def prims(self, q):
        g = self.gamma
        rho, rU, rV, rW, rE = q
        U, V, W = rU / rho, rV / rho, rW / rho
        p = (g - 1) * (rE - 0.5 * rho * (U ** 2 + V ** 2 + W ** 2))
        return {"U": U, "V": V, "W": W, "p": p, "T": p / (rho * self.R)}

    def intermediates(self, q, qs=None):
        """qs: the state of the X set (None: the stage state itself)"""
        rho, rU, rV, rW, rE = q
        v = {"rho": rho, "rU": rU, "rV": rV, "rW": rW, "rE": rE}
        v.update(self.prims(q))
        for k, val in self.prims(q if qs is None else qs).items():
            v[XSET[k]] = val
        v["mu"] = self.mu_ref * v["T_X"] ** 0.76
        v["amu"] = 0.76 * self.mu_ref * v["T_X"] ** (0.76 - 1.0)
        for f in FIRST:
            v[f] = self.op(1, DIR[f[1]], v[self.src(f)])
        for f in SECOND:
            v[f] = self.op(2, DIR[f[1]], v[self.src(f)])
        for m, (inner, d) in MIXED.items():
            v[m] = self.op(1, d, v[inner])
        for j in "xyz":
            v["Mu" + j.upper()] = v["amu"] * v["T" + j]
        for d, u in (("X", "U"), ("Y", "V"), ("Z", "W")):
            vel, pr = self.flux_prims(v, d, u)
            v["cont" + d] = self.op(1, DIR[d], {"X": rU, "Y": rV, "Z": rW}[d])
            fl = {"mx": rU * v[vel], "my": rV * v[vel], "mz": rW * v[vel],
                  "e": rE * v[vel] + v[vel] * v[pr]}
            fl[{"X": "mx", "Y": "my", "Z": "mz"}[d]] += v[pr]
            for k, t in fl.items():
                v[k + d] = self.op(1, DIR[d], t)
        return v

    @staticmethod
    def src(f):
        """the field a first or second derivative name differentiates: x
        derivatives read the X set"""
        return XSET[f[0]] if f[1] == "x" else f[0]

    @staticmethod
    def flux_prims(v, d, u):
        """names of the velocity and pressure in the fluxes along d"""
        return (XSET[u], "p_X") if d == "X" else (u, "p")

    def _coef(self, c):
        return self.cpPr if c == "cpPr" else c

    def rhs(self, q, dt, A=None, qs=None):
        """dt * RHS(q) with the sponge source toward the averages A, zeroed
        on the sponge face planes: the oracle's k2 for stage state q (qs: the
        X set's state at the first stage after a sponge update)"""
        v = self.intermediates(q, qs)
        out = []
        for c, terms in enumerate(RHS_TERMS):
            acc = np.zeros(self.shape)
            for coef, fac in terms:
                t = self._coef(coef)
                for f in fac:
                    t = t * v[f]
                acc = acc + t
            if self.sponge:
                spg = self.sigma * (A[c] - q[c])
                acc = acc + (2.0 * spg if c == 4 else spg)
            out.append(dt * acc)
        if self.walls:
            # postStepBCHandling: a wall plane keeps -dt D_n(rho u_n)
            for c in range(5):
                out[c] = np.where(self.rowdir >= 0, 0.0, out[c])
            for d, D in enumerate("XYZ"):
                out[0] = np.where(self.rowdir == d, -dt * v["cont" + D], out[0])
        # then the sponge face planes are zeroed, wall points on them too
        return [x * self.zmask for x in out]

    # ---------------------------------------------------------------- reverse
    def _chain(self, b, v, names, rho, out):
        """cotangents of the primitives `names` (U V W p T) of a state with
        density rho and velocities v[names U..W] into its conserved
        cotangents out (dict)"""
        g = self.gamma
        zero = np.zeros(self.shape)
        nU, nV, nW, np_, nT = (names[k] for k in ("U", "V", "W", "p", "T"))

        def add(k, val):
            out[k] = out.get(k, 0.0) + val

        Tb = b.get(nT, zero)
        pb = b.get(np_, zero) + Tb / (rho * self.R)
        add("rho", -Tb * v[nT] / rho)
        add("rE", (g - 1) * pb)
        add("rho", -(g - 1) * 0.5 * (v[nU] ** 2 + v[nV] ** 2 + v[nW] ** 2) * pb)
        for u, m in ((nU, "rU"), (nV, "rV"), (nW, "rW")):
            ub = b.get(u, zero) - (g - 1) * rho * v[u] * pb
            add(m, ub / rho)
            add("rho", -ub * v[u] / rho)

    def rhs_vjp_full(self, q, lam, dt, A=None, qs=None):
        """(d(dt R)/dq)^T lam, (d(dt R)/dqs)^T lam (None without qs) and
        (d(dt R)/dA)^T lam (None without sponges)"""
        v = self.intermediates(q, qs)
        b = {}

        def add(name, val):
            b[name] = b.get(name, 0.0) + val

        w = [dt * self.zmask * np.where(self.rowdir >= 0, 0.0, x) for x in lam]
        for d, D in enumerate("XYZ"):
            if (self.rowdir == d).any():
                b["cont" + D] = np.where(self.rowdir == d, -dt * lam[0] * self.zmask, 0.0)
        for c, terms in enumerate(RHS_TERMS):
            for coef, fac in terms:
                cf = self._coef(coef)
                for k, f in enumerate(fac):
                    t = cf * w[c]
                    for kk, ff in enumerate(fac):
                        if kk != k:
                            t = t * v[ff]
                    add(f, t)
        zero = np.zeros(self.shape)
        for d, u in (("X", "U"), ("Y", "V"), ("Z", "W")):
            dd = DIR[d]
            vel, pr = self.flux_prims(v, d, u)
            add({"X": "rU", "Y": "rV", "Z": "rW"}[d], self.op(1, dd, b.get("cont" + d, zero), True))
            for k, m in (("mx", "rU"), ("my", "rV"), ("mz", "rW")):
                tb = self.op(1, dd, b.get(k + d, zero), True)
                add(m, tb * v[vel])
                add(vel, tb * v[m])
                if k == {"X": "mx", "Y": "my", "Z": "mz"}[d]:
                    add(pr, tb)
            tb = self.op(1, dd, b.get("e" + d, zero), True)
            add("rE", tb * v[vel])
            add(vel, tb * (v["rE"] + v[pr]))
            add(pr, tb * v[vel])
        for j in "xyz":
            mb = b.get("Mu" + j.upper(), zero)
            add("amu", mb * v["T" + j])
            add("T" + j, mb * v["amu"])
        for m, (inner, d) in MIXED.items():
            add(inner, self.op(1, d, b.get(m, zero), True))
        for f in SECOND:
            add(self.src(f), self.op(2, DIR[f[1]], b.get(f, zero), True))
        for f in FIRST:
            add(self.src(f), self.op(1, DIR[f[1]], b.get(f, zero), True))
        TX = v["T_X"]
        add("T_X", b.get("mu", zero) * 0.76 * v["mu"] / TX)
        add("T_X", b.get("amu", zero) * (0.76 - 1.0) * v["amu"] / TX)
        # conserved terms, then the primitives of the stage state and of the
        # X set's state
        qb = {k: b.get(k, zero) for k in ("rho", "rU", "rV", "rW", "rE")}
        self._chain(b, v, {k: k for k in XSET}, q[0], qb)
        if qs is None:
            self._chain(b, v, XSET, q[0], qb)
            qsb = None
        else:
            d = {}
            self._chain(b, v, XSET, qs[0], d)
            qsb = [d.get(k, zero) for k in ("rho", "rU", "rV", "rW", "rE")]
        out = [qb[k] for k in ("rho", "rU", "rV", "rW", "rE")]
        Ab = None
        if self.sponge:
            Ab = []
            for c in range(5):
                s = (2.0 if c == 4 else 1.0) * self.sigma * w[c]
                out[c] = out[c] - s
                Ab.append(s)
        return out, qsb, Ab

    def rhs_vjp(self, q, lam, dt, A=None, qs=None):
        """(d(dt R)/dq)^T lam"""
        return self.rhs_vjp_full(q, lam, dt, A, qs)[0]

    # ---------------------------------------------------------------- sponge
    def sponge_update(self, qf, A, dt):
        """updateSponge (CSolver_AWS.cpp:1937-2031) after the filtered state
        qf: the averages relax toward qf with eps = 1/(avgT/dt + 1), the
        energy average blends toward the sponge pressure 1/gamma at the
        averages' density and momenta, and the face planes of the state are
        reset to the averages.  Returns (new state, new averages)."""
        eps = 1.0 / (self.avgT / dt + 1.0)
        a = [x + eps * (y - x) for x, y in zip(A, qf)]
        a[4] = self.epsP * a[4] + (1.0 - self.epsP) * (
            self.spP / (self.gamma - 1.0) + 0.5 * (a[1] ** 2 + a[2] ** 2 + a[3] ** 2) / a[0])
        qn = [np.where(self.onplane, x, y) for x, y in zip(a, qf)]
        return qn, a

    def sponge_update_vjp(self, qf, A, dt, lq, lA):
        """cotangents of (qf, A) from those of the new state lq and the new
        averages lA"""
        eps = 1.0 / (self.avgT / dt + 1.0)
        a = [x + eps * (y - x) for x, y in zip(A, qf)]
        lqf = [np.where(self.onplane, 0.0, x) for x in lq]
        la = [y + np.where(self.onplane, x, 0.0) for x, y in zip(lq, lA)]
        l4 = la[4]
        c = (1.0 - self.epsP) * l4
        la[4] = self.epsP * l4
        for m in (1, 2, 3):
            la[m] = la[m] + c * a[m] / a[0]
        la[0] = la[0] - c * 0.5 * (a[1] ** 2 + a[2] ** 2 + a[3] ** 2) / a[0] ** 2
        lAo = [(1.0 - eps) * x for x in la]
        lqf = [x + eps * y for x, y in zip(lqf, la)]
        return lqf, lAo

    # ---------------------------------------------------------------- RK4
    @staticmethod
    def filter_order(count):
        m = count % 3
        return (0, 1, 2) if m == 1 else (1, 2, 0) if m == 2 else (2, 0, 1)

    def rk4_step(self, q, dt, order, A=None, qs=None):
        """one RK4 step (RK4.hpp:35-118) followed by the filter in `order`
        (None: no filter); the filtered state before any sponge update.
        With walls every stage state is wall-corrected first (the first in
        place: the steps build on it)"""
        W = self.wall_pre if self.walls else (lambda x: x)
        q, qs = self.wall_first(q, qs)
        k1 = self.rhs(q, dt, A, qs)
        k2 = self.rhs(W([a + 0.5 * k for a, k in zip(q, k1)]), dt, A)
        k3 = self.rhs(W([a + 0.5 * k for a, k in zip(q, k2)]), dt, A)
        k4 = self.rhs(W([a + k for a, k in zip(q, k3)]), dt, A)
        acc = [a + k1_ / 6 + k2_ / 3 + k3_ / 3 + k4_ / 6
               for a, k1_, k2_, k3_, k4_ in zip(q, k1, k2, k3, k4)]
        return self.filt(acc, order) if order is not None else acc

    def tvdrk3_step(self, q, dt, order, A=None, qs=None):
        """one TVDRK3 step (TVDRK3.hpp:38-96: Q2 = q + k1, Q3 = 3/4 q +
        1/4 (Q2 + k2), q' = 1/3 q + 2/3 (Q3 + k3)) followed by the filter in
        `order` (None: no filter); with walls q, Q2 and Q3 wall-corrected"""
        W = self.wall_pre if self.walls else (lambda x: x)
        q, qs = self.wall_first(q, qs)
        k1 = self.rhs(q, dt, A, qs)
        Q2 = W([a + k for a, k in zip(q, k1)])
        k2 = self.rhs(Q2, dt, A)
        Q3 = W([0.75 * a + 0.25 * b + 0.25 * k for a, b, k in zip(q, Q2, k2)])
        k3 = self.rhs(Q3, dt, A)
        acc = [a / 3 + 2.0 / 3.0 * b + 2.0 / 3.0 * k for a, b, k in zip(q, Q3, k3)]
        return self.filt(acc, order) if order is not None else acc

    def wall_first(self, q, qs):
        """the first stage's wall correction: of the state, and of the X
        set's state with it at the first stage after a sponge update"""
        if not self.walls:
            return q, qs
        if qs is None:
            return self.wall_pre(q), None
        return self.wall_pre(q, qs=qs)

    def wall_first_vjp(self, q, qs, lam, lqs):
        """the reverse of wall_first: cotangents of (q, qs) (lqs None
        without qs)"""
        if not self.walls:
            return lam, lqs
        if qs is None:
            return self.wall_pre_vjp(q, lam), None
        return self.wall_pre_vjp(q, lam, qs, lqs)

    def step(self, q, dt, order, A=None, qs=None):
        if self.rk == 4:
            return self.rk4_step(q, dt, order, A, qs)
        return self.tvdrk3_step(q, dt, order, A, qs)

    def step_vjp_full(self, q, dt, order, lam, A=None, qs=None):
        """cotangents of (q, qs, A) of a step (before its sponge update)
        from that of its filtered result"""
        f = self.rk4_step_vjp if self.rk == 4 else self.tvdrk3_step_vjp
        return f(q, dt, order, lam, A, qs)

    def step_vjp(self, q, dt, order, lam):
        return self.step_vjp_full(q, dt, order, lam)[0]

    def conserved_from_ic(self, rho, U, V, W, p):
        g = self.gamma
        return [rho, rho * U, rho * V, rho * W, p / (g - 1) + 0.5 * rho * (U * U + V * V + W * W)]

    def run(self, q, nsteps, dt, filter_step=1):
        """the forward map of nsteps steps from state q (the averages start
        from q, as set_initial_conditions sets them); returns the final
        state, the final averages and the per-step records (state,
        averages, X-set state, filter order, filtered state)"""
        A = [x.copy() for x in q] if self.sponge else None
        qs = None
        recs = []
        fcount = 0
        for n in range(1, nsteps + 1):
            order = None
            if n % filter_step == 0:
                fcount += 1
                order = self.filter_order(fcount)
            qf = self.step(q, dt, order, A, qs)
            recs.append((q, A, qs, order, qf))
            if self.sponge:
                q, A = self.sponge_update(qf, A, dt)
                qs = qf
            else:
                q = qf
        return q, A, recs

    def ke_gradient(self, ic, nsteps, dt, filter_step=1, q0=None):
        """J = <rho |u|^2 / 2> after nsteps steps of the configured scheme
        (RK4 or TVDRK3) with fixed dt and the filter every filter_step
        steps; returns J, dJ/dU0, dJ/dV0, dJ/dW0 at fixed initial density
        and pressure.  q0: the initial conserved state (default: formed from
        ic; with walls pass the oracle's, whose set_initial_conditions also
        corrects the wall faces)"""
        if q0 is None:
            q0 = self.conserved_from_ic(*[np.asarray(a, dtype=np.float64) for a in ic])
        q, A, recs = self.run(q0, nsteps, dt, filter_step)
        N = q[0].size
        rho = q[0]
        u = [q[i] / rho for i in (1, 2, 3)]
        J = float(np.sum(0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / N)
        zero = np.zeros(self.shape)
        lam = [-0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / N, u[0] / N, u[1] / N, u[2] / N,
               zero]
        lA = [zero] * 5 if self.sponge else None
        lqs = None  # cotangent of a step's filtered state from the next step's X set
        for n in range(nsteps - 1, -1, -1):
            qn, An, qsn, order, qf = recs[n]
            if self.sponge:
                lam, lA = self.sponge_update_vjp(qf, An, dt, lam, lA)
                if lqs is not None:
                    lam = [a + b for a, b in zip(lam, lqs)]
            lam, lqs, lAs = self.step_vjp_full(qn, dt, order, lam, An, qsn)
            if lAs is not None:
                lA = [a + b for a, b in zip(lA, lAs)]
        if self.sponge:
            lam = [a + b for a, b in zip(lam, lA)]  # the averages start as q0
        # m_i = rho u_i, E = p/(g-1) + rho |u|^2 / 2 at fixed rho and p, with
        # rho and u of the initial state
        r0 = q0[0]
        grads = [r0 * (lam[i + 1] + (q0[i + 1] / r0) * lam[4]) for i in range(3)]
        return J, grads[0], grads[1], grads[2]

    @staticmethod
    def _acc(x, y):
        if y is None:
            return x
        return y if x is None else [a + b for a, b in zip(x, y)]

    def rk4_step_vjp(self, q, dt, order, lam, A=None, qs=None):
        """cotangents of the step's start state (and of qs, A) from that of
        its filtered result; with q1 = W(q), Q2 = q1 + k1/2, Q3 = q1 + k2/2,
        Q4 = q1 + k3, k_s = dt R(W(Q_s)) (W: the wall correction, the
        identity without walls) and the filtered acc = q1 + k1/6 + k2/3 +
        k3/3 + k4/6"""
        W = self.wall_pre if self.walls else (lambda x: x)
        WT = self.wall_pre_vjp
        q1, qs1 = self.wall_first(q, qs)
        k1 = self.rhs(q1, dt, A, qs1)
        Q2 = [a + 0.5 * k for a, k in zip(q1, k1)]
        k2 = self.rhs(W(Q2), dt, A)
        Q3 = [a + 0.5 * k for a, k in zip(q1, k2)]
        k3 = self.rhs(W(Q3), dt, A)
        Q4 = [a + k for a, k in zip(q1, k3)]
        la = self.filt(lam, order, transpose=True) if order is not None else list(lam)
        qb = list(la)
        Q4b, _, Ab = self.rhs_vjp_full(W(Q4), [x / 6 for x in la], dt, A)
        Q4b = WT(Q4, Q4b)
        qb = [a + c for a, c in zip(qb, Q4b)]
        Q3b, _, Ab3 = self.rhs_vjp_full(W(Q3), [x / 3 + c for x, c in zip(la, Q4b)], dt, A)
        Q3b = WT(Q3, Q3b)
        qb = [a + c for a, c in zip(qb, Q3b)]
        Q2b, _, Ab2 = self.rhs_vjp_full(W(Q2), [x / 3 + 0.5 * c for x, c in zip(la, Q3b)], dt, A)
        Q2b = WT(Q2, Q2b)
        qb = [a + c for a, c in zip(qb, Q2b)]
        Q1b, qsb, Ab1 = self.rhs_vjp_full(q1, [x / 6 + 0.5 * c for x, c in zip(la, Q2b)], dt, A,
                                          qs1)
        for x in (Ab3, Ab2, Ab1):
            Ab = self._acc(Ab, x)
        lq, lqs = self.wall_first_vjp(q, qs, [a + c for a, c in zip(qb, Q1b)], qsb)
        return lq, lqs, Ab

    def tvdrk3_step_vjp(self, q, dt, order, lam, A=None, qs=None):
        """cotangents of the step's start state (and of qs, A) from that of
        its filtered result for tvdrk3_step: with la the cotangent of the
        unfiltered result, Q3b = 2/3 la + J3^T (2/3 la), Q2b = 1/4 Q3b +
        J2^T (1/4 Q3b), qb = 1/3 la + 3/4 Q3b + Q2b + J1^T Q2b (J_s^T:
        rhs_vjp at Q_s); with walls each Q_s b taken back through the wall
        correction of its stage"""
        W = self.wall_pre if self.walls else (lambda x: x)
        WT = self.wall_pre_vjp
        q1, qs1 = self.wall_first(q, qs)
        k1 = self.rhs(q1, dt, A, qs1)
        Q2 = [a + k for a, k in zip(q1, k1)]
        Q2w = W(Q2)
        k2 = self.rhs(Q2w, dt, A)
        Q3 = [0.75 * a + 0.25 * b + 0.25 * k for a, b, k in zip(q1, Q2w, k2)]
        Q3w = W(Q3)
        la = self.filt(lam, order, transpose=True) if order is not None else list(lam)
        t3 = [2.0 / 3.0 * x for x in la]
        J3, _, Ab = self.rhs_vjp_full(Q3w, t3, dt, A)
        Q3b = WT(Q3, [a + c for a, c in zip(t3, J3)])
        t2 = [0.25 * x for x in Q3b]
        J2, _, Ab2 = self.rhs_vjp_full(Q2w, t2, dt, A)
        Q2b = WT(Q2, [a + c for a, c in zip(t2, J2)])
        J1, qsb, Ab1 = self.rhs_vjp_full(q1, Q2b, dt, A, qs1)
        for x in (Ab2, Ab1):
            Ab = self._acc(Ab, x)
        out = [x / 3 + 0.75 * a + b + c for x, a, b, c in zip(la, Q3b, Q2b, J1)]
        lq, lqs = self.wall_first_vjp(q, qs, out, qsb)
        return lq, lqs, Ab
