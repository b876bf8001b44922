"""FP64 restatement of the reference's 2-D fluid-solid coupled problem.

TEST INFRASTRUCTURE ONLY (like pyoracle): a checker for a coupled GPU path;
nothing in the product imports it.  Parity unpinned by the reference: the
reference holds no coupled test data, so only the single-physics parts are
pinned (they are the oracle's own, pinned elsewhere) together with the
checks in tests/test_oracle_coupled.py.

A mesh of quads whose elements are fluid (Scalar) or solid (Elastic2D); all
fields live on every node, fluid u, v, a, a_ and solid ux, uy, ... (the
fields of physicsToFields, Order2Newmark.cpp:7-30).  It is held as two
single-physics oracle problems (the fluid and the solid elements) joined at
their shared nodes:

* mass: one `mi` assembled over all elements (Order2Newmark.cpp:32-66), so
  at a shared node both fields use the inverse of the fluid plus the solid
  mass;
* per step (Problem.cpp:143-215; Main.cpp:45-66) every element adds
  f - Ku + s, where s is the coupling surface integral of coupled elements
  (elements with a vertex on an element of the other physics, Mesh.cpp:505-521):
  - fluid (ElasticToAcoustic.cpp:17-41): over every edge of the element
    that has a neighbouring element (Mesh.cpp:480-503, whatever its
    physics), the edge integral of (vx, vy) . n with n the element's
    outward edge normal (TensorQuad.cpp:400-418);
  - solid (AcousticToElastic2D.cpp:17-60): over every edge whose neighbour
    is fluid, -rho_0 times the edge integral of the fluid v, added to both
    the x and the y row (no normal), rho_0 the mean of the neighbour's
    vertex RHO (ExodusModel lookup at the neighbour's centre);
  - the edge integral (TensorQuad.cpp:474-515) weights the edge nodes by
    w_j |edge| / 2;
* the velocities in s are those after the previous step;
* Newmark as Order2Newmark.cpp:68-115 on every field; with
  `newmark_quirk` the reference's literal field list {ax, ay, ax, a}
  (Order2Newmark.cpp:93-96): the solid's x advanced twice, y once, the
  fluid's a once.
"""
import numpy as np

from oracle import pyoracle as po

# first node and stride of the local edge from vertex i to i+1 (TensorQuad.cpp:476-503)
def _edge_nodes(i, n):
    start, stride = {0: (0, 1), 1: (n - 1, n), 2: (n * n - 1, -1), 3: (n * (n - 1), -n)}[i]
    return [start + j * stride for j in range(n)]


class CoupledSim:
    def __init__(self, order, cells, coords, fluid, par_fluid, rho_fluid, par_solid,
                 newmark_quirk=False):
        """cells E x 4 (reference vertex order), coords V x 2, fluid[E] bool;
        par_fluid / rho_fluid: (fluid elements) x 4 x 1 VP and x 4 RHO;
        par_solid: (solid elements) x 4 x 5 (RHO, C11, C13, C33, C55)."""
        cells = np.asarray(cells)
        coords = np.asarray(coords, dtype=np.float64)
        fluid = np.asarray(fluid, dtype=bool)
        self.order, self.n = order, order + 1
        self.newmark_quirk = bool(newmark_quirk)
        self.cells, self.coords, self.fluid = cells, coords, fluid
        self.ef = np.flatnonzero(fluid)
        self.es = np.flatnonzero(~fluid)
        self.rho_fluid = np.asarray(rho_fluid, dtype=np.float64)
        self.part = {}
        for key, idx, phys, par in (("f", self.ef, po.ACOUSTIC, par_fluid), ("s", self.es, po.ELASTIC, par_solid)):
            if len(idx) == 0:
                continue
            sub = cells[idx]
            used = np.unique(sub)
            remap = np.full(len(coords), -1)
            remap[used] = np.arange(len(used))
            S = po.Sim(2, order, phys, remap[sub], coords[used], par)
            self.part[key] = dict(S=S, idx=idx, conn=S.conn(), mi=S.inverse_mass(), xy=S.node_coords(),
                                  c=S.c, G=S.G, u=np.zeros((S.c, S.G)), v=np.zeros((S.c, S.G)),
                                  ap=np.zeros((S.c, S.G)), src=[])
        # shared nodes and the combined mass
        self.f2s, self.s2f = {}, {}
        if len(self.part) == 2:
            F, Sd = self.part["f"], self.part["s"]
            key = {tuple(np.round(p, 6)): g for g, p in enumerate(Sd["xy"])}
            for gf, p in enumerate(F["xy"]):
                gs = key.get(tuple(np.round(p, 6)))
                if gs is not None:
                    self.f2s[gf], self.s2f[gs] = gs, gf
            for gf, gs in self.f2s.items():
                m = 1.0 / F["mi"][gf] + 1.0 / Sd["mi"][gs]
                F["mi"][gf] = Sd["mi"][gs] = 1.0 / m
        # element edges, neighbours and the coupled elements' edge lists
        owners = {}
        for e, cv in enumerate(cells):
            for i in range(4):
                owners.setdefault(frozenset((cv[i], cv[(i + 1) % 4])), []).append(e)
        vtx_fluid = np.zeros(len(coords), dtype=bool)
        vtx_solid = np.zeros(len(coords), dtype=bool)
        vtx_fluid[cells[fluid].ravel()] = True
        vtx_solid[cells[~fluid].ravel()] = True
        self.edges = {"f": [], "s": []}  # (local element index, edge i, rho_0)
        for key, idx in (("f", self.ef), ("s", self.es)):
            other = vtx_solid if key == "f" else vtx_fluid
            for le, e in enumerate(idx):
                cv = cells[e]
                if not other[cv].any():
                    continue  # no coupling field in the element's closure
                for i in range(4):
                    nb = [o for o in owners[frozenset((cv[i], cv[(i + 1) % 4]))] if o != e]
                    for o in nb:
                        if key == "f":
                            self.edges["f"].append((le, i, 0.0))
                        elif fluid[o]:
                            lo = int(np.searchsorted(self.ef, o))
                            self.edges["s"].append((le, i, float(self.rho_fluid[lo].mean())))

    def edge_geometry(self, e, i):
        """(half length, outward unit normal) of local edge i of element e."""
        cv = self.cells[e]
        (x0, y0), (x1, y1) = self.coords[cv[i]], self.coords[cv[(i + 1) % 4]]
        nrm = np.array([y1 - y0, -(x1 - x0)])
        length = np.hypot(x1 - x0, y1 - y0)
        return length / 2.0, nrm / length

    def add_source(self, part, loc, amp, f0, t0, direction=None):
        P = self.part[part]
        k = P["S"].add_source(loc, amp, f0, t0, direction)
        e, _ = P["S"].source_info(k)
        P["src"].append((e, P["S"].source_load(k), amp, f0, t0))

    def surface(self):
        """Coupling terms s of the current velocities, per part (c x G)."""
        out = {k: np.zeros((P["c"], P["G"])) for k, P in self.part.items()}
        if len(self.part) < 2:
            return out
        F, Sd = self.part["f"], self.part["s"]
        w = po.gll(self.order)[1]
        for le, i, _ in self.edges["f"]:
            half, nrm = self.edge_geometry(self.ef[le], i)
            for j, q in enumerate(_edge_nodes(i, self.n)):
                gf = F["conn"][le][q]
                gs = self.f2s.get(gf)
                vn = 0.0 if gs is None else Sd["v"][0, gs] * nrm[0] + Sd["v"][1, gs] * nrm[1]
                out["f"][0, gf] += vn * half * w[j]
        for le, i, rho0 in self.edges["s"]:
            half, _ = self.edge_geometry(self.es[le], i)
            for j, q in enumerate(_edge_nodes(i, self.n)):
                gs = Sd["conn"][le][q]
                gf = self.s2f.get(gs)
                vf = 0.0 if gf is None else F["v"][0, gf]
                val = -rho0 * vf * half * w[j]
                out["s"][0, gs] += val
                out["s"][1, gs] += val
        return out

    def run(self, dt, nsteps):
        acl, dsp = 0.5 * dt, 0.5 * dt * dt
        t = 0.0
        for _ in range(nsteps):
            s = self.surface()
            acc = {}
            for key, P in self.part.items():
                a = -P["S"].apply_stiffness(P["u"]) + s[key]
                for e, load, amp, f0, t0 in P["src"]:
                    a[:, P["conn"][e]] += load * po.ricker(t, amp, f0, t0)
                acc[key] = a * P["mi"][None, :]
            for key, P in self.part.items():
                a = acc[key]
                # components in the order the reference's field list visits them
                comps = [0, 1, 0] if (self.newmark_quirk and P["c"] >= 2) else range(P["c"])
                for k in comps:
                    P["v"][k] += acl * a[k] + acl * P["ap"][k]
                    P["u"][k] += dt * P["v"][k] + dsp * a[k]
                    P["ap"][k] = a[k]
            t += dt
        return {k: P["u"].copy() for k, P in self.part.items()}
