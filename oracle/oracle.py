"""ctypes loader for the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product. See oracle.cpp for what each function restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KIND_NONE, KIND_BB, KIND_FLUID = 0, 1, 2

# the hook-variant oracle: built with the reference's own example kinetics headers
# (`make -C oracle ref-hooks`, only where the reference tree is present)
REF_HOOKS_PATH = os.path.join(HERE, "_ref", "liboracle_ref_hooks.so")

_libs = {}


def load(path=None):
    """The oracle library at path (default: liboracle.so, built with the package's own
    kinetics plugins); each is loaded once."""
    path = LIB_PATH if path is None else path
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    vp, ci, cd, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    sigs = {
        "orc_flow_init": ([ci, ci, ci, cd, cd, vp], None),
        "orc_flow_run": ([ci, ci, ci, vp, vp, vp, cd, cd, vp, ci, vp], None),
        "orc_flow_velocity": ([ci, ci, ci, vp, vp, cd, cd, vp, vp, vp], None),
        "orc_speciate_pcf": ([ci, ci, vp, vp, vp, cl, vp, cd, ci, vp, vp, vp, vp, vp], None),
        "orc_pcf_residual": ([ci, ci, vp, vp, cl, vp, vp, vp], None),
        "orc_pcf_residual_var": ([ci, ci, vp, vp, ci, vp, vp, cl, vp, vp, vp], None),
        "orc_component_totals": ([ci, ci, vp, cl, vp, vp], None),
        "orc_speciate_newton_reserve": ([ci, ci, vp, vp, vp, cl, vp, cd, ci, vp, vp, vp, vp], None),
        "orc_ade_init": ([cl, ci, vp, vp], None),
        "orc_ade_density": ([cl, ci, vp, vp], None),
        "orc_ade_run": ([vp, vp, vp, vp, vp, ci, vp], None),
        "orc_node_conversion": ([cl, vp, vp, vp, ci, ci, ci, ci, ci, cd, cd], cl),
        "orc_ca_init": ([cl, vp, ci, ci, ci, vp, vp, vp], None),
        "orc_ca_redistribute": ([ci, ci, ci, ci, vp, vp, vp, vp, vp, ci, ci, ci, ci, cd, cd, ci,
                                 ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp], None),
        "orc_ca_update_age": ([ci, ci, ci, vp, vp, vp, cd, ci], ci),
        "orc_fd_biomass": ([ci, ci, ci, ci, vp, vp, vp, ci, ci, ci, ci, cd, cd, vp, vp, vp], None),
        "orc_classify_solid": ([ci, ci, ci, vp, vp, ci, ci], None),
        "orc_mlp_forward": ([ci, ci, ci, vp, cl, vp, vp], None),
        "orc_mlp_grad": ([ci, ci, ci, vp, cl, vp, vp, vp], cd),
        "orc_kinetics_plugin": ([vp, ci, vp, ci, vp, vp, ctypes.c_longlong], None),
        "orc_abiotic_plugin": ([vp, ci, vp, ctypes.c_longlong], None),
        "orc_num_threads": ([], ci),
        "orc_srg_net_eval": ([vp, vp, ci], cd),
        "orc_srg_voxel": ([vp, ci, ci, vp, vp, cd, vp, vp], None),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _libs[path] = lib
    return lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class AdeParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("S", ctypes.c_int),
        ("kind", ctypes.c_void_p), ("omega", ctypes.c_void_p), ("bc_type", ctypes.c_void_p),
        ("bc_value", ctypes.c_void_p), ("u", ctypes.c_void_p),
        ("solid_code", ctypes.c_int), ("bb_code", ctypes.c_int),
        ("kinetics", ctypes.c_void_p), ("dt", ctypes.c_double), ("nb", ctypes.c_int),
        ("nc", ctypes.c_int), ("eqS", ctypes.c_void_p), ("eqLogK", ctypes.c_void_p),
        ("eqCompSpecies", ctypes.c_void_p), ("eqTol", ctypes.c_double), ("eqMaxIter", ctypes.c_int),
        ("mask_update", ctypes.c_int), ("bmass_threshold", ctypes.c_double),
        ("pore_code", ctypes.c_int), ("biofilm_code", ctypes.c_int),
        ("immobile", ctypes.c_void_p), ("abiotic", ctypes.c_void_p), ("abiotic_surface", ctypes.c_int),
        ("dissol", ctypes.c_void_p), ("phase", ctypes.c_void_p), ("phase_sub", ctypes.c_void_p),
        ("dissol_surface", ctypes.c_int),
        ("srg", ctypes.c_void_p),
        ("mask_rule", ctypes.c_void_p),
    ]


def mask_rule(pore_codes, bio_codes, sum_labels=True):
    """The rule of several material types for the oracle: {npore, pore[8], bio[4], sum_labels}
    (int32), as the step and the CA functions take it."""
    v = np.zeros(14, dtype=np.int32)
    v[0] = len(pore_codes)
    v[1:1 + len(pore_codes)] = pore_codes
    v[9:9 + len(bio_codes)] = bio_codes
    v[13] = 1 if sum_labels else 0
    return v


class SrgNetParams(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int), ("in_sub", ctypes.c_void_p), ("n_layers", ctypes.c_int),
        ("width", ctypes.c_void_p), ("params", ctypes.c_void_p), ("x_offset", ctypes.c_void_p),
        ("x_gain", ctypes.c_void_p), ("x_ymin", ctypes.c_double), ("y_ymin", ctypes.c_double),
        ("y_gain", ctypes.c_double), ("y_offset", ctypes.c_double), ("log_output", ctypes.c_int),
        ("closure", ctypes.c_int), ("yield_", ctypes.c_void_p),
    ]


class SrgParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("id", ctypes.c_void_p), ("Kc", ctypes.c_void_p), ("decay", ctypes.c_void_p),
        ("vmax", ctypes.c_void_p), ("maxup", ctypes.c_void_p), ("mw", ctypes.c_void_p),
        ("fixc", ctypes.c_void_p), ("fixlb", ctypes.c_void_p), ("use_totals", ctypes.c_int),
        ("nc", ctypes.c_int), ("eqS", ctypes.c_void_p), ("eqCompSpecies", ctypes.c_void_p),
        ("nets", ctypes.c_void_p),
    ]


# ------------------------------------------------------------------ flow
def flow_run(nx, ny, nz, mask, kind, omega, rho_in, rho_out, nsteps, energy=False, f=None):
    lib = load()
    N = nx * ny * nz
    if f is None:
        f = np.zeros(19 * N)
        lib.orc_flow_init(nx, ny, nz, rho_in, rho_out, _p(f))
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    k = np.ascontiguousarray(kind, dtype=np.int32)
    o = _f(omega)
    e = np.zeros(nsteps) if energy else None
    lib.orc_flow_run(nx, ny, nz, _p(m), _p(k), _p(o), rho_in, rho_out, _p(f), nsteps, _p(e))
    return f, e


def flow_velocity(nx, ny, nz, mask, kind, rho_in, rho_out, f):
    lib = load()
    N = nx * ny * nz
    u = np.zeros(3 * N)
    r = np.zeros(N)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    k = np.ascontiguousarray(kind, dtype=np.int32)
    lib.orc_flow_velocity(nx, ny, nz, _p(m), _p(k), rho_in, rho_out, _p(f), _p(u), _p(r))
    return u.reshape(N, 3), r


# ------------------------------------------------------------------ speciation
def speciate_pcf(S, logK, comp_species, conc, tol=1e-10, max_iter=200):
    lib = load()
    S = _f(S)
    ns, nc = S.shape
    conc = _f(conc)
    n = conc.shape[0]
    out = np.zeros_like(conc)
    it = np.zeros(n, dtype=np.int32)
    r = np.zeros(n)
    cv = np.zeros(n, dtype=np.uint8)
    lc = np.zeros((n, nc))
    cs = np.ascontiguousarray(comp_species, dtype=np.int32)
    lib.orc_speciate_pcf(nc, ns, _p(S), _p(_f(logK)), _p(cs), n, _p(conc), tol, max_iter, _p(out), _p(it),
                         _p(r), _p(cv), _p(lc))
    return out, it, r, cv, lc


def speciate_newton_reserve(S, logK, comp_species, conc, tol=1e-12, max_iter=100):
    lib = load()
    S = _f(S)
    ns, nc = S.shape
    conc = _f(conc)
    n = conc.shape[0]
    out = np.zeros_like(conc)
    it = np.zeros(n, dtype=np.int32)
    r = np.zeros(n)
    cv = np.zeros(n, dtype=np.uint8)
    cs = np.ascontiguousarray(comp_species, dtype=np.int32)
    lib.orc_speciate_newton_reserve(nc, ns, _p(S), _p(_f(logK)), _p(cs), n, _p(conc), tol, max_iter, _p(out),
                                    _p(it), _p(r), _p(cv))
    return out, it, r, cv


def pcf_residual(S, logK, logC, T, var_species=None, logk_var=None):
    """The reference's judge per voxel; var_species / logk_var[n, nvar] give a per-voxel logK
    to those species."""
    lib = load()
    S = _f(S)
    ns, nc = S.shape
    logC = _f(logC)
    T = _f(T)
    n = logC.shape[0]
    f = np.zeros((n, nc))
    if var_species is None:
        lib.orc_pcf_residual(nc, ns, _p(S), _p(_f(logK)), n, _p(logC), _p(T), _p(f))
    else:
        vs = np.ascontiguousarray(var_species, dtype=np.int32)
        lkv = _f(logk_var)
        assert lkv.shape == (n, len(vs))
        lib.orc_pcf_residual_var(nc, ns, _p(S), _p(_f(logK)), len(vs), _p(vs), _p(lkv), n, _p(logC), _p(T),
                                 _p(f))
    return f


def component_totals(S, conc):
    lib = load()
    S = _f(S)
    ns, nc = S.shape
    conc = _f(conc)
    n = conc.shape[0]
    T = np.zeros((n, nc))
    lib.orc_component_totals(nc, ns, _p(S), n, _p(conc), _p(T))
    return T


# ------------------------------------------------------------------ transport + reactions
_KIN_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_longlong)


def kinetics_plugin_ptr(library=None):
    lib = load(library)
    return ctypes.cast(lib.orc_kinetics_plugin, ctypes.c_void_p).value


def abiotic_plugin_ptr(library=None):
    lib = load(library)
    return ctypes.cast(lib.orc_abiotic_plugin, ctypes.c_void_p).value


def dissolution_plugin_ptr(library=None):
    lib = load(library)
    return ctypes.cast(lib.orc_dissolution_plugin, ctypes.c_void_p).value


def abiotic_plugin(C, mask_code, library=None):
    lib = load(library)
    C = _f(C)
    sr = np.zeros(C.size)
    lib.orc_abiotic_plugin(_p(C), C.size, _p(sr), mask_code)
    return sr


def kinetics_plugin(B, C, mask_code, library=None):
    lib = load(library)
    B = _f(B)
    C = _f(C)
    sr = np.zeros(C.size)
    br = np.zeros(B.size)
    lib.orc_kinetics_plugin(_p(B), B.size, _p(C), C.size, _p(sr), _p(br), mask_code)
    return sr, br


def node_conversion(mask, phase, mineral, pore_code, solid_code, bb_code, precip_code, phase_id, precip_max,
                    reopen):
    """Precipitation / dissolution node conversion on copies of mask and phase; returns
    (mask, phase, number of voxels whose mask code changed)."""
    lib = load()
    m = np.ascontiguousarray(mask, dtype=np.uint8).copy()
    ph = np.ascontiguousarray(phase, dtype=np.uint8).copy()
    mi = _f(mineral).ravel()
    n = lib.orc_node_conversion(m.size, _p(m), _p(ph), _p(mi), pore_code, solid_code, bb_code, precip_code, phase_id,
                                precip_max, reopen)
    return m, ph, n


def ca_init(mask, solid=0, bb=1, pore=2, rule=None):
    """createAgeDomain3D and the distance field's sign: (age int32, dpos uint8). rule: a
    mask_rule(...) of several pore types, or None (pore)."""
    lib = load()
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    age = np.zeros(m.size, dtype=np.int32)
    dpos = np.zeros(m.size, dtype=np.uint8)
    lib.orc_ca_init(m.size, _p(m), solid, bb, pore, _p(age), _p(dpos), None if rule is None else _p(rule))
    return age, dpos


class CaState:
    """The CA's lattices for the oracle: biomass (nb, N) FP64, mask, age, distance sign and
    total biomass, advanced in place by redistribute()."""

    def __init__(self, nx, ny, nz, b, mask, solid=0, bb=1, pore=2, biofilm=3, bthr=10.0, rule=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.b = np.ascontiguousarray(np.atleast_2d(b), dtype=np.float64).copy()
        self.mask = np.ascontiguousarray(mask, dtype=np.uint8).ravel().copy()
        self.rule = rule
        self.age, self.dpos = ca_init(self.mask, solid, bb, pore, rule)
        self.btot = np.zeros(self.mask.size)
        self.codes = (solid, bb, pore, biofilm)
        self.bthr = bthr
        self.sweep = 0

    def redistribute(self, bmax, half=0, seed=1):
        """The sweep loop; returns (sweeps, age passes, status, mask flips, max before, max after)."""
        lib = load()
        rep = np.zeros(4, dtype=np.int64)
        mx = np.zeros(2)
        s, bb, p, bf = self.codes
        lib.orc_ca_redistribute(self.nx, self.ny, self.nz, self.b.shape[0], _p(self.b), _p(self.mask), _p(self.age),
                                _p(self.dpos), _p(self.btot), s, bb, p, bf, self.bthr, bmax, half, seed, self.sweep,
                                _p(rep), _p(mx), None if self.rule is None else _p(self.rule))
        self.sweep += int(rep[0])
        return int(rep[0]), int(rep[1]), int(rep[2]), int(rep[3]), float(mx[0]), float(mx[1])

    def fd_biomass(self, nu):
        """The FD biofilm solver's block: mask/totals, one fdDiffusion3D step, mask/totals;
        returns (mask flips, max total before, max total after)."""
        rep = np.zeros(1, dtype=np.int64)
        mx = np.zeros(2)
        s, bb, p, bf = self.codes
        load().orc_fd_biomass(self.nx, self.ny, self.nz, self.b.shape[0], _p(self.b), _p(self.mask), _p(self.btot),
                              s, bb, p, bf, self.bthr, nu, _p(rep), _p(mx),
                              None if self.rule is None else _p(self.rule))
        return int(rep[0]), float(mx[0]), float(mx[1])

    def update_age(self, bmax, max_passes=1000):
        return load().orc_ca_update_age(self.nx, self.ny, self.nz, _p(self.age), _p(self.dpos), _p(self.btot), bmax,
                                        max_passes)


class AdeRun:
    """Holds the oracle's state for a coupled transport run (populations, biomass, mask)."""

    def __init__(self, nx, ny, nz, mask, kind, omega, bc_type, bc_value, u, c0, solid_code=0, bb_code=1,
                 kinetics=False, dt=1.0, biomass=None, tableau=None, eq_tol=1e-10, eq_max_iter=200,
                 mask_update=None, immobile=None, abiotic=0, dissolution=None, surrogate=None, library=None,
                 rule=None):
        self.lib = load(library)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = N = nx * ny * nz
        self.mask = np.ascontiguousarray(mask, dtype=np.uint8).copy()
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        self.omega = _f(omega).ravel()
        self.S = self.omega.size // 256
        self.bc_type = np.ascontiguousarray(bc_type, dtype=np.int32).ravel()
        self.bc_value = _f(bc_value).ravel()
        self.u = _f(u).ravel()
        self.g = np.zeros(self.S * 7 * N)
        c0 = _f(c0)
        self.lib.orc_ade_init(N, self.S, _p(c0), _p(self.g))
        self.bc_rho = np.repeat(c0, 2 * ny * nz).copy()
        self.bmass = None if biomass is None else _f(biomass).ravel().copy()
        nb = 0 if biomass is None else self.bmass.size // N
        P = AdeParams()
        P.nx, P.ny, P.nz, P.S = nx, ny, nz, self.S
        P.kind, P.omega = _p(self.kind).value, _p(self.omega).value
        P.bc_type, P.bc_value, P.u = _p(self.bc_type).value, _p(self.bc_value).value, _p(self.u).value
        P.solid_code, P.bb_code = solid_code, bb_code
        P.kinetics = kinetics_plugin_ptr(library) if kinetics else None
        P.dt, P.nb = dt, nb
        if tableau is not None:
            Sm, lk, cs = tableau
            self.eqS = _f(Sm)
            self.eqLk = _f(lk)
            self.eqCs = np.ascontiguousarray(cs, dtype=np.int32)
            P.nc = self.eqS.shape[1]
            P.eqS, P.eqLogK, P.eqCompSpecies = _p(self.eqS).value, _p(self.eqLk).value, _p(self.eqCs).value
            P.eqTol, P.eqMaxIter = eq_tol, eq_max_iter
        else:
            P.nc = 0
        if mask_update is not None:
            thr, pore, bio = mask_update
            P.mask_update, P.bmass_threshold, P.pore_code, P.biofilm_code = 1, thr, pore, bio
        self.rule = rule  # mask_rule(...) of several material types, kept alive with P
        if rule is not None:
            P.mask_rule = _p(rule).value
        self.immobile = np.zeros(self.S, dtype=np.int32) if immobile is None else \
            np.ascontiguousarray(immobile, dtype=np.int32)
        P.immobile = _p(self.immobile).value
        if abiotic:
            P.abiotic = abiotic_plugin_ptr(library)
            P.abiotic_surface = 1 if abiotic == 2 else 0
        if dissolution is not None:
            # (phase ids per voxel, {phase id: immobile solute}, surface_only)
            ph, subs, surf = dissolution
            self.phase = np.ascontiguousarray(ph, dtype=np.uint8).ravel().copy()
            self.phase_sub = np.full(16, -1, dtype=np.int32)
            for pid, sub in subs.items():
                self.phase_sub[pid] = sub
            P.dissol = dissolution_plugin_ptr(library)
            P.phase, P.phase_sub = _p(self.phase).value, _p(self.phase_sub).value
            P.dissol_surface = 1 if surf else 0
        if surrogate is not None:
            self.surrogate = surrogate  # keeps its arrays alive
            P.srg = ctypes.addressof(surrogate.P)
        self.P = P
        self.eq_stats = np.zeros(3, dtype=np.int64)

    def set_field(self, solute, c):
        """One solute's field at rest: an immobile solute in every voxel, a mobile one in the
        fluid voxels (f_i - t_i = t_i (c - 1))."""
        c = _f(c).ravel()
        w = np.array([0.25] + [0.125] * 6)
        sel = np.ones(self.N, dtype=bool) if self.immobile[solute] else (self.kind[self.mask] == KIND_FLUID)
        for i in range(7):
            blk = self.g[(solute * 7 + i) * self.N:(solute * 7 + i + 1) * self.N]
            blk[sel] = w[i] * (c[sel] - 1.0)

    def run(self, nsteps):
        self.lib.orc_ade_run(ctypes.byref(self.P), _p(self.mask), _p(self.g), _p(self.bmass), _p(self.bc_rho),
                             nsteps, _p(self.eq_stats))

    def conc(self):
        """Concentration per node. On fluid nodes of the x faces the entering population is
        only rebuilt at the next collision, so these report what the boundary prescribes
        (the Dirichlet value, or for a Neumann face the density it copies), as the
        boundary dynamics' computeDensity does. Wall nodes report 0 for mobile solutes:
        Palabos' BounceBack::computeDensity returns a fixed fictitious density (0 by
        default), not the sum of the reflected populations the node holds. Immobile
        solutes report their amount everywhere (a precipitated voxel keeps its mineral)."""
        c = np.zeros(self.S * self.N)
        self.lib.orc_ade_density(self.N, self.S, _p(self.g), _p(c))
        c = c.reshape(self.S, self.nz, self.ny, self.nx)
        fluid = (self.kind[self.mask] == KIND_FLUID).reshape(self.nz, self.ny, self.nx)
        rho = self.bc_rho.reshape(self.S, 2, self.nz, self.ny)
        for s in range(self.S):
            if self.immobile[s]:
                continue
            c[s][~fluid] = 0.0
            for side, xb in ((0, 0), (1, self.nx - 1)):
                val = self.bc_value[2 * s + side] if self.bc_type[2 * s + side] == 0 else rho[s, side]
                c[s, :, :, xb] = np.where(fluid[:, :, xb], val, c[s, :, :, xb])
        return c.reshape(self.S, self.N)


# ------------------------------------------------------------------ surrogate processor
class Surrogate:
    """The surrogate processor's settings for the oracle (surrogate.cpp). nets: one entry per
    microbe, a dict (n_in, in_sub, widths, params, x_offset, x_gain, x_ymin, y_ymin, y_gain,
    y_offset, log_output; optional closure, yield) or None for no network."""

    def __init__(self, S, ids, Kc, decay, nets, vmax=None, max_uptake=None, molar_mass=None, fix_conc=None,
                 fix_lb=None, use_totals=False, tableau=None):
        n = len(ids)
        self._keep = []
        def arr(a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt).ravel()
            self._keep.append(a)
            return a.ctypes.data
        P = SrgParams()
        P.n = n
        P.id = arr(ids, np.int32)
        P.Kc = arr(np.reshape(Kc, (n, S)))
        P.decay = arr(decay)
        P.vmax = arr(np.zeros((n, S)) if vmax is None else np.abs(np.reshape(vmax, (n, S))))
        P.maxup = arr(np.full((n, S), -1e30) if max_uptake is None else np.reshape(max_uptake, (n, S)))
        P.mw = arr(np.full(n, 24.6) if molar_mass is None else molar_mass)
        P.fixc = arr(np.zeros(S) if fix_conc is None else fix_conc, np.int32)
        P.fixlb = arr(np.zeros(S) if fix_lb is None else fix_lb, np.int32)
        P.use_totals = 1 if use_totals else 0
        if tableau is not None:
            Sm, cs = tableau
            Sm = np.asarray(Sm, dtype=np.float64)
            P.nc = Sm.shape[1]
            P.eqS = arr(Sm)
            P.eqCompSpecies = arr(cs, np.int32)
        self.net_structs = (SrgNetParams * max(n, 1))()
        for k, net in enumerate(nets):
            q = self.net_structs[k]
            if net is None:
                q.n_layers = 0
                continue
            q.n_in = net["n_in"]
            q.in_sub = arr(net["in_sub"], np.int32)
            q.n_layers = len(net["widths"])
            q.width = arr(net["widths"], np.int32)
            q.params = arr(net["params"])
            q.x_offset = arr(net["x_offset"])
            q.x_gain = arr(net["x_gain"])
            q.x_ymin, q.y_ymin, q.y_gain, q.y_offset = net["x_ymin"], net["y_ymin"], net["y_gain"], net["y_offset"]
            q.log_output = 1 if net.get("log_output") else 0
            q.closure = net.get("closure", 0)
            q.yield_ = arr(net["yield"]) if net.get("yield") is not None else None
        P.nets = ctypes.addressof(self.net_structs)
        self.P = P
        self.S = S

    def voxel(self, conc, bmass, dt):
        """One voxel's deposited (dC[S], dB[nb]) from its post-collision state."""
        conc = _f(conc)
        bmass = _f(bmass)
        dC = np.zeros(self.S)
        dB = np.zeros(bmass.size)
        load().orc_srg_voxel(ctypes.byref(self.P), self.S, bmass.size, _p(conc), _p(bmass), dt, _p(dC), _p(dB))
        return dC, dB

    def growth(self, k, fin):
        """Microbe k's network on an uptake row fin[S]."""
        fin = _f(fin)
        return load().orc_srg_net_eval(ctypes.byref(self.net_structs[k]), _p(fin), fin.size)


# ------------------------------------------------------------------ MLP
def mlp_forward(din, H, dout, theta, X):
    lib = load()
    X = _f(X)
    Y = np.zeros((X.shape[0], dout))
    lib.orc_mlp_forward(din, H, dout, _p(_f(theta)), X.shape[0], _p(X), _p(Y))
    return Y


def mlp_grad(din, H, dout, theta, X, T):
    lib = load()
    theta = _f(theta)
    X = _f(X)
    T = _f(T)
    g = np.zeros_like(theta)
    loss = lib.orc_mlp_grad(din, H, dout, _p(theta), X.shape[0], _p(X), _p(T), _p(g))
    return g, loss


def num_threads():
    return load().orc_num_threads()
