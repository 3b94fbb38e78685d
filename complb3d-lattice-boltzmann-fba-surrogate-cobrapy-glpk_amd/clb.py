"""Python binding of the C ABI in include/clb.h (libclb.so, built in this directory).

This is the host-side mirror used by the tests and the benchmark. It loads the HIP
library and nothing else: there is no CPU fallback, and every call that needs the GPU
raises ClbError when the library or the device is missing.

Names follow the reference's operator surface where one exists:
  * ``EquilibriumChemistry`` mirrors the setters / calculate_species_concentrations of
    EquilibriumChemistry<T> (src/complab3d_processors_part4_eqsolver.hh:186-1214), solved
    by the GPU Newton kernel.
  * ``Lattice.step`` runs the reference's per-iteration order (src/complab.cpp:1252-1849).
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclb.so")

CLB_OK = 0
PREC_FP64, PREC_FP32 = 0, 1
KIND_NONE, KIND_BB, KIND_FLUID = 0, 1, 2
BC_DIRICHLET, BC_NEUMANN = 0, 1
KINETICS_NONE, KINETICS_PLUGIN = 0, 1
WARM_REF, WARM_USER, WARM_MLP = 0, 1, 2
BATCH_AUTO, BATCH_FULL, BATCH_POLISH = 0, 1, 2
SPEC_FUSED, SPEC_SEPARATE = 0, 1
CA_DONE, CA_STALLED, CA_SWEEP_CAP, CA_AGE_CAP, CA_NONFINITE = 0, 1, 2, 3, 4
KERNEL_FLOW, KERNEL_ADE, KERNEL_FUSED = 0, 1, 2
LAYOUT_COMPACT, LAYOUT_DENSE = 0, 1


class ClbError(RuntimeError):
    pass


class TuneReport(ctypes.Structure):
    _fields_ = [("flow_bx", ctypes.c_int), ("flow_by", ctypes.c_int), ("flow_ms", ctypes.c_double),
                ("ade_bx", ctypes.c_int), ("ade_by", ctypes.c_int), ("ade_ms", ctypes.c_double),
                ("fused_bx", ctypes.c_int), ("fused_by", ctypes.c_int), ("fused_budget", ctypes.c_int),
                ("fused_ms", ctypes.c_double), ("trials", ctypes.c_int)]


class CaReport(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_int), ("age_passes", ctypes.c_int), ("status", ctypes.c_int),
                ("mask_changes", ctypes.c_longlong), ("bmax_before", ctypes.c_double),
                ("bmax_after", ctypes.c_double)]


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


class Perf(ctypes.Structure):
    _fields_ = [("flow_ms", ctypes.c_double), ("flow_launches", ctypes.c_longlong),
                ("ade_ms", ctypes.c_double), ("ade_launches", ctypes.c_longlong),
                ("spec_ms", ctypes.c_double), ("spec_launches", ctypes.c_longlong),
                ("mlp_ms", ctypes.c_double), ("mlp_launches", ctypes.c_longlong),
                ("fluid_nodes", ctypes.c_longlong), ("reactive_nodes", ctypes.c_longlong),
                ("srg_ms", ctypes.c_double), ("srg_launches", ctypes.c_longlong),
                ("stragglers", ctypes.c_longlong)]


_libs = {}


def load(path=None):
    """Load libclb.so, or the library at path: a build of the same sources with other kinetics
    plugins compiled in (make KINETICS=... ABIOTIC=... OUT=path). Each is loaded once; raises
    ClbError if it has not been built."""
    path = LIB_PATH if path is None else os.path.abspath(path)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise ClbError(f"{path} is missing: build it first (__graft_entry__.build() or make)")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    c_int, c_double, c_long, c_u64 = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_uint64
    vp = ctypes.c_void_p
    sig = {
        "clb_create": [P(_Config), P(vp)],
        "clb_destroy": [vp],
        "clb_sync": [vp],
        "clb_device_name": [ctypes.c_char_p, c_int],
        "clb_upload_mask": [vp, vp],
        "clb_download_mask": [vp, vp],
        "clb_set_dynamics": [vp, vp, vp],
        "clb_set_solute_dynamics": [vp, vp],
        "clb_flow_init": [vp, c_double, c_double],
        "clb_flow_steps": [vp, c_int, vp],
        "clb_flow_tracer": [vp, c_double, c_double, c_double],
        "clb_flow_resolve": [vp, c_int, vp, vp, vp],
        "clb_flow_download_velocity": [vp, vp, vp],
        "clb_set_solutes": [vp, c_int, vp, vp, vp],
        "clb_solutes_init": [vp, vp],
        "clb_couple_velocity": [vp],
        "clb_set_velocity": [vp, vp],
        "clb_set_biomass": [vp, c_int, vp],
        "clb_download_conc": [vp, vp],
        "clb_download_biomass": [vp, vp],
        "clb_set_kinetics": [vp, c_int, c_double, c_int, c_int],
        "clb_set_tableau": [vp, c_int, c_int, vp, vp, vp, c_double, c_int],
        "clb_set_mask_update": [vp, c_int, c_double, c_int, c_int],
        "clb_set_mask_codes": [vp, c_int, vp, c_int, vp],
        "clb_step": [vp, c_int],
        "clb_speciation_stats": [vp, vp],
        "clb_batch_load": [vp, c_long, vp, c_int, vp, vp],
        "clb_batch_solve": [vp, c_int, vp],
        "clb_set_batch_method": [vp, c_int],
        "clb_batch_fetch": [vp, vp, vp, vp, vp],
        "clb_mlp_config": [vp, c_int, c_int, c_int],
        "clb_mlp_set_params": [vp, vp],
        "clb_mlp_get_params": [vp, vp],
        "clb_mlp_set_scaling": [vp, vp, vp, vp, vp],
        "clb_mlp_forward": [vp, c_long, vp, vp, c_int],
        "clb_mlp_grad": [vp, c_long, vp, vp, vp, vp],
        "clb_mlp_train": [vp, c_long, vp, vp, c_int, c_int, c_double, c_u64, vp],
        "clb_mlp_train_from_batch": [vp, c_int, c_int, c_double, c_u64, vp],
        "clb_set_immobile": [vp, vp],
        "clb_set_abiotic": [vp, c_int],
        "clb_set_phase": [vp, vp],
        "clb_download_phase": [vp, vp],
        "clb_node_conversion": [vp, c_int, c_int, c_int, c_int, c_double, c_double, vp],
        "clb_set_speciation_mode": [vp, c_int],
        "clb_set_fused_budget": [vp, c_int],
        "clb_set_straggler_budget": [vp, c_int],
        "clb_set_layout": [vp, c_int],
        "clb_set_compact_order": [vp, c_int, c_int, c_int],
        "clb_set_block_shape": [vp, c_int, c_int, c_int],
        "clb_autotune": [vp, c_int, vp],
        "clb_ca_init": [vp],
        "clb_save_state": [vp, ctypes.c_char_p],
        "clb_load_state": [vp, ctypes.c_char_p],
        "clb_set_dissolution": [vp, c_int, c_int, vp, vp, c_int],
        "clb_set_solute_field": [vp, c_int, vp],
        "clb_set_surrogate": [vp, c_int, vp, vp, vp, vp, vp, vp, vp, vp, c_int],
        "clb_set_surrogate_net": [vp, c_int, c_int, vp, c_int, vp, vp, vp, vp, c_double, c_double, c_double,
                                  c_double, c_int, c_int, vp],
        "clb_ca_redistribute": [vp, c_double, c_int, ctypes.c_ulonglong, vp],
        "clb_ca_update_age": [vp, c_double, c_int, vp],
        "clb_fd_biomass": [vp, c_double, vp],
        "clb_ca_download_age": [vp, vp],
        "clb_ca_set_age": [vp, vp],
        "clb_flow_stats": [vp, vp],
        "clb_solutes_reset": [vp, vp],
        "clb_set_graph": [vp, c_int, c_int],
        "clb_set_timing": [vp, c_int],
        "clb_get_perf": [vp, P(Perf)],
        "clb_reset_perf": [vp],
        "clb_gen_sphere_pack": [c_int, c_int, c_int, c_double, c_double, c_double, c_u64, c_int, c_int, c_int, vp],
        "clb_gen_tableau": [c_int, c_int, c_double, c_double, c_int, c_int, c_u64, c_double, c_double, vp, vp],
        "clb_gen_totals_field": [c_int, c_int, c_int, c_int, c_double, c_double, c_double, c_u64, vp],
        "clb_gen_biomass_seed": [c_long, vp, c_int, c_double, c_double, c_double, c_u64, c_u64, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    lib.clb_last_error.argtypes = [vp]
    lib.clb_last_error.restype = ctypes.c_char_p
    _libs[path] = lib
    return lib


def exported_symbols():
    """Names the header declares, for the ABI test."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "clb.h")
    with open(hdr) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(clb_[a-z0-9_]+)\s*\(", text)))


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ generators (host)
# ------------------------------------------------------------------ files (host only)
GEOM_X_FASTEST = "x_fastest"  # one integer per line, x fastest, then y, then z (makeGeometry.py)
GEOM_X_SLICES = "x_slices"    # x slowest, each x slice z fastest (readGeometry's slice reads)


def read_geometry(path, nx, ny, nz, order=GEOM_X_FASTEST):
    """readGeometry (src/complab_functions.hh:127-167): the nx*ny*nz material numbers of a
    geometry file, returned as the (nx + 2)*ny*nz mask of the lattice (x fastest), whose
    planes x = 0 and x = nx + 1 copy x = 1 and x = nx, as the reference's reader does.

    The reference's reader takes the file slice by slice along x (each slice in Palabos'
    field order, z fastest: GEOM_X_SLICES), while its makeGeometry.py writes x fastest
    (GEOM_X_FASTEST); `order` says which the file follows."""
    v = np.loadtxt(path, dtype=np.int64, ndmin=1).ravel()
    if v.size != nx * ny * nz:
        raise ClbError(f"{path}: {v.size} values, expected {nx}*{ny}*{nz}")
    if order == GEOM_X_FASTEST:
        g = v.reshape(nz, ny, nx)
    elif order == GEOM_X_SLICES:
        g = v.reshape(nx, ny, nz).transpose(2, 1, 0)
    else:
        raise ClbError(f"unknown geometry order {order!r}")
    if g.min() < 0 or g.max() > 255:
        raise ClbError(f"{path}: material numbers must be 0..255")
    out = np.empty((nz, ny, nx + 2), dtype=np.uint8)
    out[:, :, 1:-1] = g
    out[:, :, 0] = g[:, :, 0]
    out[:, :, -1] = g[:, :, -1]
    return out.ravel()


def write_vti(path, nx, ny, nz, fields, crop_x=True):
    """VTK XML ImageData of lattice fields (each (N,) or (N, 3), x fastest), base64 binary,
    over the planes 1 <= x <= nx-2 when crop_x (the reference's writers, writeAdvVTI /
    writeNsVTI / saveGeometry, write Box3D(1, nx-2, 0, ny-1, 0, nz-1)). float32 / float64 /
    integer arrays keep their type."""
    import base64
    x0, x1 = (1, nx - 2) if crop_x else (0, nx - 1)
    types = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32",
             np.dtype(np.uint8): "UInt8", np.dtype(np.int64): "Int64"}
    parts = []
    for name, a in fields.items():
        a = np.asarray(a)
        comps = 1 if a.ndim == 1 else a.shape[1]
        a = a.reshape(nz, ny, nx, comps)[:, :, x0:x1 + 1, :]
        if a.dtype not in types:
            a = a.astype(np.float64)
        raw = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes()
        blob = base64.b64encode(np.uint32(len(raw)).astype("<u4").tobytes() + raw).decode("ascii")
        parts.append(f'        <DataArray type="{types[a.dtype]}" Name="{name}" NumberOfComponents="{comps}" '
                     f'format="binary">\n{blob}\n        </DataArray>\n')
    ext = f"0 {x1 - x0} 0 {ny - 1} 0 {nz - 1}"
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" header_type="UInt32">\n')
        f.write(f'  <ImageData WholeExtent="{ext}" Origin="0 0 0" Spacing="1 1 1">\n')
        f.write(f'    <Piece Extent="{ext}">\n      <PointData>\n')
        f.writelines(parts)
        f.write("      </PointData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n")


def _strip_comments(text):
    import re
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _braced(text, start):
    """The text between the brace at `start` and its match."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise ClbError("unbalanced braces in surrogate header")


def _parse_network(text):
    import re
    num = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"

    def values(name):
        m = re.search(r"\b" + name + r"\s*(?:\[[^\]]*\]\s*)*=\s*(\{|" + num + r")", text)
        if not m:
            return None
        if m.group(1) != "{":
            return [[float(m.group(1))]]
        body = _braced(text, m.start(1))
        rows = re.findall(r"\{([^{}]*)\}", body)
        if not rows:
            rows = [body]
        return [[float(v) for v in re.findall(num, r)] for r in rows]

    def scalar(name, default=None):
        v = values(name)
        if v is None:
            if default is None:
                raise ClbError(f"surrogate header: {name} not found")
            return default
        return v[0][0]

    layers = []
    W = values("IW1_1")
    if W is None:
        raise ClbError("surrogate header: no IW1_1 (input weights) found")
    k = 1
    while W is not None:
        b = values(f"b{k}")
        if b is None:
            raise ClbError(f"surrogate header: b{k} not found")
        b = [v for row in b for v in row]
        if len(W) != len(b) or len({len(r) for r in W}) != 1:
            raise ClbError(f"surrogate header: layer {k} shapes do not match")
        layers.append((np.array(W, dtype=np.float64), np.array(b, dtype=np.float64)))
        W = values(f"LW{k + 1}_{k}")
        k += 1
    n_in = layers[0][0].shape[1]
    for (Wa, _), (Wb, _) in zip(layers, layers[1:]):
        if Wb.shape[1] != Wa.shape[0]:
            raise ClbError("surrogate header: consecutive layers do not chain")
    if layers[-1][0].shape[0] != 1:
        raise ClbError("surrogate header: the output layer must have one neuron")
    xo = values("x_offset")
    if xo is not None:  # the trainer's generated header: arrays, inputs in substrate order
        x_offset = [v for r in xo for v in r]
        x_gain = [v for r in values("x_gain") for v in r]
        in_sub = list(range(n_in))
    else:  # the model file: one scalar per input, input k taken from Fin[iM][j]
        x_offset = [scalar(f"x_offset{j}") for j in range(n_in)]
        x_gain = [scalar(f"x_gain{j}") for j in range(n_in)]
        in_sub = list(range(n_in))
        for kk, j in re.findall(r"Xp0\s*\[\s*(\d+)\s*\]\s*=\s*\(\s*Fin\s*\[[^\]]*\]\s*\[\s*(\d+)\s*\]", text):
            in_sub[int(kk)] = int(j)
    m = re.search(r"\blogOutput\s*=\s*(true|false)", text)
    return {
        "n_in": n_in,
        "in_sub": in_sub,
        "widths": [Wl.shape[0] for Wl, _ in layers],
        "params": np.concatenate([np.concatenate([Wl.ravel(), bl]) for Wl, bl in layers]),
        "x_offset": x_offset,
        "x_gain": x_gain,
        "x_ymin": scalar("x_ymin", -1.0),
        "y_ymin": scalar("y_ymin", -1.0),
        "y_gain": scalar("y_gain"),
        "y_offset": scalar("y_offset", 0.0),
        "log_output": bool(m and m.group(1) == "true"),
    }


def read_surrogate_header(path):
    """The networks of a surrogate model header, as {microbe id: network} for
    Lattice.set_surrogate_net. Reads both forms the reference produces: the model file
    (surrogateModel.hh:252-500: std::vector weight tables IW1_1, LW2_1, ..., b1, ...,
    scalar x_offsetK / x_gainK, one block per `microbeId == k` branch, inputs named by
    its Fin[iM][j] lines) and the trainer's generated header
    (surrogate_training/trainSurrogate.py:245-395: array weights, x_offset[] / x_gain[],
    logOutput, inputs in substrate order, one network, returned as id 0)."""
    import re
    with open(path) as fh:
        text = _strip_comments(fh.read())
    marks = [(m.start(), int(m.group(1))) for m in re.finditer(r"microbeId\s*==\s*(\d+)", text)]
    if not marks:
        return {0: _parse_network(text)}
    nets = {}
    for i, (pos, mid) in enumerate(marks):
        end = marks[i + 1][0] if i + 1 < len(marks) else len(text)
        if re.search(r"\bIW1_1\s*(?:\[[^\]]*\]\s*)*=", text[pos:end]):
            nets[mid] = _parse_network(text[pos:end])
    return nets


def gen_sphere_pack(nx, ny, nz, rmin=3.0, rmax=5.0, porosity=0.45, seed=0, pore=2, solid=0, bb=1):
    lib = load()
    m = np.zeros(nx * ny * nz, dtype=np.uint8)
    rc = lib.clb_gen_sphere_pack(nx, ny, nz, rmin, rmax, porosity, seed, pore, solid, bb, _ptr(m))
    if rc != CLB_OK:
        raise ClbError(f"clb_gen_sphere_pack failed ({rc})")
    return m


def gen_tableau(nc, ns, logk_min=-14.0, logk_max=14.0, smin=-2, smax=2, seed=3, log_t_min=-6.0, log_t_max=-2.0):
    lib = load()
    S = np.zeros(ns * nc)
    lk = np.zeros(ns)
    rc = lib.clb_gen_tableau(nc, ns, logk_min, logk_max, smin, smax, seed, log_t_min, log_t_max, _ptr(S), _ptr(lk))
    if rc != CLB_OK:
        raise ClbError(f"clb_gen_tableau failed ({rc})")
    return S.reshape(ns, nc), lk


def gen_totals_field(nx, ny, nz, nc, sigma=4.0, log_t_min=-6.0, log_t_max=-2.0, seed=4):
    lib = load()
    t = np.zeros(nx * ny * nz * nc)
    rc = lib.clb_gen_totals_field(nx, ny, nz, nc, sigma, log_t_min, log_t_max, seed, _ptr(t))
    if rc != CLB_OK:
        raise ClbError(f"clb_gen_totals_field failed ({rc})")
    return t.reshape(-1, nc)


def gen_biomass_seed(mask, pore=2, fraction=0.01, bmin=10.0, bmax=100.0, seed_sites=1, seed_density=2):
    lib = load()
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    b = np.zeros(m.size)
    rc = lib.clb_gen_biomass_seed(m.size, _ptr(m), pore, fraction, bmin, bmax, seed_sites, seed_density, _ptr(b))
    if rc != CLB_OK:
        raise ClbError(f"clb_gen_biomass_seed failed ({rc})")
    return b


def device_name():
    lib = load()
    buf = ctypes.create_string_buffer(256)
    if lib.clb_device_name(buf, 256) != CLB_OK:
        raise ClbError("no HIP device")
    return buf.value.decode()


# ------------------------------------------------------------------ flow convergence (host)
class ValueTracer:
    """Palabos' util::ValueTracer (Palabos v2.3.0, src/core/util.h; not in the reference tree),
    restated from its published behaviour, as the reference's flow loops use it
    (src/complab.cpp:593, 609-610, 1149, 1797-1798): ValueTracer(u, L, epsilon) keeps the last
    deltaT = int(L / u / 2) values; it has converged once that window is full and
    |stddev / mean| < epsilon, stddev with n - 1; a NaN ratio counts as converged (Palabos
    reports "simulation diverged" and stops). The driver builds its tracers with u = 1,
    L = 1000: a window of 500 values."""

    def __init__(self, u=1.0, L=1000.0, epsilon=1e-4):
        self.delta_t = int(L / u / 2.0)
        self.epsilon = epsilon
        self.values = []
        self.t = 0

    def take_value(self, v):
        self.values.append(v)
        if len(self.values) > abs(self.delta_t):
            self.values.pop(0)
        self.t += 1

    def reset(self):
        self.values = []
        self.t = 0

    def average(self):
        return sum(self.values) / len(self.values)

    def std_dev(self, avg):
        n = len(self.values)
        return math.sqrt(sum((v - avg) * (v - avg) for v in self.values) / (n - 1))

    def has_converged(self):
        if len(self.values) < abs(self.delta_t):
            return False
        avg = self.average()
        sd = self.std_dev(avg)
        # IEEE division, as in C++: 0/0 is NaN, x/0 infinite
        ratio = sd / avg if avg != 0.0 else (float("nan") if sd == 0.0 else math.copysign(float("inf"), sd))
        if math.isnan(ratio):
            return True
        return abs(ratio) < self.epsilon


# ------------------------------------------------------------------ context
class Lattice:
    """One GPU's lattice: mask, D3Q19 flow, D3Q7 solutes, reactions, batched speciation."""

    def __init__(self, nx, ny, nz, precision=PREC_FP64, device=0, library=None):
        """library: the path of another build of the library (other kinetics plugins), or None
        for libclb.so."""
        self.lib = load(library)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny * nz
        cfg = _Config(device, precision, nx, ny, nz)
        h = ctypes.c_void_p()
        rc = self.lib.clb_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != CLB_OK:
            raise ClbError(f"clb_create failed ({rc}): is a GPU visible and the library built for it?")
        self.h = h
        self.S = 0
        self.nb = 0
        self.nc = 0

    def close(self):
        if self.h:
            self.lib.clb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != CLB_OK:
            msg = self.lib.clb_last_error(self.h)
            raise ClbError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # geometry / dynamics
    def upload_mask(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        assert m.size == self.N
        self._chk(self.lib.clb_upload_mask(self.h, _ptr(m)), "upload_mask")

    def download_mask(self):
        m = np.zeros(self.N, dtype=np.uint8)
        self._chk(self.lib.clb_download_mask(self.h, _ptr(m)), "download_mask")
        return m

    def set_dynamics(self, kind, flow_omega):
        k = np.ascontiguousarray(kind, dtype=np.int32)
        o = _f64(flow_omega)
        assert k.size == 256 and o.size == 256
        self._chk(self.lib.clb_set_dynamics(self.h, _ptr(k), _ptr(o)), "set_dynamics")

    def set_solute_dynamics(self, kind):
        """The solute lattices' own per-code dynamics (soluteDomainSetup): a biofilm code
        stays advection-diffusion here even when it is bounce-back for the flow."""
        k = np.ascontiguousarray(kind, dtype=np.int32)
        assert k.size == 256
        self._chk(self.lib.clb_set_solute_dynamics(self.h, _ptr(k)), "set_solute_dynamics")

    # flow
    def flow_init(self, rho_in, rho_out):
        self._chk(self.lib.clb_flow_init(self.h, rho_in, rho_out), "flow_init")

    def flow_steps(self, n, energy=False):
        e = np.zeros(n) if energy else None
        self._chk(self.lib.clb_flow_steps(self.h, n, _ptr(e)), "flow_steps")
        return e

    def velocity(self):
        u = np.zeros(3 * self.N)
        r = np.zeros(self.N)
        self._chk(self.lib.clb_flow_download_velocity(self.h, _ptr(u), _ptr(r)), "velocity")
        return u.reshape(self.N, 3), r

    # solutes
    def set_solutes(self, omega, bc_type, bc_value):
        o = _f64(omega).ravel()
        S = o.size // 256
        bt = np.ascontiguousarray(bc_type, dtype=np.int32).ravel()
        bv = _f64(bc_value).ravel()
        assert bt.size == 2 * S and bv.size == 2 * S
        self._chk(self.lib.clb_set_solutes(self.h, S, _ptr(o), _ptr(bt), _ptr(bv)), "set_solutes")
        self.S = S

    def solutes_init(self, c0):
        c = _f64(c0)
        self._chk(self.lib.clb_solutes_init(self.h, _ptr(c)), "solutes_init")

    def couple_velocity(self):
        self._chk(self.lib.clb_couple_velocity(self.h), "couple_velocity")

    def set_velocity(self, u):
        uu = _f64(u).ravel()
        assert uu.size == 3 * self.N
        self._chk(self.lib.clb_set_velocity(self.h, _ptr(uu)), "set_velocity")

    def set_biomass(self, b):
        bb = _f64(b).ravel()
        nb = bb.size // self.N
        self._chk(self.lib.clb_set_biomass(self.h, nb, _ptr(bb)), "set_biomass")
        self.nb = nb

    def conc(self):
        c = np.zeros(self.S * self.N)
        self._chk(self.lib.clb_download_conc(self.h, _ptr(c)), "download_conc")
        return c.reshape(self.S, self.N)

    def biomass(self):
        b = np.zeros(max(self.nb, 1) * self.N)
        self._chk(self.lib.clb_download_biomass(self.h, _ptr(b)), "download_biomass")
        return b.reshape(-1, self.N)[: self.nb]

    def set_kinetics(self, kind, dt, solid_code=0, bb_code=1):
        self._chk(self.lib.clb_set_kinetics(self.h, kind, dt, solid_code, bb_code), "set_kinetics")

    def set_tableau(self, S, logK, comp_species=None, tol=1e-10, max_iter=200):
        S = _f64(S)
        ns, nc = S.shape
        lk = _f64(logK)
        cs = None if comp_species is None else np.ascontiguousarray(comp_species, dtype=np.int32)
        self._chk(self.lib.clb_set_tableau(self.h, nc, ns, _ptr(S.ravel()), _ptr(lk), _ptr(cs), tol, max_iter),
                  "set_tableau")
        self.nc = nc

    def set_mask_update(self, enable, threshold, pore_code=2, biofilm_code=3):
        self._chk(self.lib.clb_set_mask_update(self.h, int(enable), threshold, pore_code, biofilm_code),
                  "set_mask_update")

    def set_mask_codes(self, pore_codes, bio_codes):
        """The pore <-> biofilm rule with several material types (pore_dynamics, and
        bio_dynamics[i][0] per biomass field i): see clb_set_mask_codes."""
        p = np.ascontiguousarray(pore_codes, dtype=np.int32)
        b = np.ascontiguousarray(bio_codes, dtype=np.int32)
        self._chk(self.lib.clb_set_mask_codes(self.h, p.size, _ptr(p), b.size, _ptr(b)), "set_mask_codes")

    def step(self, n=1):
        self._chk(self.lib.clb_step(self.h, n), "step")

    def set_immobile(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        assert f.size == self.S
        self._chk(self.lib.clb_set_immobile(self.h, _ptr(f)), "set_immobile")

    def set_abiotic(self, mode):
        self._chk(self.lib.clb_set_abiotic(self.h, int(mode)), "set_abiotic")

    def set_phase(self, phase):
        p = np.ascontiguousarray(phase, dtype=np.uint8).ravel()
        assert p.size == self.N
        self._chk(self.lib.clb_set_phase(self.h, _ptr(p)), "set_phase")

    def phase(self):
        p = np.zeros(self.N, dtype=np.uint8)
        self._chk(self.lib.clb_download_phase(self.h, _ptr(p)), "download_phase")
        return p

    def node_conversion(self, mineral_solute, pore_code, precip_code, phase_id, precip_max, reopen):
        ch = ctypes.c_longlong()
        self._chk(self.lib.clb_node_conversion(self.h, mineral_solute, pore_code, precip_code, phase_id, precip_max,
                                               reopen, ctypes.byref(ch)), "node_conversion")
        return ch.value

    def flow_stats(self):
        s = np.zeros(4)
        self._chk(self.lib.clb_flow_stats(self.h, _ptr(s)), "flow_stats")
        return s

    def flow_resolve(self, tracer, max_steps):
        """A flow solve of the reference's loops (src/complab.cpp:607-620, 1794-1800,
        1836-1840): collideAndStream, the stored average energy into the tracer, stop when it
        has converged, at most max_steps. The tracer keeps its window across calls, as the
        driver's ns_convg2 does (declared once, src/complab.cpp:1149, never reset). Steps that
        cannot end the solve (the window not yet full) run as one batch. Returns (steps run,
        converged)."""
        done = 0
        while done < max_steps:
            k = min(max(1, tracer.delta_t - len(tracer.values)), max_steps - done)
            for v in self.flow_steps(k, energy=True):
                tracer.take_value(float(v))
            done += k
            if tracer.has_converged():
                return done, True
        return done, False

    def flow_tracer(self, u=1.0, L=1000.0, epsilon=1e-4):
        """Start the device tracer (Palabos' ValueTracer(u, L, epsilon), as ValueTracer above)
        with an empty window; clb_flow_resolve's steps feed it until the next call."""
        self._chk(self.lib.clb_flow_tracer(self.h, float(u), float(L), float(epsilon)), "flow_tracer")

    def flow_resolve_device(self, max_steps, energy=False):
        """flow_resolve with the tracer on the device (flow_tracer): the same steps and
        convergence as the host loop with the same tracer, without a host round trip per step.
        Returns (steps run, converged) or, with energy, (steps run, converged, energies)."""
        steps, conv = ctypes.c_int(), ctypes.c_int()
        e = np.zeros(max(int(max_steps), 1)) if energy else None
        self._chk(self.lib.clb_flow_resolve(self.h, int(max_steps), ctypes.byref(steps), ctypes.byref(conv),
                                            _ptr(e) if energy else None), "flow_resolve")
        if energy:
            return steps.value, bool(conv.value), e[:steps.value].copy()
        return steps.value, bool(conv.value)

    def percolates(self, thrd=1e-12):
        """The percolation test after a re-solve that used all its steps (src/complab.cpp:
        1801-1805): the mean x velocity over the plane x = nx-2 above thrd."""
        return float(self.flow_stats()[2]) > thrd

    def solutes_reset(self, c0):
        c = _f64(c0)
        self._chk(self.lib.clb_solutes_reset(self.h, _ptr(c)), "solutes_reset")

    def set_graph(self, enable, steps_per_graph=100):
        self._chk(self.lib.clb_set_graph(self.h, int(enable), steps_per_graph), "set_graph")

    def set_speciation_mode(self, mode):
        self._chk(self.lib.clb_set_speciation_mode(self.h, int(mode)), "set_speciation_mode")

    def set_layout(self, layout):
        self._chk(self.lib.clb_set_layout(self.h, layout), "set_layout")

    def set_compact_order(self, bx, by, bz):
        self._chk(self.lib.clb_set_compact_order(self.h, bx, by, bz), "set_compact_order")

    def set_fused_budget(self, evals):
        self._chk(self.lib.clb_set_fused_budget(self.h, int(evals)), "set_fused_budget")

    def set_straggler_budget(self, evals):
        """Budget of the stragglers' first pass (0: the full iteration takes every straggler)."""
        self._chk(self.lib.clb_set_straggler_budget(self.h, int(evals)), "set_straggler_budget")

    def set_block_shape(self, kernel, bx, by):
        self._chk(self.lib.clb_set_block_shape(self.h, int(kernel), int(bx), int(by)), "set_block_shape")

    # checkpoint
    def save_state(self, path):
        self._chk(self.lib.clb_save_state(self.h, os.fsencode(path)), "save_state")

    def load_state(self, path):
        self._chk(self.lib.clb_load_state(self.h, os.fsencode(path)), "load_state")

    # mineral dissolution
    def set_dissolution(self, enable, phase_ids=(), substrates=(), surface_only=True):
        ids = np.ascontiguousarray(phase_ids, dtype=np.int32)
        sub = np.ascontiguousarray(substrates, dtype=np.int32)
        self._chk(self.lib.clb_set_dissolution(self.h, int(enable), ids.size, _ptr(ids), _ptr(sub), int(surface_only)),
                  "set_dissolution")

    def set_surrogate(self, ids, Kc, decay, vmax=None, max_uptake=None, molar_mass=None, fix_conc=None,
                      fix_lb=None, use_totals=False):
        """Microbes grown by the surrogate processor (clb_set_surrogate); ids = their biomass
        fields; Kc, vmax, max_uptake: (n, S); decay, molar_mass: (n,); fix_*: (S,)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32).ravel()
        n = ids.size
        S = self.S

        def mat(a):
            return None if a is None else _f64(np.reshape(a, (n, S)))

        def flags(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.int32).ravel()
        kc, vm, mu = mat(Kc), mat(vmax), mat(max_uptake)
        dec = _f64(np.reshape(decay, (n,))) if n else None
        mw = None if molar_mass is None else _f64(np.reshape(molar_mass, (n,)))
        fc, fl = flags(fix_conc), flags(fix_lb)
        self._chk(self.lib.clb_set_surrogate(self.h, n, _ptr(ids), _ptr(kc), _ptr(dec), _ptr(vm), _ptr(mu), _ptr(mw),
                                             _ptr(fc), _ptr(fl), int(bool(use_totals))), "set_surrogate")

    def set_surrogate_net(self, microbe, net, closure=0, yields=None):
        """Microbe `microbe`'s network (a dict as read_surrogate_header returns), or None to
        remove it; closure 0 / 1 / 2 with yields (S,) in gDW per mmol."""
        if net is None:
            self._chk(self.lib.clb_set_surrogate_net(self.h, int(microbe), 0, None, 0, None, None, None, None, 0.0,
                                                     0.0, 1.0, 0.0, 0, 0, None), "set_surrogate_net")
            return
        ins = np.ascontiguousarray(net["in_sub"], dtype=np.int32)
        wid = np.ascontiguousarray(net["widths"], dtype=np.int32)
        prm = _f64(net["params"]).ravel()
        xo, xg = _f64(net["x_offset"]).ravel(), _f64(net["x_gain"]).ravel()
        y = None if yields is None else _f64(yields).ravel()
        self._chk(self.lib.clb_set_surrogate_net(self.h, int(microbe), int(net["n_in"]), _ptr(ins), wid.size, _ptr(wid),
                                                 _ptr(prm), _ptr(xo), _ptr(xg), float(net["x_ymin"]),
                                                 float(net["y_ymin"]), float(net["y_gain"]), float(net["y_offset"]),
                                                 int(bool(net.get("log_output"))), int(closure), _ptr(y)),
                  "set_surrogate_net")

    def set_solute_field(self, solute, c):
        cc = _f64(c).ravel()
        assert cc.size == self.N
        self._chk(self.lib.clb_set_solute_field(self.h, int(solute), _ptr(cc)), "set_solute_field")

    # CA biofilm: biomass redistribution
    def ca_init(self):
        self._chk(self.lib.clb_ca_init(self.h), "ca_init")

    def ca_redistribute(self, bmax, half=0, seed=1):
        """Sweep until the max total biomass is within 1e-12 of bmax; returns the report as a dict."""
        r = CaReport()
        self._chk(self.lib.clb_ca_redistribute(self.h, float(bmax), int(half), int(seed), ctypes.addressof(r)),
                  "ca_redistribute")
        return {k: getattr(r, k) for k, _ in CaReport._fields_}

    def fd_biomass(self, nu):
        """One step of the FD biofilm solver (mask/totals, explicit diffusion, mask/totals);
        returns the report as a dict."""
        r = CaReport()
        self._chk(self.lib.clb_fd_biomass(self.h, float(nu), ctypes.addressof(r)), "fd_biomass")
        return {k: getattr(r, k) for k, _ in CaReport._fields_}

    def ca_update_age(self, bmax, max_passes=1000):
        p = ctypes.c_int(0)
        self._chk(self.lib.clb_ca_update_age(self.h, float(bmax), int(max_passes), ctypes.addressof(p)),
                  "ca_update_age")
        return p.value

    def ca_age(self):
        a = np.zeros(self.N, dtype=np.int32)
        self._chk(self.lib.clb_ca_download_age(self.h, _ptr(a)), "ca_download_age")
        return a

    def ca_set_age(self, age):
        a = np.ascontiguousarray(age, dtype=np.int32).ravel()
        assert a.size == self.N
        self._chk(self.lib.clb_ca_set_age(self.h, _ptr(a)), "ca_set_age")

    def autotune(self, steps=10):
        """Pick the fastest block shapes (and fused budget) on the current state; the state
        is left unchanged. Returns the report as a dict."""
        r = TuneReport()
        self._chk(self.lib.clb_autotune(self.h, int(steps), ctypes.addressof(r)), "autotune")
        return {k: getattr(r, k) for k, _ in TuneReport._fields_}

    def speciation_stats(self):
        s = np.zeros(3, dtype=np.int64)
        self._chk(self.lib.clb_speciation_stats(self.h, _ptr(s)), "speciation_stats")
        return s

    # batched speciation
    def batch_load(self, totals, var_species=None, logk_var=None):
        t = _f64(totals)
        n = t.shape[0]
        nvar = 0 if var_species is None else len(var_species)
        vs = None if nvar == 0 else np.ascontiguousarray(var_species, dtype=np.int32)
        lk = None if nvar == 0 else _f64(logk_var)
        self._chk(self.lib.clb_batch_load(self.h, n, _ptr(t.ravel()), nvar, _ptr(vs),
                                          None if lk is None else _ptr(lk.ravel())), "batch_load")
        self.bn = n

    def batch_solve(self, warm=WARM_REF, logc_in=None):
        lc = None if logc_in is None else _f64(logc_in).ravel()
        self._chk(self.lib.clb_batch_solve(self.h, warm, _ptr(lc)), "batch_solve")

    def set_batch_method(self, method):
        """BATCH_AUTO (the polish from caller / network starts), BATCH_FULL or BATCH_POLISH."""
        self._chk(self.lib.clb_set_batch_method(self.h, int(method)), "set_batch_method")

    def batch_fetch(self):
        n, nc = self.bn, self.nc
        lc = np.zeros(n * nc)
        it = np.zeros(n, dtype=np.int32)
        r = np.zeros(n)
        cv = np.zeros(n, dtype=np.uint8)
        self._chk(self.lib.clb_batch_fetch(self.h, _ptr(lc), _ptr(it), _ptr(r), _ptr(cv)), "batch_fetch")
        return lc.reshape(n, nc), it, r, cv

    # MLP
    def mlp_config(self, din, hidden, dout):
        self._chk(self.lib.clb_mlp_config(self.h, din, hidden, dout), "mlp_config")
        self.mlp_dims = (din, hidden, dout)

    def mlp_nparams(self):
        din, H, dout = self.mlp_dims
        return H * din + H + H * H + H + dout * H + dout

    def mlp_set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == self.mlp_nparams()
        self._chk(self.lib.clb_mlp_set_params(self.h, _ptr(t)), "mlp_set_params")

    def mlp_get_params(self):
        t = np.zeros(self.mlp_nparams(), dtype=np.float32)
        self._chk(self.lib.clb_mlp_get_params(self.h, _ptr(t)), "mlp_get_params")
        return t

    def mlp_set_scaling(self, xmin, xmax, ymin, ymax):
        a = [_f64(v) for v in (xmin, xmax, ymin, ymax)]
        self._chk(self.lib.clb_mlp_set_scaling(self.h, *[_ptr(v) for v in a]), "mlp_set_scaling")

    def mlp_forward(self, X, bf16=False):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        Y = np.zeros((n, self.mlp_dims[2]), dtype=np.float32)
        self._chk(self.lib.clb_mlp_forward(self.h, n, _ptr(X), _ptr(Y), int(bf16)), "mlp_forward")
        return Y

    def mlp_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.mlp_nparams(), dtype=np.float32)
        loss = ctypes.c_double()
        self._chk(self.lib.clb_mlp_grad(self.h, X.shape[0], _ptr(X), _ptr(T), _ptr(g), ctypes.byref(loss)),
                  "mlp_grad")
        return g, loss.value

    def mlp_train(self, X, T, epochs, batch, lr, seed=7):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = ctypes.c_double()
        self._chk(self.lib.clb_mlp_train(self.h, X.shape[0], _ptr(X), _ptr(T), epochs, batch, lr, seed,
                                         ctypes.byref(loss)), "mlp_train")
        return loss.value

    def mlp_train_from_batch(self, epochs, batch, lr, seed=7):
        loss = ctypes.c_double()
        self._chk(self.lib.clb_mlp_train_from_batch(self.h, epochs, batch, lr, seed, ctypes.byref(loss)),
                  "mlp_train_from_batch")
        return loss.value

    # measurement
    def set_timing(self, on):
        self._chk(self.lib.clb_set_timing(self.h, int(on)), "set_timing")

    def perf(self):
        p = Perf()
        self._chk(self.lib.clb_get_perf(self.h, ctypes.byref(p)), "get_perf")
        return p

    def reset_perf(self):
        self._chk(self.lib.clb_reset_perf(self.h), "reset_perf")

    def sync(self):
        self._chk(self.lib.clb_sync(self.h), "sync")


class EquilibriumChemistry:
    """GPU-backed mirror of EquilibriumChemistry<T>
    (src/complab3d_processors_part4_eqsolver.hh:186-1271): same setters, same defaults
    (maxIterations 200, tolerance 1e-8), same convergence judge (PCF residual 2-norm <
    tolerance); calculate_species_concentrations takes a batch of species vectors."""

    def __init__(self, lattice=None):
        self.lat = lattice if lattice is not None else Lattice(4, 1, 1)
        self.species = []
        self.components = []
        self.logK = None
        self.S = None
        self.maxIterations = 200
        self.tolerance = 1e-8
        self.lastIterations = None
        self.lastResidual = None
        self.lastConverged = None

    def setSpeciesNames(self, names):
        self.species = [n for n in names if n not in ("H2O", "h2o")]

    def setComponentNames(self, names):
        self.components = list(names)

    def setLogK(self, logK):
        self.logK = np.asarray(logK, dtype=np.float64)

    def setStoichiometryMatrix(self, S):
        self.S = np.asarray(S, dtype=np.float64)

    def setMaxIterations(self, n):
        self.maxIterations = int(n)

    def setTolerance(self, tol):
        self.tolerance = float(tol)

    def _comp_species(self):
        return [self.species.index(c) if c in self.species else -1 for c in self.components]

    def calculate_species_concentrations(self, conc):
        """conc: (n, ns) species concentrations -> (n, ns) equilibrium concentrations."""
        conc = np.atleast_2d(np.asarray(conc, dtype=np.float64))
        S, lk = self.S, self.logK
        ns, nc = S.shape
        eq = np.array([np.any(np.abs(S[i]) > 1e-10) for i in range(ns)])
        cs = np.clip(conc, 1e-30, 10.0)
        T = np.maximum(cs[:, eq] @ S[eq], 1e-30)
        guess = np.empty((conc.shape[0], nc))
        for j, sp in enumerate(self._comp_species()):
            guess[:, j] = np.log10(cs[:, sp]) if sp >= 0 else -7.0
        self.lat.set_tableau(S, lk, self._comp_species(), self.tolerance, self.maxIterations)
        self.lat.batch_load(T)
        self.lat.batch_solve(WARM_USER, guess)
        x, it, r, cv = self.lat.batch_fetch()
        self.lastIterations, self.lastResidual, self.lastConverged = it, r, cv.astype(bool)
        lc = np.clip(lk[None, :] + x @ S.T, -30.0, 1.0)
        out = np.where(eq[None, :], 10.0 ** lc, np.where(conc > 1e-30, np.minimum(conc, 10.0), 1e-30))
        return np.clip(out, 1e-30, 10.0)
