"""ctypes mirror of include/incompflow.h (the C-ABI of the solver library).

Only data layouts and constants live here; the library binding is in ``_lib.py``.
Field-for-field correspondence with the reference's compile-time contract:
``ifl_config`` <-> include/config.h:6-81, face/slot/BC codes <-> include/constants.h:7-36.
"""
import ctypes as C

IFL_ABI_VERSION = 3

# status codes
IFL_OK, IFL_ERR_ARG, IFL_ERR_DEVICE, IFL_ERR_DIVERGED, IFL_ERR_NOT_CONVERGED, IFL_ERR_IO, IFL_ERR_NOMEM = range(7)

# faces (constants.h:7-12)
EAST, WEST, NORTH, SOUTH, TOP, BOTTOM = range(6)
# coefficient slots (constants.h:14-20)
SLOT_C, SLOT_E, SLOT_W, SLOT_N, SLOT_S, SLOT_T, SLOT_B = range(7)
# BC codes (constants.h:34-36)
WALL, INLET, OUTLET = range(3)
# temperature BC codes (config.h:46-53)
TEMP_DIRICHLET, TEMP_NEUMANN = range(2)

SMOOTHER_RBGS, SMOOTHER_JACOBI, SMOOTHER_LEX = range(3)
STOP_NONE, STOP_FAITHFUL, STOP_FIXED, STOP_SCALED = range(4)
PSOLVE_SWEEPS, PSOLVE_MG = range(2)

(FIELD_U, FIELD_V, FIELD_W, FIELD_P, FIELD_TEMP, FIELD_PCORR, FIELD_UF, FIELD_VF, FIELD_WF,
 FIELD_MCOEF, FIELD_PCOEF, FIELD_BU, FIELD_BV, FIELD_BW, FIELD_BP, FIELD_TCOEF, FIELD_BT) = range(17)
FIELD_NAMES = ["u", "v", "w", "p", "temp", "pcorr", "uf", "vf", "wf", "mcoef", "pcoef",
               "bu", "bv", "bw", "bp", "tcoef", "bt"]

(STAGE_INIT_FACES, STAGE_FACE_BC, STAGE_MOM_ASSEMBLE, STAGE_MOM_RESIDUAL, STAGE_SMOOTH_U,
 STAGE_SMOOTH_V, STAGE_SMOOTH_W, STAGE_RHIE_CHOW, STAGE_PCORR_ASSEMBLE, STAGE_SMOOTH_P,
 STAGE_UPDATE, STAGE_OUTER, STAGE_SMOOTH_T) = range(13)

TILE_AUTO, TILE_64x16, TILE_64x8, TILE_32x8, TILE_32x4, TILE_128x32 = range(6)
_TILE_SHAPE = {TILE_64x16: (64, 16), TILE_64x8: (64, 8), TILE_32x8: (32, 8), TILE_32x4: (32, 4), TILE_128x32: (128, 32)}


def tile_count(nx, ny, tile):
    tx, ty = _TILE_SHAPE[tile]
    return -(-nx // tx) * -(-ny // ty)


def default_tile(nx, ny):
    """The 2D sweep tile a single case gets with tile = 0 (the library's rule, sweep2d.hip)."""
    if tile_count(nx, ny, TILE_128x32) >= 256:
        return TILE_128x32
    if tile_count(nx, ny, TILE_64x16) >= 768:
        return TILE_64x16
    if tile_count(nx, ny, TILE_64x8) >= 512:
        return TILE_64x8
    if tile_count(nx, ny, TILE_32x8) >= 256:
        return TILE_32x8
    return TILE_32x4

(KERNEL_SWEEP, KERNEL_MOM_ASSEMBLE, KERNEL_RHIE_CHOW, KERNEL_PCORR_ASSEMBLE, KERNEL_UPDATE, KERNEL_MOM_SWEEP,
 KERNEL_TEMP_SWEEP, KERNEL_PSOLVE) = range(8)


class Config(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xmin", C.c_double), ("xmax", C.c_double),
        ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("density", C.c_double),
        ("viscosity", C.c_double),
        ("bc_type", C.c_int32 * 6),
        ("bc_val", (C.c_double * 3) * 6),
        ("num_outer_iter", C.c_int32),
        ("n_iter", C.c_int32 * 4),
        ("relax", C.c_double * 4),
        ("tol", C.c_double * 4),
        ("tol_mass", C.c_double),
        ("init", C.c_double * 5),
        ("smoother", C.c_int32),
        ("stop_rule", C.c_int32),
        ("tol_residual", C.c_double),
        ("p_refine", C.c_int32),
        ("log_every", C.c_int32),
        ("check_every", C.c_int32),
        ("tile", C.c_int32),
        ("sweep_block", C.c_int32),
        ("energy", C.c_int32),
        ("n_iter_t", C.c_int32),
        ("conductivity", C.c_double),
        ("specific_heat", C.c_double),
        ("relax_t", C.c_double),
        ("tol_t", C.c_double),
        ("temp_bc_type", C.c_int32 * 6),
        ("temp_bc_val", C.c_double * 6),
        ("sweep_block_mom", C.c_int32),
        ("p_solver", C.c_int32),
        ("mg_cycles", C.c_int32),
        ("mg_sweeps", C.c_int32),
        ("mg_coarse_sweeps", C.c_int32),
        ("mg_over", C.c_double),
        ("fail_if_not_converged", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]

    def copy(self):
        c = Config()
        C.pointer(c)[0] = self
        return c

    def field_size(self, field):
        nx, ny, nz, dim = self.nx, self.ny, self.nz, self.dim
        n = nx * ny * nz
        if field in (FIELD_U, FIELD_V, FIELD_W, FIELD_P, FIELD_TEMP, FIELD_PCORR,
                     FIELD_BU, FIELD_BV, FIELD_BW, FIELD_BP, FIELD_BT):
            return n
        if field == FIELD_UF:
            return (nx + 1) * ny * nz
        if field == FIELD_VF:
            return nx * (ny + 1) * nz
        if field == FIELD_WF:
            return nx * ny * (nz + 1)
        if field in (FIELD_MCOEF, FIELD_PCOEF, FIELD_TCOEF):
            return n * (1 + 2 * dim)
        raise ValueError(field)

    def shape(self, field):
        """numpy shape (slowest first) of a field."""
        nx, ny, nz, dim = self.nx, self.ny, self.nz, self.dim
        if field == FIELD_UF:
            return (nz, ny, nx + 1)
        if field == FIELD_VF:
            return (nz, ny + 1, nx)
        if field == FIELD_WF:
            return (nz + 1, ny, nx)
        if field in (FIELD_MCOEF, FIELD_PCOEF, FIELD_TCOEF):
            return (1 + 2 * dim, nz, ny, nx)
        return (nz, ny, nx)


class IterRecord(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "p_norm", "u_norm", "v_norm", "w_norm", "uf_norm", "vf_norm", "wf_norm", "mass_norm",
        "res_cont", "res_u", "res_v", "res_w", "scaled",
        "rel_p", "rel_u", "rel_v", "rel_w", "rel_uf", "rel_vf", "rel_wf")] + [
        ("sweeps", C.c_int32 * 4), ("converged", C.c_int32), ("sweeps_t", C.c_int32), ("res_t", C.c_double)]


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("status", C.c_int32),
                ("pad", C.c_int32), ("elapsed_ms", C.c_double), ("final_scaled", C.c_double),
                ("sweeps_total", C.c_double)]


TUNE_LAUNCH, TUNE_SWEEPS, TUNE_PSOLVE, TUNE_PSOLVE_TARGET = 1, 2, 4, 8
OBJ_PRESSURE_DROP, OBJ_LINEAR = 0, 1


class AdjointOpts(C.Structure):
    _fields_ = [("objective", C.c_int32), ("max_iter", C.c_int32), ("tol", C.c_double), ("grad_tol", C.c_double),
                ("weights", C.POINTER(C.c_float) * 4), ("warm_start", C.c_int32), ("check_every", C.c_int32),
                ("tile", C.c_int32), ("reserved", C.c_int32)]


class AdjointReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("J", C.c_double),
                ("final_scaled", C.c_double), ("elapsed_ms", C.c_double), ("grad_change", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("tile", C.c_int32), ("sweep_block", C.c_int32), ("n_iter", C.c_int32 * 4),
                ("us_per_iter", C.c_double), ("decades_per_s", C.c_double), ("trials", C.c_int32),
                ("sweep_block_mom", C.c_int32), ("p_solver", C.c_int32), ("mg_cycles", C.c_int32), ("mg_over", C.c_double)]


def default_config():
    """The reference's shipped include/config.h (1024x256 channel on [-2,2]x[-0.5,0.5])."""
    c = Config()
    c.dim = 2
    c.nx, c.ny, c.nz = 1024, 256, 1
    c.xmin, c.xmax, c.ymin, c.ymax, c.zmin, c.zmax = -2.0, 2.0, -0.5, 0.5, 0.0, 1.0
    c.density, c.viscosity = 50.0, 1.0
    for f, t in enumerate((OUTLET, INLET, WALL, WALL, WALL, WALL)):
        c.bc_type[f] = t
        for q in range(3):
            c.bc_val[f][q] = 0.0
    c.bc_val[WEST][0] = 1.0
    c.num_outer_iter = 20000
    for q, (n, r, t) in enumerate(((10, 0.75, 1e-5), (10, 0.75, 1e-5), (10, 0.75, 1e-5), (20, 0.25, 1e-3))):
        c.n_iter[q], c.relax[q], c.tol[q] = n, r, t
    c.tol_mass = 1e-9
    for q, val in enumerate((0.0, 0.0, 0.0, 0.0, 273.0)):
        c.init[q] = val
    c.smoother = SMOOTHER_RBGS
    c.stop_rule = STOP_FAITHFUL
    c.tol_residual = 1e-5
    c.p_refine = 0
    c.log_every = 0
    c.check_every = 0
    # config.h:19-20, 45-63: thermal properties and TempBCs (declared, not solved by the reference)
    c.energy = 0
    c.n_iter_t, c.relax_t, c.tol_t = 10, 0.75, 1e-5
    c.conductivity, c.specific_heat = 1.0, 1.0
    for f, (t, v) in enumerate(((TEMP_NEUMANN, 0.0), (TEMP_NEUMANN, 0.0), (TEMP_DIRICHLET, 293.0),
                                (TEMP_DIRICHLET, 373.0), (TEMP_NEUMANN, 0.0), (TEMP_NEUMANN, 0.0))):
        c.temp_bc_type[f], c.temp_bc_val[f] = t, v
    return c


def _set_relax(c, alpha_u, alpha_p):
    for q in range(3):
        c.relax[q] = alpha_u
    c.relax[3] = alpha_p


def channel_config(nx=512, ny=128, iters=3000):
    """Plain channel: README defaults grid/physics (README.md:187-280) with alpha 0.7/0.3.

    512x128 on x in [-0.0625, 0.0625], y in [-0.0125, 0.0125]; rho=998.2, mu=1.002e-3;
    west inlet u=0.001, east outlet, north/south walls; fixed outer iteration count.
    """
    c = default_config()
    c.nx, c.ny = nx, ny
    c.xmin, c.xmax, c.ymin, c.ymax = -0.0625, 0.0625, -0.0125, 0.0125
    c.density, c.viscosity = 998.2, 1.002e-3
    c.bc_val[WEST][0] = 0.001
    _set_relax(c, 0.7, 0.3)
    c.num_outer_iter = iters
    c.stop_rule = STOP_NONE
    c.init[4] = 0.0
    return c


def cavity_config(n=1024, re=100.0, p_refine=1, tol_residual=1e-5, max_iter=200000):
    """2D lid-driven cavity on the unit square, rho=1, mu=1/Re, lid u=1 on the north wall."""
    c = default_config()
    c.nx, c.ny = n, n
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 1.0, 0.0, 1.0
    c.density, c.viscosity = 1.0, 1.0 / re
    for f in range(6):
        c.bc_type[f] = WALL
        for q in range(3):
            c.bc_val[f][q] = 0.0
    c.bc_val[NORTH][0] = 1.0
    _set_relax(c, 0.7, 0.3)
    c.stop_rule = STOP_SCALED
    c.tol_residual = tol_residual
    c.num_outer_iter = max_iter
    c.p_refine = p_refine
    c.init[4] = 0.0
    return c


def cavity3d_config(n=256, re=400.0, tol_residual=1e-5, max_iter=200000):
    """3D lid-driven cavity on the unit cube, rho=1, mu=1/Re, lid u=1 on the top (z=max) face."""
    c = cavity_config(n=n, re=re, p_refine=0, tol_residual=tol_residual, max_iter=max_iter)
    c.dim = 3
    c.nz = n
    c.bc_val[NORTH][0] = 0.0
    c.bc_val[TOP][0] = 1.0
    return c


def adjoint_channel_config(nx=1024, ny=256):
    """Adjoint channel: [0,0.5]x[0,0.05], rho=1, mu=1e-3, west inlet profile, east outlet."""
    c = default_config()
    c.nx, c.ny = nx, ny
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 0.5, 0.0, 0.05
    c.density, c.viscosity = 1.0, 1e-3
    c.bc_val[WEST][0] = 0.02
    _set_relax(c, 0.7, 0.3)
    c.stop_rule = STOP_SCALED
    c.tol_residual = 1e-6
    c.num_outer_iter = 400000
    c.init[4] = 0.0
    return c


def splitmix64(seed):
    """SplitMix64 stream of uniforms in [0, 1) (53-bit mantissa)."""
    m = (1 << 64) - 1
    x = seed & m
    while True:
        x = (x + 0x9E3779B97F4A7C15) & m
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        z ^= z >> 31
        yield (z >> 11) * (1.0 / (1 << 53))


def uneven_batch_configs(n=64, seed=20261203, tol_residual=1e-4, max_iter=100000):
    """Seeded batch of uneven 2D cases.  Per case, in draw order: type (u < 0.5 cavity, else
    channel), nx = 16*(4 + floor(29u)), ny = 16*(4 + floor(29u)), Re = 10 + 990u; rho = 1,
    mu = 1/Re, U = 1.  Cavity: unit square, lid u = 1 on the north wall.  Channel: height 1,
    length nx/ny (square cells), uniform inlet u = 1 on the west face, east outlet, north/south
    walls.  Scaled-residual stop rule."""
    g = splitmix64(seed)
    out = []
    for _ in range(n):
        is_cavity = next(g) < 0.5
        nx = 16 * (4 + int(29 * next(g)))
        ny = 16 * (4 + int(29 * next(g)))
        re = 10.0 + 990.0 * next(g)
        c = cavity_config(n=nx, re=re, p_refine=0, tol_residual=tol_residual, max_iter=max_iter)
        c.nx, c.ny = nx, ny
        if not is_cavity:
            c.xmin, c.xmax, c.ymin, c.ymax = 0.0, nx / ny, 0.0, 1.0
            for f in range(6):
                c.bc_type[f] = WALL
                for q in range(3):
                    c.bc_val[f][q] = 0.0
            c.bc_type[EAST] = OUTLET
            c.bc_type[WEST] = INLET
            c.bc_val[WEST][0] = 1.0
        out.append(c)
    return out


def adjoint_inlet_profile(cfg, u_max=0.02, sigma=1e-3, seed=20261203):
    """Inlet profile of the adjoint channel: u_j = 4 u_max y_j (H - y_j) / H^2 + sigma xi_j with
    y_j = (j + 1/2) dy and xi_j standard normal from SplitMix64 uniforms through Box-Muller."""
    import math
    H = cfg.ymax - cfg.ymin
    dy = H / cfg.ny
    g = splitmix64(seed)
    out = []
    for j in range(cfg.ny):
        y = (j + 0.5) * dy
        u1 = next(g)
        u2 = next(g)
        xi = math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(2.0 * math.pi * u2)
        out.append(4.0 * u_max * y * (H - y) / (H * H) + sigma * xi)
    return out


# ---- binary state files (ifl_save_state / ifl_load_state) ----
STATE_MAGIC = b"IFLSTATE"
STATE_VERSION = 1
STATE_FIELDS = (FIELD_U, FIELD_V, FIELD_W, FIELD_P, FIELD_TEMP, FIELD_UF, FIELD_VF, FIELD_WF)


class StateHeader(C.LittleEndianStructure):
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("dim", C.c_int32), ("nx", C.c_int32),
                ("ny", C.c_int32), ("nz", C.c_int32), ("last_converged", C.c_int32), ("iter", C.c_int64),
                ("norm_max", C.c_double * 7), ("res_max", C.c_double * 5), ("last_scaled", C.c_double),
                ("payload_bytes", C.c_uint64), ("checksum", C.c_uint64)]


def fnv1a64(data):
    """FNV-1a (64-bit) of a bytes-like object: the checksum of a state file's field data."""
    h = 0xcbf29ce484222325
    for x in bytes(data):
        h = ((h ^ x) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _state_sizes(dim, nx, ny, nz):
    c = Config()
    c.dim, c.nx, c.ny, c.nz = dim, nx, ny, nz
    return [(f, c.field_size(f), c.shape(f)) for f in STATE_FIELDS]


def read_state(path, verify=True):
    """Parse a state file without the library: (header, {field id: float32 array, shaped})."""
    import numpy as np
    with open(path, "rb") as fh:
        raw = fh.read()
    if len(raw) < C.sizeof(StateHeader):
        raise ValueError(f"{path}: shorter than a state header")
    hd = StateHeader.from_buffer_copy(raw[:C.sizeof(StateHeader)])
    if hd.magic != STATE_MAGIC or hd.version != STATE_VERSION:
        raise ValueError(f"{path}: not a version-{STATE_VERSION} state file")
    payload = raw[C.sizeof(StateHeader):]
    sizes = _state_sizes(hd.dim, hd.nx, hd.ny, hd.nz)
    if len(payload) != hd.payload_bytes or hd.payload_bytes != 4 * sum(n for _, n, _ in sizes):
        raise ValueError(f"{path}: payload size does not match the grid")
    if verify and fnv1a64(payload) != hd.checksum:
        raise ValueError(f"{path}: checksum mismatch")
    data = np.frombuffer(payload, dtype="<f4")
    out, off = {}, 0
    for f, n, shp in sizes:
        out[f] = data[off:off + n].reshape(shp).copy()
        off += n
    return hd, out


def write_state(path, dim, nx, ny, nz, fields, iter=0, norm_max=None, res_max=None, last_scaled=0.0,
                last_converged=0):
    """Write a state file from host arrays (fields: {field id: array}; missing fields are zero)."""
    import numpy as np
    sizes = _state_sizes(dim, nx, ny, nz)
    parts = []
    for f, n, _ in sizes:
        a = np.zeros(n, dtype="<f4") if f not in fields else np.ascontiguousarray(fields[f], dtype="<f4").ravel()
        if a.size != n:
            raise ValueError(f"field {FIELD_NAMES[f]}: {a.size} values, the grid needs {n}")
        parts.append(a.tobytes())
    payload = b"".join(parts)
    hd = StateHeader()
    hd.magic, hd.version = STATE_MAGIC, STATE_VERSION
    hd.dim, hd.nx, hd.ny, hd.nz = dim, nx, ny, nz
    hd.last_converged, hd.iter = last_converged, iter
    for q, v in enumerate(norm_max or [0.0] * 7):
        hd.norm_max[q] = v
    for q, v in enumerate(res_max or [0.0] * 5):
        hd.res_max[q] = v
    hd.last_scaled = last_scaled
    hd.payload_bytes = len(payload)
    hd.checksum = fnv1a64(payload)
    with open(path, "wb") as fh:
        fh.write(bytes(hd))
        fh.write(payload)
