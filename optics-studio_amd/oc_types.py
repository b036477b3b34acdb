"""ctypes mirrors of the POD structs and status codes in include/optics_gpu.h.

Kept free of any library loading so that both the product binding
(optics_gpu.py) and the test checker can share one definition of the ABI.
Each struct cites the Rust type of the reference it mirrors.
"""
import ctypes as C

c_u64 = C.c_uint64
c_i32 = C.c_int32
c_f64 = C.c_double

STATUS = {
    0: "OC_OK",
    1: "OC_INVALID_DISTANCE",
    2: "OC_INVALID_WAVELENGTH",
    3: "OC_INVALID_MEDIUM",
    4: "OC_EMPTY_GRID",
    5: "OC_NON_FINITE_SAMPLE",
    6: "OC_ALIASING_RISK",
    7: "OC_REGIME_VIOLATION",
    20: "OC_INVALID_APERTURE",
    21: "OC_OUTSIDE_FIELD",
    22: "OC_INVALID_AXIS",
    23: "OC_INVALID_PHASE",
    24: "OC_INVALID_ORIENTATION",
    25: "OC_INVALID_GRID",
    26: "OC_INVALID_GRID_PITCH",
    40: "OC_INV_NOT_FINITE",
    41: "OC_INV_OUT_OF_BOUNDS",
    42: "OC_INV_WRONG_DIMENSION",
    43: "OC_INV_CARRIER",
    44: "OC_INV_RENDER",
    45: "OC_INV_ADJOINT",
    100: "OC_ERR_INVALID_ARGUMENT",
    101: "OC_ERR_UNSUPPORTED",
    102: "OC_ERR_DEVICE",
    103: "OC_ERR_OUT_OF_MEMORY",
}
CODE = {v: k for k, v in STATUS.items()}

APERTURE_FIELD, APERTURE_CIRCLE, APERTURE_ELLIPSE, APERTURE_RECTANGLE = 0, 1, 2, 3
ENGINE_ASM, ENGINE_FRESNEL, ENGINE_RS, ENGINE_FRESNEL_CONV = 0, 1, 2, 7


class BandGeom(C.Structure):
    """BandField geometry (src/field.rs:99-105, 287-290)."""
    _fields_ = [("nx", c_u64), ("ny", c_u64), ("dx", c_f64), ("dy", c_f64), ("lambda_", c_f64)]


class AsmConfig(C.Structure):
    """AsmConfig (src/propagation/asm.rs:100-137)."""
    _fields_ = [("padding", c_f64), ("band_limited", c_i32), ("strict", c_i32),
                ("max_padding", c_f64), ("medium", c_f64)]

    @classmethod
    def default(cls, **kw):
        c = cls(2.0, 1, 1, 8.0, 1.0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class AsmReport(C.Structure):
    """SamplingReport (src/propagation/asm.rs:144-179)."""
    _fields_ = [("lambda_", c_f64), ("medium", c_f64), ("z", c_f64), ("nx", c_u64), ("ny", c_u64),
                ("dx", c_f64), ("dy", c_f64), ("requested_padding", c_f64), ("used_padding", c_f64),
                ("nx_pad", c_u64), ("ny_pad", c_u64), ("band_limit_fx", c_f64), ("band_limit_fy", c_f64),
                ("max_phase_step_x", c_f64), ("max_phase_step_y", c_f64), ("nyquist_ok", c_i32),
                ("_pad0", c_i32)]


class FresnelConfig(C.Structure):
    """FresnelConfig (src/propagation/fresnel.rs:99-123)."""
    _fields_ = [("padding", c_f64), ("strict", c_i32), ("_pad0", c_i32), ("max_padding", c_f64)]

    @classmethod
    def default(cls, **kw):
        c = cls(2.0, 1, 0, 8.0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class FresnelReport(C.Structure):
    """FresnelSamplingReport (src/propagation/fresnel.rs:130-217)."""
    _fields_ = [("lambda_", c_f64), ("z", c_f64), ("nx", c_u64), ("ny", c_u64), ("dx", c_f64), ("dy", c_f64),
                ("requested_padding", c_f64), ("used_padding", c_f64), ("nx_pad", c_u64), ("ny_pad", c_u64),
                ("max_phase_step_x", c_f64), ("max_phase_step_y", c_f64), ("nyquist_ok", c_i32),
                ("kernel_nyquist_ok", c_i32), ("max_kernel_step_x", c_f64), ("max_kernel_step_y", c_f64),
                ("max_phase_error_vs_asm", c_f64), ("paraxial_ok", c_i32), ("_pad0", c_i32)]


class FraunhoferConfig(C.Structure):
    """FraunhoferConfig (src/propagation/fraunhofer.rs:117-170); the Option
    fields are a flag plus value."""
    _fields_ = [("strict", c_i32), ("has_source_pos", c_i32), ("max_n_f", c_f64), ("source_x", c_f64),
                ("source_y", c_f64), ("has_aperture_extent", c_i32), ("_pad0", c_i32), ("aperture_extent", c_f64)]

    @classmethod
    def default(cls, strict=True, max_n_f=1.0, source_pos=None, aperture_extent=None):
        c = cls()
        c.strict = 1 if strict else 0
        c.max_n_f = max_n_f
        if source_pos is not None:
            c.has_source_pos, c.source_x, c.source_y = 1, source_pos[0], source_pos[1]
        if aperture_extent is not None:
            c.has_aperture_extent, c.aperture_extent = 1, aperture_extent
        return c


class FraunhoferReport(C.Structure):
    """FraunhoferReport (src/propagation/fraunhofer.rs:178-222)."""
    _fields_ = [("lambda_", c_f64), ("z", c_f64), ("nx", c_u64), ("ny", c_u64), ("dx", c_f64), ("dy", c_f64),
                ("dx_out", c_f64), ("dy_out", c_f64), ("n_f_x", c_f64), ("n_f_y", c_f64), ("source_x", c_f64),
                ("source_y", c_f64), ("extent_used", c_f64), ("extent_is_aperture", c_i32), ("regime_ok", c_i32),
                ("quad_phase_error_rad", c_f64), ("max_n_f", c_f64)]


class RsConfig(C.Structure):
    """RayleighSommerfeldConfig (src/propagation/rayleigh_sommerfeld.rs:95-116)."""
    _fields_ = [("padding", c_f64), ("strict", c_i32), ("_pad0", c_i32)]

    @classmethod
    def default(cls, **kw):
        c = cls(2.0, 1, 0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class RsReport(C.Structure):
    """RayleighSommerfeldSamplingReport (rayleigh_sommerfeld.rs:123-189)."""
    _fields_ = [("lambda_", c_f64), ("z", c_f64), ("nx", c_u64), ("ny", c_u64), ("dx", c_f64), ("dy", c_f64),
                ("padding", c_f64), ("nx_pad", c_u64), ("ny_pad", c_u64), ("max_kernel_step_x", c_f64),
                ("max_kernel_step_y", c_f64), ("kernel_nyquist_ok", c_i32), ("paraxial_ok", c_i32),
                ("max_phase_error_vs_fresnel", c_f64)]


class ClearAperture(C.Structure):
    """ClearAperture (src/element.rs:384-406)."""
    _fields_ = [("kind", c_i32), ("_pad0", c_i32), ("a", c_f64), ("b", c_f64)]


class PhaseMask(C.Structure):
    """PhaseMask::sampled arguments (src/element.rs:1321-1368)."""
    _fields_ = [("values", C.c_void_p), ("nx", c_u64), ("ny", c_u64), ("dx", c_f64), ("dy", c_f64),
                ("orientation", c_f64), ("aperture", ClearAperture), ("cx", c_f64), ("cy", c_f64)]


class PropDesc(C.Structure):
    """One problem of a batched propagation (no single reference type: a
    (BandField, z, engine config) triple of asm/fresnel/rayleigh_sommerfeld)."""
    _fields_ = [("geom", BandGeom), ("z", c_f64), ("engine", c_i32), ("band_limited", c_i32),
                ("strict", c_i32), ("status", c_i32), ("padding", c_f64), ("max_padding", c_f64),
                ("medium", c_f64), ("mask", C.POINTER(PhaseMask)), ("in_", C.c_void_p), ("out", C.c_void_p)]


SCHED_PAD, SCHED_FAST, SCHED_LDS = 0, 1, 2
SCHED_KINDS = {SCHED_PAD: "pad", SCHED_FAST: "fast", SCHED_LDS: "lds"}


class Schedule(C.Structure):
    """oc_schedule: one pick of the batch schedule tuner (no reference counterpart)."""
    _fields_ = [("length", c_u64), ("double_precision", c_i32), ("kind", c_i32), ("problems", c_u64),
                ("run_length", c_u64), ("variant", c_i32), ("nradix", c_i32), ("radix", c_i32 * 12), ("ms", c_f64)]


def schedule_dict(s):
    d = dict(length=s.length, double=bool(s.double_precision), kind=SCHED_KINDS[s.kind], problems=s.problems,
             ms=s.ms)
    if s.kind == SCHED_PAD:
        d["run_length"] = s.run_length
    elif s.kind == SCHED_FAST:
        d["variant"] = s.variant
    else:
        d["radix"] = list(s.radix[: s.nradix])
    return d


def schedule_from(d):
    s = Schedule()
    s.length = d["length"]
    s.double_precision = 1 if d.get("double") else 0
    s.kind = {v: k for k, v in SCHED_KINDS.items()}[d["kind"]]
    s.problems = d.get("problems", 1)
    s.run_length = d.get("run_length", 0)
    s.variant = d.get("variant", 0)
    rad = d.get("radix", [])
    s.nradix = len(rad)
    for i, r in enumerate(rad):
        s.radix[i] = r
    return s


class GsProblem(C.Structure):
    """GsProblem + GsConfig (crates/optics-inverse/src/gs.rs:125-305), GridPhase carrier, MseToTarget merit."""
    _fields_ = [("geom", BandGeom), ("input", C.c_void_p), ("z_total", c_f64), ("target_amplitude", C.c_void_p),
                ("merit_target", C.c_void_p), ("asm_cfg", AsmConfig), ("aperture", ClearAperture),
                ("cx", c_f64), ("cy", c_f64), ("lo", c_f64), ("hi", c_f64), ("carrier_values", C.c_void_p),
                ("max_iterations", c_u64), ("patience", c_u64), ("merit_tolerance", c_f64),
                ("variant", c_i32), ("init_kind", c_i32), ("beta", c_f64), ("init_phase", C.c_void_p),
                ("seed", c_u64)]


class GsRecord(C.Structure):
    """GsRecord (gs.rs:254-268)."""
    _fields_ = [("iteration", c_u64), ("merit", c_f64), ("best_merit", c_f64), ("amplitude_mse", c_f64),
                ("target_error_rms", c_f64), ("phase_update_rms", c_f64)]


class GsSolution(C.Structure):
    """GsSolution (gs.rs:271-305) with caller-allocated arrays."""
    _fields_ = [("phase", C.c_void_p), ("merit", c_f64), ("best_iteration", c_u64), ("amplitude_mse", c_f64),
                ("converged", c_i32), ("_pad0", c_i32), ("iterations", c_u64), ("history", C.c_void_p),
                ("target_intensity", C.c_void_p), ("target_amplitude_error", C.c_void_p),
                ("phase_error", C.c_void_p), ("illumination_amplitude", C.c_void_p)]


class DesignStep(C.Structure):
    """Design step (crates/optics-inverse/src/carrier.rs:35-43) with a PhaseMask element."""
    _fields_ = [("has_element", c_i32), ("_pad0", c_i32), ("mask", PhaseMask), ("z_after", c_f64),
                ("medium", c_f64)]


class RenderConfig(C.Structure):
    """Renderer derivative channels and FD steps (observe.rs:50-58, 152-170)."""
    _fields_ = [("deriv_x", c_i32), ("deriv_y", c_i32), ("deriv_lambda", c_i32), ("_pad0", c_i32),
                ("scene_step", c_f64), ("lambda_step", c_f64)]

    @staticmethod
    def make(x=False, y=False, lam=False, scene_step=1e-6, lambda_step=1e-9):
        """Renderer::new defaults: no derivatives, 1 um scene step, 1 nm band step."""
        return RenderConfig(int(x), int(y), int(lam), 0, float(scene_step), float(lambda_step))

    def renders(self):
        return 1 + 2 * (self.deriv_x + self.deriv_y + self.deriv_lambda)

    def variants(self):
        return 3 if self.deriv_lambda else 1

    def channels(self):
        return 1 + self.deriv_x + self.deriv_y + self.deriv_lambda


class AdamConfig(C.Structure):
    """AdamConfig (crates/optics-inverse/src/fd.rs:62-93); has_clip = 0 is clip_norm = None."""
    _fields_ = [("lr_frac", c_f64), ("beta1", c_f64), ("beta2", c_f64), ("eps", c_f64), ("has_clip", c_i32),
                ("_pad0", c_i32), ("clip_norm", c_f64)]

    @staticmethod
    def default(**kw):
        c = AdamConfig(0.02, 0.9, 0.999, 1.0e-8, 1, 0, 1.0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


def geom(nx, ny, dx, dy, lam):
    return BandGeom(int(nx), int(ny), float(dx), float(dy), float(lam))


def report_dict(rep):
    return {name: getattr(rep, name) for name, _ in rep._fields_ if not name.startswith("_pad")}
