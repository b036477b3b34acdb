"""ctypes binding of liboptics_gpu.so (the C ABI in include/optics_gpu.h).

The binding mirrors the reference's Rust call shapes (asm::propagate,
fresnel::propagate, rayleigh_sommerfeld::propagate, PhaseMask::apply, ...):
each takes a complex field plus grid pitch, wavelength, distance and the
engine config, and returns (status, out, report).  Host-memory functions run
the f64 drop-in; `*_dev_*` functions take torch CUDA tensors.

There is no CPU fallback: importing this module loads the HIP library and
raises if it is absent.
"""
import ctypes as C
import os

import numpy as np

import oc_types as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboptics_gpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (hipcc, gfx950)")

# One HIP runtime per process: when PyTorch (used for device buffers) is
# present, let it load its runtime first so the library binds to the same one.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = C.CDLL(LIB_PATH)

vp, c_f64, c_i32, c_u64, c_sz = C.c_void_p, C.c_double, C.c_int32, C.c_uint64, C.c_size_t
P = C.POINTER

EXPORTS = {
    "oc_context_create": ([C.c_int, P(vp)], C.c_int),
    "oc_context_destroy": ([vp], None),
    "oc_context_set_stream": ([vp, vp], C.c_int),
    "oc_context_stream": ([vp], vp),
    "oc_context_synchronize": ([vp], C.c_int),
    "oc_last_error": ([vp], C.c_char_p),
    "oc_status_name": ([C.c_int], C.c_char_p),
    "oc_version": ([], C.c_char_p),
    "oc_context_enable_timing": ([vp, C.c_int], C.c_int),
    "oc_context_pass_timing": ([vp, P(c_f64), P(c_u64), C.c_int], C.c_int),
    "oc_context_set_timing_mask": ([vp, C.c_uint32], C.c_int),
    "oc_autotune_fft": ([vp, c_u64, C.c_int, P(C.c_int), P(c_f64)], C.c_int),
    "oc_context_set_fft_variant": ([vp, c_u64, C.c_int, C.c_int], C.c_int),
    "oc_context_fft_variant": ([vp, c_u64, C.c_int], C.c_int),
    "oc_context_set_fresnel_sep_layout": ([vp, c_u64, C.c_int, C.c_int], C.c_int),
    "oc_context_fresnel_sep_layout": ([vp, c_u64, P(C.c_int), P(C.c_int)], C.c_int),
    "oc_context_force_generic": ([vp, C.c_int], C.c_int),
    "oc_asm_sampling_report": ([P(T.BandGeom), c_f64, P(T.AsmConfig), P(T.AsmReport)], C.c_int),
    "oc_asm_propagate": ([vp, P(T.BandGeom), vp, c_f64, P(T.AsmConfig), vp, P(T.AsmReport)], C.c_int),
    "oc_asm_propagate_dev_c128": ([vp, P(T.BandGeom), vp, c_f64, P(T.AsmConfig), P(T.PhaseMask), vp,
                                   P(T.AsmReport)], C.c_int),
    "oc_asm_propagate_dev_c64": ([vp, P(T.BandGeom), vp, c_f64, P(T.AsmConfig), P(T.PhaseMask), vp,
                                  P(T.AsmReport)], C.c_int),
    "oc_fresnel_sampling_report": ([P(T.BandGeom), c_f64, P(T.FresnelConfig), P(T.FresnelReport)], C.c_int),
    "oc_fresnel_propagate_convolution": ([vp, P(T.BandGeom), vp, c_f64, P(T.FresnelConfig), vp,
                                          P(T.FresnelReport)], C.c_int),
    "oc_fresnel_propagate_convolution_dev_c128": ([vp, P(T.BandGeom), vp, c_f64, P(T.FresnelConfig), vp,
                                                   P(T.FresnelReport)], C.c_int),
    "oc_fresnel_propagate_convolution_dev_c64": ([vp, P(T.BandGeom), vp, c_f64, P(T.FresnelConfig), vp,
                                                  P(T.FresnelReport)], C.c_int),
    "oc_fraunhofer_regime_report": ([P(T.BandGeom), c_f64, P(T.FraunhoferConfig), P(T.FraunhoferReport)],
                                    C.c_int),
    "oc_fraunhofer_propagate": ([vp, P(T.BandGeom), vp, c_f64, P(T.FraunhoferConfig), vp,
                                 P(T.FraunhoferReport)], C.c_int),
    "oc_fraunhofer_propagate_dev_c128": ([vp, P(T.BandGeom), vp, c_f64, P(T.FraunhoferConfig), vp,
                                          P(T.FraunhoferReport)], C.c_int),
    "oc_fraunhofer_propagate_dev_c64": ([vp, P(T.BandGeom), vp, c_f64, P(T.FraunhoferConfig), vp,
                                         P(T.FraunhoferReport)], C.c_int),
    "oc_fresnel_propagate": ([vp, P(T.BandGeom), vp, c_f64, P(T.FresnelConfig), vp, P(T.FresnelReport)],
                             C.c_int),
    "oc_fresnel_propagate_dev_c128": ([vp, P(T.BandGeom), vp, c_f64, P(T.FresnelConfig), P(T.PhaseMask), vp,
                                       P(T.FresnelReport)], C.c_int),
    "oc_fresnel_propagate_dev_c64": ([vp, P(T.BandGeom), vp, c_f64, P(T.FresnelConfig), P(T.PhaseMask), vp,
                                      P(T.FresnelReport)], C.c_int),
    "oc_rs_sampling_report": ([P(T.BandGeom), c_f64, P(T.RsConfig), P(T.RsReport)], C.c_int),
    "oc_rs_propagate": ([vp, P(T.BandGeom), vp, c_f64, P(T.RsConfig), vp, P(T.RsReport)], C.c_int),
    "oc_rs_propagate_dev_c128": ([vp, P(T.BandGeom), vp, c_f64, P(T.RsConfig), vp, P(T.RsReport)], C.c_int),
    "oc_rs_propagate_dev_c64": ([vp, P(T.BandGeom), vp, c_f64, P(T.RsConfig), vp, P(T.RsReport)], C.c_int),
    "oc_propagate_batch_dev_c64": ([vp, P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_batch_dev_c128": ([vp, P(T.PropDesc), c_sz], C.c_int),
    "oc_phasemask_validate": ([P(T.PhaseMask)], C.c_int),
    "oc_phasemask_apply": ([vp, P(T.BandGeom), vp, P(T.PhaseMask), vp], C.c_int),
    "oc_phasemask_apply_dev_c128": ([vp, P(T.BandGeom), vp, P(T.PhaseMask), vp], C.c_int),
    "oc_phasemask_apply_dev_c64": ([vp, P(T.BandGeom), vp, P(T.PhaseMask), vp], C.c_int),
    "oc_gs_run": ([vp, P(T.GsProblem), P(T.GsSolution)], C.c_int),
    "oc_gs_run_batch": ([vp, P(T.GsProblem), P(T.GsSolution), c_sz, c_u64], C.c_int),
    "oc_gs_fft_run_dev_c64": ([vp, c_u64, c_u64, vp, vp, vp, c_u64, c_u64, P(c_f64)], C.c_int),
    "oc_gs_fft_run_dev_c128": ([vp, c_u64, c_u64, vp, vp, vp, c_u64, c_u64, P(c_f64)], C.c_int),
    "oc_adjoint_asm": ([vp, P(T.BandGeom), vp, c_f64, P(T.AsmConfig), vp, P(T.AsmReport)], C.c_int),
    "oc_adjoint_transmission": ([vp, P(T.BandGeom), P(T.PhaseMask), vp], C.c_int),
    "oc_adjoint_push_back": ([vp, P(T.DesignStep), c_sz, P(T.BandGeom), vp, P(T.AsmConfig), vp, vp], C.c_int),
    "oc_adjoint_grid_gradient_dev_c128": ([vp, P(T.DesignStep), c_sz, P(T.BandGeom), vp, P(T.AsmConfig), vp, vp,
                                           P(c_f64)], C.c_int),
    "oc_adjoint_grid_gradient_frames_dev_c128": ([vp, P(T.BandGeom), vp, c_sz, c_sz, P(T.DesignStep), c_sz,
                                                  P(T.AsmConfig), vp, vp, c_i32, c_i32, vp, P(c_f64)], C.c_int),
    "oc_adjoint_grid_gradient_frames_dev_c64": ([vp, P(T.BandGeom), vp, c_sz, c_sz, P(T.DesignStep), c_sz,
                                                 P(T.AsmConfig), vp, vp, c_i32, c_i32, vp, P(c_f64)], C.c_int),
    "oc_fd_gradient_dev_c128": ([vp, P(T.BandGeom), P(T.DesignStep), c_sz, c_sz, vp, vp, vp, vp, P(T.AsmConfig), vp,
                                 vp, vp, vp, P(c_sz)], C.c_int),
    "oc_fd_gradient_dev_c64": ([vp, P(T.BandGeom), P(T.DesignStep), c_sz, c_sz, vp, vp, vp, vp, P(T.AsmConfig), vp,
                                vp, vp, vp, P(c_sz)], C.c_int),
    "oc_adam_step_dev": ([vp, c_sz, vp, vp, vp, vp, vp, vp, C.c_uint64, P(T.AdamConfig)], C.c_int),
    "oc_propagate_chain_dev_c64": ([vp, P(T.PropDesc), P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_chain_dev_c128": ([vp, P(T.PropDesc), P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_chain_submit_dev_c64": ([vp, P(T.PropDesc), P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_chain_submit_dev_c128": ([vp, P(T.PropDesc), P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_batch_noalloc_dev_c64": ([vp, P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_batch_submit_dev_c64": ([vp, P(T.PropDesc), c_sz], C.c_int),
    "oc_propagate_batch_submit_dev_c128": ([vp, P(T.PropDesc), c_sz], C.c_int),
    "oc_context_wait": ([vp], C.c_int),
    "oc_autotune_batch_dev_c64": ([vp, P(T.PropDesc), c_sz, c_i32, P(T.Schedule), c_sz, P(c_sz), P(c_f64)], C.c_int),
    "oc_autotune_batch_dev_c128": ([vp, P(T.PropDesc), c_sz, c_i32, P(T.Schedule), c_sz, P(c_sz), P(c_f64)],
                                   C.c_int),
    "oc_schedule_set": ([vp, P(T.Schedule)], C.c_int),
    "oc_schedule_get": ([vp, c_sz, P(T.Schedule)], C.c_int),
    "oc_schedule_count": ([vp], c_sz),
    "oc_schedule_clear": ([vp], C.c_int),
    "oc_schedule_save": ([vp, C.c_char_p], C.c_int),
    "oc_schedule_load": ([vp, C.c_char_p, P(c_sz)], C.c_int),
    "oc_mtf": ([vp, vp, c_sz, c_sz, c_sz, c_f64, vp, vp], C.c_int),
    "oc_mtf_dev": ([vp, vp, c_sz, c_sz, c_sz, c_f64, vp, vp], C.c_int),
    "oc_render": ([vp, P(T.BandGeom), vp, c_sz, c_sz, P(T.RenderConfig), P(T.DesignStep), c_sz, P(T.AsmConfig), vp,
                   vp], C.c_int),
    "oc_render_dev_c128": ([vp, P(T.BandGeom), vp, c_sz, c_sz, P(T.RenderConfig), P(T.DesignStep), c_sz,
                            P(T.AsmConfig), vp, vp], C.c_int),
    "oc_render_dev_c64": ([vp, P(T.BandGeom), vp, c_sz, c_sz, P(T.RenderConfig), P(T.DesignStep), c_sz,
                           P(T.AsmConfig), vp, vp], C.c_int),
}


def _declare():
    for name, (args, res) in EXPORTS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res


_declare()


class OcError(RuntimeError):
    pass


FFT_VARIANTS = 4  # OC_FFT_VARIANTS


class Context:
    """oc_context: device, stream, scratch arena, plan caches."""

    def __init__(self, device=0):
        h = vp()
        st = lib.oc_context_create(device, C.byref(h))
        if st != 0:
            raise OcError(f"oc_context_create failed: {T.STATUS.get(st, st)}")
        self.h = h

    def close(self):
        if self.h:
            lib.oc_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib.oc_context_stream(self.h)

    def synchronize(self):
        return lib.oc_context_synchronize(self.h)

    def enable_timing(self, on=True):
        lib.oc_context_enable_timing(self.h, 1 if on else 0)

    PASS_NAMES = ["row_fwd", "col", "row_inv", "rs_kgen", "gs", "other", "sep_col", "sep_row"]

    def set_timing_passes(self, names=None):
        """Time only the named pass kinds (None: all)."""
        mask = 0xFFFFFFFF if names is None else sum(1 << self.PASS_NAMES.index(x) for x in names)
        return lib.oc_context_set_timing_mask(self.h, mask)

    def pass_timing(self, reset=True):
        """{pass: (total_ms, launches)} for row_fwd, col, row_inv, rs_kgen, gs, other, sep_col, sep_row."""
        ms = (c_f64 * 8)()
        n = (c_u64 * 8)()
        lib.oc_context_pass_timing(self.h, ms, n, 1 if reset else 0)
        names = ["row_fwd", "col", "row_inv", "rs_kgen", "gs", "other", "sep_col", "sep_row"]
        return {names[i]: (ms[i], n[i]) for i in range(8)}

    def autotune(self, n, double=False):
        """Times every fast-kernel layout variant at padded length n; keeps the best.
        Returns (variant, [ms per variant])."""
        v = C.c_int()
        ms = (c_f64 * FFT_VARIANTS)()
        st = lib.oc_autotune_fft(self.h, n, 1 if double else 0, C.byref(v), ms)
        if st:
            raise OcError(f"autotune failed: {T.STATUS.get(st, st)} {self.last_error()}")
        return v.value, list(ms)

    def set_fft_variant(self, n, variant, double=False):
        return lib.oc_context_set_fft_variant(self.h, n, 1 if double else 0, variant)

    def fft_variant(self, n, double=False):
        return lib.oc_context_fft_variant(self.h, n, 1 if double else 0)

    def set_fresnel_sep_layout(self, n, v_col, v_row):
        """Layouts of the two-pass complex64 Fresnel's column and row passes."""
        return lib.oc_context_set_fresnel_sep_layout(self.h, n, v_col, v_row)

    def fresnel_sep_layout(self, n):
        vc, vr = C.c_int(), C.c_int()
        st = lib.oc_context_fresnel_sep_layout(self.h, n, C.byref(vc), C.byref(vr))
        if st:
            raise OcError(f"fresnel_sep_layout: {T.STATUS.get(st, st)}")
        return vc.value, vr.value

    def force_generic(self, on=True):
        return lib.oc_context_force_generic(self.h, 1 if on else 0)

    def last_error(self):
        return lib.oc_last_error(self.h).decode()


_CTX = None


def context():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _geom(u, dx, dy, lam):
    ny, nx = u.shape
    return T.geom(nx, ny, dx, dy, lam)


class Mask:
    """PhaseMask::sampled arguments with HOST lattice values (f64 drop-in)."""

    def __init__(self, values, dx, dy, orientation=0.0, aperture=(T.APERTURE_FIELD, 0.0, 0.0),
                 center=(0.0, 0.0)):
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        ny, nx = self.values.shape
        kind, a, b = aperture
        self.struct = T.PhaseMask(self.values.ctypes.data, nx, ny, dx, dy, orientation,
                                  T.ClearAperture(kind, 0, a, b), center[0], center[1])


# ---------------------------------------------------------------------------
# f64 drop-in (host memory in, host memory out)
# ---------------------------------------------------------------------------

def asm_propagate(u, dx, dy, lam, z, cfg, ctx=None):
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    rep = T.AsmReport()
    st = lib.oc_asm_propagate(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, z, C.byref(cfg),
                              out.ctypes.data, C.byref(rep))
    return st, out, rep


def fresnel_propagate(u, dx, dy, lam, z, cfg, ctx=None):
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    rep = T.FresnelReport()
    st = lib.oc_fresnel_propagate(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, z, C.byref(cfg),
                                  out.ctypes.data, C.byref(rep))
    return st, out, rep


def fresnel_propagate_convolution(u, dx, dy, lam, z, cfg, ctx=None):
    """fresnel::propagate_convolution, host f64."""
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    rep = T.FresnelReport()
    st = lib.oc_fresnel_propagate_convolution(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, z,
                                              C.byref(cfg), out.ctypes.data, C.byref(rep))
    return st, out, rep


def fraunhofer_propagate(u, dx, dy, lam, z, cfg, ctx=None):
    """fraunhofer::propagate, host f64; the far field's pitch is
    (report.dx_out, report.dy_out)."""
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    rep = T.FraunhoferReport()
    st = lib.oc_fraunhofer_propagate(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, z, C.byref(cfg),
                                     out.ctypes.data, C.byref(rep))
    return st, out, rep


def fraunhofer_regime_report(shape, dx, dy, lam, z, cfg):
    ny, nx = shape
    rep = T.FraunhoferReport()
    lib.oc_fraunhofer_regime_report(C.byref(T.geom(nx, ny, dx, dy, lam)), z, C.byref(cfg), C.byref(rep))
    return rep


def rs_propagate(u, dx, dy, lam, z, cfg, ctx=None):
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    rep = T.RsReport()
    st = lib.oc_rs_propagate(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, z, C.byref(cfg),
                             out.ctypes.data, C.byref(rep))
    return st, out, rep


def phasemask_apply(u, dx, dy, lam, mask, ctx=None):
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    st = lib.oc_phasemask_apply(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, C.byref(mask.struct),
                                out.ctypes.data)
    return st, out


def phasemask_validate(mask):
    return lib.oc_phasemask_validate(C.byref(mask.struct))


def asm_sampling_report(shape, dx, dy, lam, z, cfg):
    ny, nx = shape
    rep = T.AsmReport()
    lib.oc_asm_sampling_report(C.byref(T.geom(nx, ny, dx, dy, lam)), z, C.byref(cfg), C.byref(rep))
    return rep


def fresnel_sampling_report(shape, dx, dy, lam, z, cfg):
    ny, nx = shape
    rep = T.FresnelReport()
    lib.oc_fresnel_sampling_report(C.byref(T.geom(nx, ny, dx, dy, lam)), z, C.byref(cfg), C.byref(rep))
    return rep


def rs_sampling_report(shape, dx, dy, lam, z, cfg):
    ny, nx = shape
    rep = T.RsReport()
    lib.oc_rs_sampling_report(C.byref(T.geom(nx, ny, dx, dy, lam)), z, C.byref(cfg), C.byref(rep))
    return rep


# ---------------------------------------------------------------------------
# Device-resident forms over torch CUDA tensors
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def dev_mask(values, dx, dy, orientation=0.0, aperture=(T.APERTURE_FIELD, 0.0, 0.0), center=(0.0, 0.0)):
    """PhaseMask with its lattice copied to the device; returns (struct, keepalive tensor)."""
    torch = _torch()
    v = torch.as_tensor(np.ascontiguousarray(values, np.float64)).cuda()
    ny, nx = v.shape
    kind, a, b = aperture
    s = T.PhaseMask(v.data_ptr(), nx, ny, dx, dy, orientation, T.ClearAperture(kind, 0, a, b), center[0],
                    center[1])
    return s, v


def dev_mask_from_tensor(values_dev, dx, dy, orientation=0.0, aperture=(T.APERTURE_FIELD, 0.0, 0.0),
                         center=(0.0, 0.0)):
    """PhaseMask over a lattice already on the device (contiguous float64
    CUDA tensor, (ny, nx)); the caller keeps the tensor alive while the
    struct is in use."""
    torch = _torch()
    if values_dev.dtype != torch.float64 or not values_dev.is_cuda or not values_dev.is_contiguous():
        raise ValueError("lattice must be a contiguous float64 CUDA tensor")
    ny, nx = values_dev.shape
    kind, a, b = aperture
    return T.PhaseMask(values_dev.data_ptr(), nx, ny, dx, dy, orientation, T.ClearAperture(kind, 0, a, b),
                       center[0], center[1])


def _dev_call(fn, u_dev, out_dev, geom, z, cfg, rep, mask_struct, ctx):
    torch = _torch()
    torch.cuda.synchronize()
    args = [ctx.h, C.byref(geom), u_dev.data_ptr(), z, C.byref(cfg)]
    if mask_struct is not False:
        args.append(C.byref(mask_struct) if mask_struct is not None else None)
    args += [out_dev.data_ptr(), C.byref(rep)]
    return fn(*args)


def asm_propagate_dev(u_dev, dx, dy, lam, z, cfg, mask=None, out=None, ctx=None):
    """u_dev: torch complex64 or complex128 CUDA tensor (ny, nx)."""
    torch = _torch()
    ctx = ctx or context()
    out = torch.empty_like(u_dev) if out is None else out
    ny, nx = u_dev.shape
    fn = lib.oc_asm_propagate_dev_c64 if u_dev.dtype == torch.complex64 else lib.oc_asm_propagate_dev_c128
    rep = T.AsmReport()
    st = _dev_call(fn, u_dev, out, T.geom(nx, ny, dx, dy, lam), z, cfg, rep, mask, ctx)
    return st, out, rep


def fresnel_propagate_dev(u_dev, dx, dy, lam, z, cfg, mask=None, out=None, ctx=None):
    torch = _torch()
    ctx = ctx or context()
    out = torch.empty_like(u_dev) if out is None else out
    ny, nx = u_dev.shape
    fn = lib.oc_fresnel_propagate_dev_c64 if u_dev.dtype == torch.complex64 else lib.oc_fresnel_propagate_dev_c128
    rep = T.FresnelReport()
    st = _dev_call(fn, u_dev, out, T.geom(nx, ny, dx, dy, lam), z, cfg, rep, mask, ctx)
    return st, out, rep


def fresnel_propagate_convolution_dev(u_dev, dx, dy, lam, z, cfg, out=None, ctx=None):
    torch = _torch()
    ctx = ctx or context()
    out = torch.empty_like(u_dev) if out is None else out
    ny, nx = u_dev.shape
    fn = (lib.oc_fresnel_propagate_convolution_dev_c64 if u_dev.dtype == torch.complex64
          else lib.oc_fresnel_propagate_convolution_dev_c128)
    rep = T.FresnelReport()
    st = _dev_call(fn, u_dev, out, T.geom(nx, ny, dx, dy, lam), z, cfg, rep, False, ctx)
    return st, out, rep


def fraunhofer_propagate_dev(u_dev, dx, dy, lam, z, cfg, out=None, ctx=None):
    torch = _torch()
    ctx = ctx or context()
    out = torch.empty_like(u_dev) if out is None else out
    ny, nx = u_dev.shape
    fn = (lib.oc_fraunhofer_propagate_dev_c64 if u_dev.dtype == torch.complex64
          else lib.oc_fraunhofer_propagate_dev_c128)
    rep = T.FraunhoferReport()
    st = _dev_call(fn, u_dev, out, T.geom(nx, ny, dx, dy, lam), z, cfg, rep, False, ctx)
    return st, out, rep


def rs_propagate_dev(u_dev, dx, dy, lam, z, cfg, out=None, ctx=None):
    torch = _torch()
    ctx = ctx or context()
    out = torch.empty_like(u_dev) if out is None else out
    ny, nx = u_dev.shape
    fn = lib.oc_rs_propagate_dev_c64 if u_dev.dtype == torch.complex64 else lib.oc_rs_propagate_dev_c128
    rep = T.RsReport()
    st = _dev_call(fn, u_dev, out, T.geom(nx, ny, dx, dy, lam), z, cfg, rep, False, ctx)
    return st, out, rep


def phasemask_apply_dev(u_dev, dx, dy, lam, mask_struct, out=None, ctx=None):
    torch = _torch()
    ctx = ctx or context()
    out = torch.empty_like(u_dev) if out is None else out
    ny, nx = u_dev.shape
    fn = lib.oc_phasemask_apply_dev_c64 if u_dev.dtype == torch.complex64 else lib.oc_phasemask_apply_dev_c128
    torch.cuda.synchronize()
    st = fn(ctx.h, C.byref(T.geom(nx, ny, dx, dy, lam)), u_dev.data_ptr(), C.byref(mask_struct), out.data_ptr())
    ctx.synchronize()
    return st, out


def propagate_batch_dev(descs, double=False, ctx=None):
    """descs: ctypes array of PropDesc (device pointers).  Returns status."""
    ctx = ctx or context()
    _torch().cuda.synchronize()
    fn = lib.oc_propagate_batch_dev_c128 if double else lib.oc_propagate_batch_dev_c64
    return fn(ctx.h, descs, len(descs))


def autotune_batch_dev(descs, double=False, sweeps=1, ctx=None):
    """oc_autotune_batch_dev_*: tunes the schedules of this batch (it runs
    the batch several times).  Returns (status, [pick dicts], search_ms)."""
    ctx = ctx or context()
    _torch().cuda.synchronize()
    fn = lib.oc_autotune_batch_dev_c128 if double else lib.oc_autotune_batch_dev_c64
    picks = (T.Schedule * 64)()
    n = c_sz(0)
    ms = c_f64(0.0)
    st = fn(ctx.h, descs, len(descs), sweeps, picks, 64, C.byref(n), C.byref(ms))
    return st, [T.schedule_dict(picks[i]) for i in range(n.value)], ms.value


def schedules(ctx=None):
    """The context's schedule picks as dicts."""
    ctx = ctx or context()
    out = []
    for i in range(lib.oc_schedule_count(ctx.h)):
        s = T.Schedule()
        if lib.oc_schedule_get(ctx.h, i, C.byref(s)) == 0:
            out.append(T.schedule_dict(s))
    return out


def schedule_set(d, ctx=None):
    ctx = ctx or context()
    return lib.oc_schedule_set(ctx.h, C.byref(T.schedule_from(d)))


def schedule_clear(ctx=None):
    ctx = ctx or context()
    return lib.oc_schedule_clear(ctx.h)


def schedule_save(path, ctx=None):
    ctx = ctx or context()
    return lib.oc_schedule_save(ctx.h, os.fsencode(path))


def schedule_load(path, ctx=None):
    """Returns (status, entries installed)."""
    ctx = ctx or context()
    n = c_sz(0)
    st = lib.oc_schedule_load(ctx.h, os.fsencode(path), C.byref(n))
    return st, n.value


def propagate_batch_noalloc_dev(descs, ctx=None):
    """oc_propagate_batch_noalloc_dev_c64: the complex64 batch on the
    library's scratch array (no device allocation).  Returns status."""
    ctx = ctx or context()
    _torch().cuda.synchronize()
    return lib.oc_propagate_batch_noalloc_dev_c64(ctx.h, descs, len(descs))


def propagate_chain_dev(first, second, double=False, ctx=None):
    """first, second: ctypes arrays of PropDesc of one length; second[i]
    propagates first[i]'s output (first[i].out may be 0).  Returns status."""
    ctx = ctx or context()
    _torch().cuda.synchronize()
    if len(first) != len(second):
        raise ValueError("first and second must have one length")
    fn = lib.oc_propagate_chain_dev_c128 if double else lib.oc_propagate_chain_dev_c64
    return fn(ctx.h, first, second, len(first))


def propagate_chain_submit_dev(first, second, double=False, ctx=None, sync_torch=True):
    """Queues a chain without waiting (first and second must stay alive
    until wait_dev); returns the first host-detected status.  sync_torch:
    synchronise torch's stream first (the caller's inputs written by torch)."""
    ctx = ctx or context()
    if sync_torch:
        _torch().cuda.synchronize()
    if len(first) != len(second):
        raise ValueError("first and second must have one length")
    fn = lib.oc_propagate_chain_submit_dev_c128 if double else lib.oc_propagate_chain_submit_dev_c64
    return fn(ctx.h, first, second, len(first))


def propagate_batch_submit_dev(descs, double=False, ctx=None):
    """Queues a batch without waiting (descs must stay alive until wait_dev);
    returns the first host-detected status."""
    ctx = ctx or context()
    _torch().cuda.synchronize()
    fn = lib.oc_propagate_batch_submit_dev_c128 if double else lib.oc_propagate_batch_submit_dev_c64
    return fn(ctx.h, descs, len(descs))


def wait_dev(ctx=None):
    """Resolves every queued batch of the context; returns the first non-OK status."""
    ctx = ctx or context()
    return lib.oc_context_wait(ctx.h)


class DevEngine:
    """Adapter giving the device-resident forms the oracle's host call shapes
    (numpy in, numpy out), for running the same parity/physics cases."""

    def __init__(self, dtype="c64"):
        torch = _torch()
        self.cdt = torch.complex64 if dtype == "c64" else torch.complex128
        self.ndt = np.complex64 if dtype == "c64" else np.complex128

    def _up(self, u):
        torch = _torch()
        return torch.as_tensor(np.ascontiguousarray(u, self.ndt)).cuda()

    @staticmethod
    def _down(t, st=0):
        """The output as complex128, or None after a failed call (the output
        buffer is unspecified then and is not read)."""
        return t.cpu().numpy().astype(np.complex128) if st == 0 else None

    def asm_propagate(self, u, dx, dy, lam, z, cfg):
        st, out, rep = asm_propagate_dev(self._up(u), dx, dy, lam, z, cfg)
        return st, self._down(out, st), rep

    def fresnel_propagate(self, u, dx, dy, lam, z, cfg):
        st, out, rep = fresnel_propagate_dev(self._up(u), dx, dy, lam, z, cfg)
        return st, self._down(out, st), rep

    def rs_propagate(self, u, dx, dy, lam, z, cfg):
        st, out, rep = rs_propagate_dev(self._up(u), dx, dy, lam, z, cfg)
        return st, self._down(out, st), rep

    def fresnel_propagate_convolution(self, u, dx, dy, lam, z, cfg):
        st, out, rep = fresnel_propagate_convolution_dev(self._up(u), dx, dy, lam, z, cfg)
        return st, self._down(out, st), rep

    def fraunhofer_propagate(self, u, dx, dy, lam, z, cfg):
        st, out, rep = fraunhofer_propagate_dev(self._up(u), dx, dy, lam, z, cfg)
        return st, self._down(out, st), rep

    @staticmethod
    def fraunhofer_regime_report(shape, dx, dy, lam, z, cfg):
        return fraunhofer_regime_report(shape, dx, dy, lam, z, cfg)

    class Mask:
        def __init__(self, values, dx, dy, orientation=0.0, aperture=(T.APERTURE_FIELD, 0.0, 0.0),
                     center=(0.0, 0.0)):
            self.struct, self.keep = dev_mask(values, dx, dy, orientation, aperture, center)

    def phasemask_apply(self, u, dx, dy, lam, mask):
        st, out = phasemask_apply_dev(self._up(u), dx, dy, lam, mask.struct)
        return st, self._down(out, st)


# ---------------------------------------------------------------------------
# Gerchberg-Saxton
# ---------------------------------------------------------------------------

def gs_run(u_in, dx, dy, lam, z_total, target_amplitude, merit_target, asm_cfg, aperture=(0, 0.0, 0.0),
           center=(0.0, 0.0), lo=-np.pi, hi=np.pi, carrier_values=None, max_iterations=100, patience=10,
           merit_tolerance=1e-4, variant=0, beta=1.0, init_phase=None, init_seed=None, ctx=None):
    """gs::run (gs.rs:347) with the MseToTarget merit; same call shape as the
    oracle's.  init_phase: GsInit::Given; init_seed: GsInit::Random{seed};
    neither: GsInit::Zero (the carrier values)."""
    ctx = ctx or context()
    u_in = _c128(u_in)
    ny, nx = u_in.shape
    ta = np.ascontiguousarray(target_amplitude, np.float64)
    mt = np.ascontiguousarray(merit_target, np.float64)
    cv = np.ascontiguousarray(carrier_values if carrier_values is not None else np.zeros((ny, nx)), np.float64)
    ip = np.ascontiguousarray(init_phase if init_phase is not None else np.zeros((ny, nx)), np.float64)
    prob = T.GsProblem(T.geom(nx, ny, dx, dy, lam), u_in.ctypes.data, z_total, ta.ctypes.data, mt.ctypes.data,
                       asm_cfg, T.ClearAperture(aperture[0], 0, aperture[1], aperture[2]), center[0], center[1],
                       lo, hi, cv.ctypes.data, max_iterations, patience, merit_tolerance, variant,
                       _init_kind(init_phase, init_seed), beta, ip.ctypes.data, init_seed or 0)
    phase, ti, tae, pe, il = (np.empty((ny, nx)) for _ in range(5))
    hist = (T.GsRecord * max_iterations)()
    sol = T.GsSolution(phase.ctypes.data, 0.0, 0, 0.0, 0, 0, 0, C.cast(hist, vp), ti.ctypes.data,
                       tae.ctypes.data, pe.ctypes.data, il.ctypes.data)
    st = lib.oc_gs_run(ctx.h, C.byref(prob), C.byref(sol))
    if st != 0:
        return st, None
    k = sol.iterations
    history = [(h.iteration, h.merit, h.best_merit, h.amplitude_mse, h.target_error_rms, h.phase_update_rms)
               for h in hist[:k]]
    return st, dict(phase=phase, merit=sol.merit, best_iteration=sol.best_iteration,
                    amplitude_mse=sol.amplitude_mse, converged=bool(sol.converged), iterations=k, history=history,
                    target_intensity=ti, target_amplitude_error=tae, phase_error=pe, illumination_amplitude=il)


def _init_kind(init_phase, init_seed):
    """GsInit as oc_gs_problem.init_kind: Given (2), Random (1) or Zero (0)."""
    return 2 if init_phase is not None else (1 if init_seed is not None else 0)


def gs_run_batch(problems, graph_chunk=0, ctx=None):
    """oc_gs_run_batch: problems = list of dicts with gs_run's arguments
    (u_in, dx, dy, lam, z_total, target_amplitude, merit_target, asm_cfg and
    the keywords); returns (status, [solution dict per problem])."""
    ctx = ctx or context()
    keep, probs, sols, outs = [], (T.GsProblem * len(problems))(), (T.GsSolution * len(problems))(), []
    for i, p in enumerate(problems):
        u_in = _c128(p["u_in"])
        ny, nx = u_in.shape
        kw = dict(aperture=(0, 0.0, 0.0), center=(0.0, 0.0), lo=-np.pi, hi=np.pi, carrier_values=None,
                  max_iterations=100, patience=10, merit_tolerance=1e-4, variant=0, beta=1.0, init_phase=None,
                  init_seed=None)
        kw.update({k: v for k, v in p.items() if k in kw})
        ta = np.ascontiguousarray(p["target_amplitude"], np.float64)
        mt = np.ascontiguousarray(p["merit_target"], np.float64)
        cv = np.ascontiguousarray(kw["carrier_values"] if kw["carrier_values"] is not None else np.zeros((ny, nx)),
                                  np.float64)
        ip = np.ascontiguousarray(kw["init_phase"] if kw["init_phase"] is not None else np.zeros((ny, nx)),
                                  np.float64)
        ap = kw["aperture"]
        probs[i] = T.GsProblem(T.geom(nx, ny, p["dx"], p["dy"], p["lam"]), u_in.ctypes.data, p["z_total"],
                               ta.ctypes.data, mt.ctypes.data, p["asm_cfg"], T.ClearAperture(ap[0], 0, ap[1], ap[2]),
                               kw["center"][0], kw["center"][1], kw["lo"], kw["hi"], cv.ctypes.data,
                               kw["max_iterations"], kw["patience"], kw["merit_tolerance"], kw["variant"],
                               _init_kind(kw["init_phase"], kw["init_seed"]), kw["beta"], ip.ctypes.data,
                               kw["init_seed"] or 0)
        arrays = [np.empty((ny, nx)) for _ in range(5)]
        hist = (T.GsRecord * kw["max_iterations"])()
        sols[i] = T.GsSolution(arrays[0].ctypes.data, 0.0, 0, 0.0, 0, 0, 0, C.cast(hist, vp), arrays[1].ctypes.data,
                               arrays[2].ctypes.data, arrays[3].ctypes.data, arrays[4].ctypes.data)
        keep += [u_in, ta, mt, cv, ip]
        outs.append((arrays, hist))
    st = lib.oc_gs_run_batch(ctx.h, probs, sols, len(problems), graph_chunk)
    if st != 0:
        return st, None
    res = []
    for i, (arrays, hist) in enumerate(outs):
        sol = sols[i]
        k = sol.iterations
        history = [(h.iteration, h.merit, h.best_merit, h.amplitude_mse, h.target_error_rms, h.phase_update_rms)
                   for h in hist[:k]]
        res.append(dict(phase=arrays[0], merit=sol.merit, best_iteration=sol.best_iteration,
                        amplitude_mse=sol.amplitude_mse, converged=bool(sol.converged), iterations=k,
                        history=history, target_intensity=arrays[1], target_amplitude_error=arrays[2],
                        phase_error=arrays[3], illumination_amplitude=arrays[4]))
    return st, res


def gs_fft_run_dev(source_amp, target_amp, phases, iterations, graph_chunk=0, ctx=None):
    """FFT-form GS over restarts on torch CUDA tensors (float32 -> c64 form,
    float64 -> c128 form).  phases (R, n, n) is updated in place.
    Returns (status, final normalised intensity error per restart)."""
    torch = _torch()
    ctx = ctx or context()
    r, n, _ = phases.shape
    err = (c_f64 * r)()
    fn = lib.oc_gs_fft_run_dev_c64 if phases.dtype == torch.float32 else lib.oc_gs_fft_run_dev_c128
    torch.cuda.synchronize()
    st = fn(ctx.h, n, r, source_amp.data_ptr(), target_amp.data_ptr(), phases.data_ptr(), iterations, graph_chunk,
            err)
    return st, np.array(err[:])


# ---------------------------------------------------------------------------
# Adjoint
# ---------------------------------------------------------------------------

def _steps(steps):
    """steps: list of (Mask or None, z_after, medium) -> ctypes array."""
    arr = (T.DesignStep * len(steps))()
    for i, (m, z, med) in enumerate(steps):
        arr[i].has_element = 1 if m is not None else 0
        if m is not None:
            arr[i].mask = m.struct
        arr[i].z_after = z
        arr[i].medium = med
    return arr


def adjoint_push_back(steps, u, dx, dy, lam, asm_cfg, dmdi, ctx=None):
    """adjoint::push_back (adjoint.rs:140) with host buffers; oracle call shape."""
    ctx = ctx or context()
    u = _c128(u)
    ne = sum(1 for s in steps if s[0] is not None)
    out = np.zeros((ne,) + u.shape)
    d = np.ascontiguousarray(dmdi, np.float64)
    st = lib.oc_adjoint_push_back(ctx.h, _steps(steps), len(steps), C.byref(_geom(u, dx, dy, lam)), u.ctypes.data,
                                  C.byref(asm_cfg), d.ctypes.data, out.ctypes.data)
    return st, out


def adjoint_asm(u, dx, dy, lam, z, cfg, ctx=None):
    """adjoint::asm_adjoint (adjoint.rs:90-92): A(z)^dagger = A(-z), host f64."""
    ctx = ctx or context()
    u = _c128(u)
    out = np.empty_like(u)
    rep = T.AsmReport()
    st = lib.oc_adjoint_asm(ctx.h, C.byref(_geom(u, dx, dy, lam)), u.ctypes.data, z, C.byref(cfg),
                            out.ctypes.data, C.byref(rep))
    return st, out, rep


def adjoint_transmission(mask, shape, dx, dy, lam, ctx=None):
    ctx = ctx or context()
    ny, nx = shape
    out = np.empty((ny, nx), np.complex128)
    st = lib.oc_adjoint_transmission(ctx.h, C.byref(T.geom(nx, ny, dx, dy, lam)), C.byref(mask.struct),
                                     out.ctypes.data)
    return st, out


def adjoint_grid_gradient_dev(dev_steps, u_dev, dx, dy, lam, asm_cfg, target_dev, grads_dev=None, ctx=None):
    """adjoint::grid_gradient (adjoint.rs:268) for one band/scene with MseToTarget.
    dev_steps: list of (mask struct with device values or None, z_after, medium).
    Returns (status, merit); gradients land in grads_dev (n_elements, ny, nx) float64."""
    torch = _torch()
    ctx = ctx or context()
    arr = (T.DesignStep * len(dev_steps))()
    for i, (ms, z, med) in enumerate(dev_steps):
        arr[i].has_element = 1 if ms is not None else 0
        if ms is not None:
            arr[i].mask = ms
        arr[i].z_after = z
        arr[i].medium = med
    ny, nx = u_dev.shape
    merit = c_f64()
    torch.cuda.synchronize()
    st = lib.oc_adjoint_grid_gradient_dev_c128(ctx.h, arr, len(dev_steps), C.byref(T.geom(nx, ny, dx, dy, lam)),
                                               u_dev.data_ptr(), C.byref(asm_cfg), target_dev.data_ptr(),
                                               grads_dev.data_ptr() if grads_dev is not None else None,
                                               C.byref(merit))
    return st, merit.value


def adjoint_grid_gradient_dev_prepared(dev_steps, u_dev, dx, dy, lam, asm_cfg, target_dev, grads_dev, ctx=None):
    """The call of adjoint_grid_gradient_dev with its arguments marshalled
    once, as a compiled caller holds them: returns a function that runs
    oc_adjoint_grid_gradient_dev_c128 and returns (status, merit).  The
    caller orders its own work on other streams before each call (no torch
    synchronisation here)."""
    ctx = ctx or context()
    arr = (T.DesignStep * len(dev_steps))()
    for i, (ms, z, med) in enumerate(dev_steps):
        arr[i].has_element = 1 if ms is not None else 0
        if ms is not None:
            arr[i].mask = ms
        arr[i].z_after = z
        arr[i].medium = med
    ny, nx = u_dev.shape
    g = T.geom(nx, ny, dx, dy, lam)
    merit = c_f64()
    fn = lib.oc_adjoint_grid_gradient_dev_c128
    args = (ctx.h, arr, len(dev_steps), C.byref(g), u_dev.data_ptr(), C.byref(asm_cfg), target_dev.data_ptr(),
            grads_dev.data_ptr(), C.byref(merit))
    keep = (arr, g, merit, asm_cfg, u_dev, target_dev, grads_dev)

    def call():
        _ = keep
        st = fn(*args)
        return st, merit.value

    return call


def adjoint_grid_gradient_frames_dev(designs, inputs_dev, dx, dy, bands, asm_cfg, target_dev, grads_dev=None,
                                     sel_band=-1, sel_scene=-1, ctx=None):
    """adjoint::grid_gradient (adjoint.rs:268-340) with MseToTarget over every
    (band, scene) frame.  designs[band] = step list (mask structs with device
    lattices); inputs_dev complex128/complex64 (bands, scenes, ny, nx);
    target_dev float64 (ny, nx); grads_dev float64 (n_elements, ny, nx) or
    None.  Returns (status, merit)."""
    torch = _torch()
    ctx = ctx or context()
    nb, ns, ny, nx = inputs_dev.shape
    if not inputs_dev.is_contiguous() or target_dev.dtype != torch.float64 or not target_dev.is_contiguous():
        raise ValueError("inputs_dev and target_dev must be contiguous (target float64)")
    fn = {torch.complex128: lib.oc_adjoint_grid_gradient_frames_dev_c128,
          torch.complex64: lib.oc_adjoint_grid_gradient_frames_dev_c64}[inputs_dev.dtype]
    nsteps = len(designs[0])
    table = _design_table([[steps] for steps in designs])
    b = np.ascontiguousarray(bands, np.float64)
    merit = c_f64()
    torch.cuda.synchronize()
    st = fn(ctx.h, C.byref(T.geom(nx, ny, dx, dy, 0.0)), b.ctypes.data, nb, ns, table, nsteps, C.byref(asm_cfg),
            inputs_dev.data_ptr(), target_dev.data_ptr(), sel_band, sel_scene,
            grads_dev.data_ptr() if grads_dev is not None else None, C.byref(merit))
    return st, merit.value


def adjoint_grid_gradient_frames_dev_prepared(designs, inputs_dev, dx, dy, bands, asm_cfg, target_dev, grads_dev,
                                              ctx=None):
    """adjoint_grid_gradient_frames_dev with its arguments marshalled once:
    returns a function that runs the C entry point and returns (status,
    merit).  No torch synchronisation per call."""
    torch = _torch()
    ctx = ctx or context()
    nb, ns, ny, nx = inputs_dev.shape
    if not inputs_dev.is_contiguous() or target_dev.dtype != torch.float64 or not target_dev.is_contiguous():
        raise ValueError("inputs_dev and target_dev must be contiguous (target float64)")
    fn = {torch.complex128: lib.oc_adjoint_grid_gradient_frames_dev_c128,
          torch.complex64: lib.oc_adjoint_grid_gradient_frames_dev_c64}[inputs_dev.dtype]
    nsteps = len(designs[0])
    table = _design_table([[steps] for steps in designs])
    b = np.ascontiguousarray(bands, np.float64)
    g = T.geom(nx, ny, dx, dy, 0.0)
    merit = c_f64()
    args = (ctx.h, C.byref(g), b.ctypes.data, nb, ns, table, nsteps, C.byref(asm_cfg), inputs_dev.data_ptr(),
            target_dev.data_ptr(), -1, -1, grads_dev.data_ptr(), C.byref(merit))
    keep = (table, b, g, merit, asm_cfg, inputs_dev, target_dev, grads_dev)

    def call():
        _ = keep
        st = fn(*args)
        return st, merit.value

    return call


def _design_table(designs):
    """designs[band][variant] = [(mask struct or PhaseMask or None, z_after, medium), ...]
    -> flat DesignStep array [band][variant][nsteps]."""
    flat = [st for per_band in designs for steps in per_band for st in steps]
    arr = (T.DesignStep * max(len(flat), 1))()
    for i, (m, z, med) in enumerate(flat):
        arr[i].has_element = 1 if m is not None else 0
        if m is not None:
            arr[i].mask = m.struct if hasattr(m, "struct") else m
        arr[i].z_after = z
        arr[i].medium = med
    return arr


def render(designs, inputs, dx, dy, bands, asm_cfg, rc, ctx=None):
    """Renderer::render (observe.rs:213-314), host buffers.  designs[band][variant]:
    step lists (variant 0 = the band's wavelength; with the lambda channel
    1 = lambda + dl, 2 = lambda - dl); inputs: complex (bands, scenes, renders, ny, nx)
    from the source closure, renders ordered base, x+, x-, y+, y-, l+, l- (enabled ones).
    Returns (status, obs (channels, bands, scenes, ny, nx) float64)."""
    ctx = ctx or context()
    u = _c128(inputs)
    nb, ns, nr, ny, nx = u.shape
    if nr != rc.renders():
        raise ValueError(f"inputs hold {nr} renders per frame, the config needs {rc.renders()}")
    nsteps = len(designs[0][0]) if nb else 0
    obs = np.zeros((rc.channels(), nb, ns, ny, nx))
    b = np.ascontiguousarray(bands, np.float64)
    st = lib.oc_render(ctx.h, C.byref(T.geom(nx, ny, dx, dy, 0.0)), b.ctypes.data, nb, ns, C.byref(rc),
                       _design_table(designs), nsteps, C.byref(asm_cfg), u.ctypes.data, obs.ctypes.data)
    return st, obs


def render_dev(designs, inputs_dev, dx, dy, bands, asm_cfg, rc, obs_dev, ctx=None):
    """Device form of render: inputs_dev complex128 or complex64 (bands, scenes,
    renders, ny, nx), mask values on the device (dev_mask), obs_dev float64
    (channels, bands, scenes, ny, nx).  Returns the status."""
    torch = _torch()
    ctx = ctx or context()
    nb, ns, nr, ny, nx = inputs_dev.shape
    if nr != rc.renders():
        raise ValueError(f"inputs hold {nr} renders per frame, the config needs {rc.renders()}")
    if tuple(obs_dev.shape) != (rc.channels(), nb, ns, ny, nx) or obs_dev.dtype != torch.float64:
        raise ValueError("obs_dev must be float64 (channels, bands, scenes, ny, nx)")
    if not (inputs_dev.is_contiguous() and obs_dev.is_contiguous()):
        raise ValueError("inputs_dev and obs_dev must be contiguous")
    fn = {torch.complex128: lib.oc_render_dev_c128, torch.complex64: lib.oc_render_dev_c64}[inputs_dev.dtype]
    nsteps = len(designs[0][0]) if nb else 0
    b = np.ascontiguousarray(bands, np.float64)
    torch.cuda.synchronize()
    return fn(ctx.h, C.byref(T.geom(nx, ny, dx, dy, 0.0)), b.ctypes.data, nb, ns, C.byref(rc),
              _design_table(designs), nsteps, C.byref(asm_cfg), inputs_dev.data_ptr(), obs_dev.data_ptr())


def fd_gradient_dev(designs, u_dev, dx, dy, lam, asm_cfg, target_dev, up, dn, lo, hi, ctx=None):
    """fd::gradient (fd.rs:139-184) for the objective trace + MseToTarget on the
    device.  designs[i] = (up design, down design) step lists of slot i (mask
    structs with device lattices); u_dev complex128/complex64 (ny, nx);
    target_dev float64 (ny, nx); up, dn, lo, hi: per-slot probe values and
    bounds.  Returns (status, grad, merits (dim, 2), evaluations)."""
    torch = _torch()
    ctx = ctx or context()
    dim = len(up)
    ny, nx = u_dev.shape
    if tuple(target_dev.shape) != (ny, nx) or target_dev.dtype != torch.float64 or not target_dev.is_contiguous():
        raise ValueError("target_dev must be contiguous float64 of the field's shape")
    if not u_dev.is_contiguous():
        raise ValueError("u_dev must be contiguous")
    fn = {torch.complex128: lib.oc_fd_gradient_dev_c128, torch.complex64: lib.oc_fd_gradient_dev_c64}[u_dev.dtype]
    nsteps = len(designs[0][0]) if dim else 0
    table = _design_table([list(pair) for pair in designs])
    a = [np.ascontiguousarray(v, np.float64) for v in (up, dn, lo, hi)]
    g = np.zeros(dim)
    mer = np.zeros((dim, 2))
    ev = c_sz(0)
    torch.cuda.synchronize()
    st = fn(ctx.h, C.byref(T.geom(nx, ny, dx, dy, lam)), table, nsteps, dim, *[v.ctypes.data for v in a],
            C.byref(asm_cfg), u_dev.data_ptr(), target_dev.data_ptr(), g.ctypes.data, mer.ctypes.data, C.byref(ev))
    return st, g, mer, ev.value


def adam_step_dev(x_dev, g_dev, m_dev, v_dev, lo_dev, hi_dev, t, cfg, ctx=None):
    """One Adam epoch (fd.rs:210-244) on contiguous float64 device vectors."""
    torch = _torch()
    ctx = ctx or context()
    ts = (x_dev, g_dev, m_dev, v_dev, lo_dev, hi_dev)
    n = x_dev.numel()
    if any(t_.dtype != torch.float64 or not t_.is_contiguous() or t_.numel() != n for t_ in ts):
        raise ValueError("x, g, m, v, lo, hi must be contiguous float64 of one length")
    torch.cuda.synchronize()
    return lib.oc_adam_step_dev(ctx.h, n, *[t_.data_ptr() for t_ in ts], int(t), C.byref(cfg))


def mtf(psfs, dx, ctx=None):
    """analysis::mtf (analysis.rs:133-150) for one PSF (ny, nx) or a stack
    (count, ny, nx) of host intensities.  Returns (status, freqs, values)."""
    ctx = ctx or context()
    p = np.ascontiguousarray(psfs, np.float64)
    one = p.ndim == 2
    if one:
        p = p[None]
    count, ny, nx = p.shape
    v = np.zeros((count, nx // 2 + 1))
    f = np.zeros(nx // 2 + 1)
    st = lib.oc_mtf(ctx.h, p.ctypes.data, count, nx, ny, dx, v.ctypes.data, f.ctypes.data)
    return st, f, (v[0] if one else v)


def mtf_dev(psfs_dev, dx, values_dev, ctx=None):
    """Device form: psfs_dev float64 (count, ny, nx), values_dev float64
    (count, nx/2 + 1).  Returns (status, freqs)."""
    torch = _torch()
    ctx = ctx or context()
    count, ny, nx = psfs_dev.shape
    if psfs_dev.dtype != torch.float64 or not psfs_dev.is_contiguous():
        raise ValueError("psfs_dev must be contiguous float64")
    if tuple(values_dev.shape) != (count, nx // 2 + 1) or values_dev.dtype != torch.float64:
        raise ValueError("values_dev must be float64 (count, nx/2 + 1)")
    f = np.zeros(nx // 2 + 1)
    torch.cuda.synchronize()
    st = lib.oc_mtf_dev(ctx.h, psfs_dev.data_ptr(), count, nx, ny, dx, values_dev.data_ptr(), f.ctypes.data)
    return st, f
