"""ctypes binding of libabt_diff.so (include/abt_diff.h).

The library is the product path: every call here goes to the HIP library, and
loading fails loudly when it has not been built (there is no CPU fallback).
Device entry points take raw device pointers; `torch` tensors are accepted and
passed by `data_ptr()` on the current stream.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .beams import BeamSet

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libabt_diff.so")

ABT_FP64 = 0
ABT_MIXED = 1
ABT_FP64_ACCURATE = 2
ABT_FP64_FAST = 3
ABT_MIXED_GATED = 4
PRECISIONS = {"fp64": ABT_FP64, "mixed": ABT_MIXED, "fp64_accurate": ABT_FP64_ACCURATE,
              "fp64_fast": ABT_FP64_FAST, "mixed_gated": ABT_MIXED_GATED,
              ABT_FP64: ABT_FP64, ABT_MIXED: ABT_MIXED, ABT_FP64_ACCURATE: ABT_FP64_ACCURATE,
              ABT_FP64_FAST: ABT_FP64_FAST, ABT_MIXED_GATED: ABT_MIXED_GATED}
PI_R4 = float(np.float32(3.14159265358979))

# every entry point declared in include/abt_diff.h
EXPORTS = [
    "abt_version", "abt_last_error", "abt_init", "abt_finalize", "abt_make_far_field_bins",
    "abt_diff_main", "abt_diffraction", "abt_finalise", "abt_record_doubles",
    "abt_vertex_record_doubles", "abt_prepare_beams_host", "abt_prepare_beams_dev",
    "abt_grid_create", "abt_grid_destroy", "abt_grid_bins", "abt_far_field_workspace",
    "abt_far_field_dev", "abt_finalise_workspace", "abt_finalise_dev",
    "abt_far_field_incoherent_dev", "abt_finalise_dev_static", "abt_finalise_static_max_bins",
    "abt_batch_plan_create", "abt_batch_plan_destroy",
    "abt_batch_workspace", "abt_batch_set_tiles", "abt_batch_get_tiles", "abt_batch_autotune",
    "abt_batch_run", "abt_batch_capture", "abt_batch_launch", "abt_batch_gate_stats", "abt_batch_set_scaling",
    "abt_adjoint_plan_create",
    "abt_adjoint_plan_destroy", "abt_adjoint_run", "abt_find_vis", "abt_find_vis_workspace",
    "abt_find_vis_dev", "abt_vis_rotations", "abt_vis_rotate_dev",
]


class AbtError(RuntimeError):
    pass


class _Beams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ampl", ctypes.c_void_p), ("vk7", ctypes.c_void_p),
                ("prop", ctypes.c_void_p), ("fov", ctypes.c_void_p), ("vert_off", ctypes.c_void_p),
                ("verts", ctypes.c_void_p)]


class _Grid(ctypes.Structure):
    _fields_ = [("n_theta", ctypes.c_int64), ("n_phi", ctypes.c_int64),
                ("theta", ctypes.c_void_p), ("phi", ctypes.c_void_p)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AbtError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "(make -C abt_amd/csrc); there is no CPU fallback")
    # One HIP runtime per process: when torch is present its libamdhip64 (same soname)
    # must be loaded first so that the library binds to it instead of loading a second
    # runtime that would then hide the GPU from torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, d, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    L.abt_version.restype = i
    L.abt_last_error.restype = ctypes.c_char_p
    L.abt_init.argtypes = [i]
    L.abt_make_far_field_bins.argtypes = [i64, vp, i64, vp, vp, vp, vp]
    L.abt_diff_main.argtypes = [ctypes.POINTER(_Grid), d, i, ctypes.POINTER(_Beams),
                                ctypes.POINTER(_Beams), i, vp, vp]
    L.abt_diffraction.argtypes = [i64, i64, vp, vp, vp, d, vp, vp, i, vp, vp, vp, d, i64, vp, i, vp]
    L.abt_record_doubles.restype = i64
    L.abt_vertex_record_doubles.restype = i64
    L.abt_prepare_beams_host.argtypes = [ctypes.POINTER(_Beams), d, d, vp, vp]
    L.abt_prepare_beams_dev.argtypes = [ctypes.POINTER(_Beams), d, d, vp, vp, vp]
    L.abt_grid_create.argtypes = [ctypes.POINTER(_Grid), d, ctypes.POINTER(vp)]
    L.abt_grid_destroy.argtypes = [vp]
    L.abt_grid_bins.argtypes = [vp]
    L.abt_grid_bins.restype = i64
    L.abt_far_field_workspace.argtypes = [vp, i64, i]
    L.abt_far_field_workspace.restype = ctypes.c_size_t
    L.abt_far_field_dev.argtypes = [vp, vp, vp, i64, i, i, vp, vp, ctypes.c_size_t, vp]
    L.abt_finalise.argtypes = [ctypes.POINTER(_Grid), d, vp, vp, i, d, d, vp, vp, vp, vp, vp,
                               vp, vp]
    L.abt_finalise_workspace.argtypes = [vp]
    L.abt_finalise_workspace.restype = ctypes.c_size_t
    L.abt_finalise_dev.argtypes = [vp, vp, vp, i, d, d, vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.abt_far_field_incoherent_dev.argtypes = [vp, vp, vp, i64, i, vp, vp]
    i32 = ctypes.c_int32
    L.abt_batch_plan_create.argtypes = [vp, i32, vp, vp, i, i, ctypes.POINTER(vp)]
    L.abt_batch_plan_destroy.argtypes = [vp]
    L.abt_batch_workspace.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.abt_batch_set_tiles.argtypes = [vp, i32, i32]
    L.abt_batch_get_tiles.argtypes = [vp] + [ctypes.POINTER(i32)] * 4
    L.abt_batch_autotune.argtypes = [vp, vp, vp, vp, vp, i32, vp, vp]
    L.abt_batch_run.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.abt_batch_capture.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.abt_batch_launch.argtypes = [vp, vp]
    L.abt_batch_gate_stats.argtypes = [vp, ctypes.POINTER(i32), vp]
    L.abt_batch_set_scaling.argtypes = [vp, vp, vp]
    L.abt_adjoint_plan_create.argtypes = [vp, i32, vp, vp, i, ctypes.POINTER(vp)]
    L.abt_adjoint_plan_destroy.argtypes = [vp]
    L.abt_adjoint_run.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = load().abt_last_error().decode(errors="replace")
        raise AbtError(f"abt error {rc}: {msg}")


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def beams_struct(b: BeamSet, fov=True, keep=None):
    s = _Beams(b.n, _ptr(b.ampl).value, _ptr(b.vk7).value, _ptr(b.prop).value,
               _ptr(b.fov).value if fov else None, _ptr(b.vert_off).value, _ptr(b.verts).value)
    return s


def grid_struct(theta, phi):
    return _Grid(len(theta), len(phi), _ptr(theta).value, _ptr(phi).value)


def raw_to_complex(raw, nphi, nth):
    """Fortran-layout (phi, theta, 2, 2) doubles -> complex [r, c, theta, phi]."""
    a = np.asarray(raw).reshape(2, 2, nth, nphi, 2)
    return (a[..., 0] + 1j * a[..., 1]).transpose(1, 0, 2, 3)


# ---------------------------------------------------------------------------
# reference interfaces (host arrays)

def make_far_field_bins(theta, phi):
    theta, phi = _f64(theta), _f64(phi)
    n = len(theta) * len(phi)
    x, y, z = np.empty(n), np.empty(n), np.empty(n)
    _check(load().abt_make_far_field_bins(len(theta), _ptr(theta), len(phi), _ptr(phi), _ptr(x),
                                          _ptr(y), _ptr(z)))
    return x, y, z


def diff_main(theta, phi, lam, beams: BeamSet, ext: BeamSet, precision="fp64", fast_diff=False,
              raw=False, trees=("beam", "ext")):
    """diff_main (diff_mod.f90:674-825): (ampl_far_beam, ampl_far_ext_diff).  trees: the trees to
    compute; an output left out is None (the C-ABI's NULL output)."""
    theta, phi = _f64(theta), _f64(phi)
    nb = len(theta) * len(phi)
    ob = np.empty(8 * nb) if "beam" in trees else None
    oe = np.empty(8 * nb) if "ext" in trees else None
    g = grid_struct(theta, phi)
    _check(load().abt_diff_main(ctypes.byref(g), lam, int(fast_diff),
                                ctypes.byref(beams_struct(beams)), ctypes.byref(beams_struct(ext)),
                                PRECISIONS[precision], _ptr(ob) if ob is not None else None,
                                _ptr(oe) if oe is not None else None))
    if raw:
        return ob, oe
    conv = lambda o: None if o is None else raw_to_complex(o, len(phi), len(theta))
    return conv(ob), conv(oe)


def diffraction(theta, phi, x, y, z, lam, beams: BeamSet, b, precision="fp64", fast_diff=False):
    theta, phi = _f64(theta), _f64(phi)
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.empty(8 * len(theta) * len(phi))
    o0, o1 = beams.vert_off[b], beams.vert_off[b + 1]
    verts = _f64(beams.verts[o0:o1])
    _check(load().abt_diffraction(len(phi), len(theta), _ptr(x), _ptr(y), _ptr(z), lam, _ptr(phi),
                                  _ptr(theta), int(fast_diff), _ptr(_f64(beams.ampl[b])),
                                  _ptr(_f64(beams.vk7[b])), _ptr(_f64(beams.prop[b])),
                                  float(beams.fov[b]), int(o1 - o0), _ptr(verts),
                                  PRECISIONS[precision], _ptr(out)))
    return raw_to_complex(out, len(phi), len(theta))


def prepare_beams_host(beams: BeamSet, lam, fov_value=None):
    """Host set-up of the per-beam records: (recs (n, REC), vrecs (nvert, VREC))."""
    L = load()
    rec, vrec = L.abt_record_doubles(), L.abt_vertex_record_doubles()
    recs = np.zeros((beams.n, rec))
    vrecs = np.zeros((max(beams.verts.shape[0], 1), vrec))
    s = beams_struct(beams, fov=fov_value is None)
    _check(L.abt_prepare_beams_host(ctypes.byref(s), 0.0 if fov_value is None else fov_value, lam,
                                    _ptr(recs), _ptr(vrecs)))
    return recs, vrecs


# ---------------------------------------------------------------------------
# device-resident interface

class DeviceGrid:
    """Far-field bins resident on the device (abt_grid_create)."""

    def __init__(self, theta, phi, lam):
        self.theta, self.phi = _f64(theta), _f64(phi)
        self.lam = float(lam)
        h = ctypes.c_void_p()
        g = grid_struct(self.theta, self.phi)
        _check(load().abt_grid_create(ctypes.byref(g), self.lam, ctypes.byref(h)))
        self.handle = h
        self.nbins = int(load().abt_grid_bins(h))

    def close(self):
        if self.handle is not None and self.handle.value:
            load().abt_grid_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        except Exception:
            return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def prepare_beams_dev(dev_beams: dict, n, lam, recs, vrecs=None, fov_value=None, stream=None):
    """dev_beams: dict of device tensors ampl, vk7, prop, (fov), vert_off, verts."""
    s = _Beams(n, dev_beams["ampl"].data_ptr(), dev_beams["vk7"].data_ptr(),
               dev_beams["prop"].data_ptr(),
               dev_beams["fov"].data_ptr() if (fov_value is None and "fov" in dev_beams) else None,
               dev_beams["vert_off"].data_ptr(), dev_beams["verts"].data_ptr())
    _check(load().abt_prepare_beams_dev(ctypes.byref(s), 0.0 if fov_value is None else fov_value,
                                        lam, _ptr(recs), _ptr(vrecs), _stream_ptr(stream)))


def far_field_workspace(grid: DeviceGrid, nb, precision="fp64"):
    return int(load().abt_far_field_workspace(grid.handle, nb, PRECISIONS[precision]))


def far_field_dev(grid: DeviceGrid, recs, nb, out, vrecs=None, precision="fp64", fast_diff=False,
                  work=None, stream=None):
    wb = 0 if work is None else work.numel() * work.element_size()
    _check(load().abt_far_field_dev(grid.handle, _ptr(recs), _ptr(vrecs), nb, PRECISIONS[precision],
                                    int(fast_diff), _ptr(out), _ptr(work), wb, _stream_ptr(stream)))


def finalise(theta, phi, lam, raw_beam, raw_ext, scaling=False, beam_energy_out=1.0,
             ext_energy_out=1.0):
    """finalise (outputs_mod.f90:63-278) on the device; raw_* are Fortran-layout
    amplitude buffers (copied, then scaled by k like the reference)."""
    theta, phi = _f64(theta), _f64(phi)
    nth, nph = len(theta), len(phi)
    nb = nth * nph
    rb = np.array(raw_beam, dtype=np.float64, copy=True)
    re = np.array(raw_ext, dtype=np.float64, copy=True)
    m = [np.empty(16 * nb) for _ in range(3)]
    m1 = [np.empty(16 * nth) for _ in range(3)]
    params = np.zeros(9)
    g = grid_struct(theta, phi)
    _check(load().abt_finalise(ctypes.byref(g), lam, _ptr(rb), _ptr(re), int(scaling),
                               beam_energy_out, ext_energy_out, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]),
                               _ptr(m1[0]), _ptr(m1[1]), _ptr(m1[2]), _ptr(params)))
    keys = ["scatt", "scatt_beam", "scatt_ext_diff", "asymmetry", "asymmetry_beam",
            "asymmetry_ext_diff", "back_scatt", "scale_beam", "scale_ext"]
    return dict(mueller=m[0].reshape(16, nb), mueller_beam=m[1].reshape(16, nb),
                mueller_ext=m[2].reshape(16, nb), mueller_1d=m1[0].reshape(16, nth),
                mueller_1d_beam=m1[1].reshape(16, nth), mueller_1d_ext=m1[2].reshape(16, nth),
                ampl_beam=rb, ampl_ext=re, **dict(zip(keys, params)))


def finalise_dev(grid: DeviceGrid, ab, ae, mueller, m1d, params, scaling=False,
                 beam_energy_out=1.0, ext_energy_out=1.0, work=None, stream=None):
    wb = 0 if work is None else work.numel() * work.element_size()
    _check(load().abt_finalise_dev(grid.handle, _ptr(ab), _ptr(ae), int(scaling), beam_energy_out,
                                   ext_energy_out, _ptr(mueller), _ptr(m1d), _ptr(params),
                                   _ptr(work), wb, _stream_ptr(stream)))


def far_field_incoherent_dev(grid: DeviceGrid, recs, nb, out, vrecs=None, fast_diff=False,
                             stream=None):
    """Sum over beams of |amplC_b|_F^2 per bin (FP64), into out[nbins]."""
    _check(load().abt_far_field_incoherent_dev(grid.handle, _ptr(recs), _ptr(vrecs), nb,
                                               int(fast_diff), _ptr(out), _stream_ptr(stream)))


# ---------------------------------------------------------------------------
# orientation-batched far field

class BatchPlan:
    """Orientation-averaged far field of many orientations in one launch (abt_batch_*).

    beam_off / ext_off: norient+1 offsets of each orientation's records in the
    beam-tree and ext-tree record arrays.  Outputs of run(): mueller (16*nbins,
    summed over orientations), mueller_1d (16*ntheta, optional) and integrals
    (6*norient: scatt, I2, scatt_beam, I2_beam, scatt_ext, I2_ext; asymmetry = I2/scatt).
    precision: "fp64_fast" (default; meets the 1e-5 floored Mueller accuracy gate), "fp64"
    (the reference's formula), "mixed" (FP32 body; faster, misses the gate), "mixed_gated",
    "fp64_accurate"."""

    TILE_SHAPES = [(tb, bpt) for bpt in (1, 2) for tb in (8, 16, 32)]

    def __init__(self, grid: DeviceGrid, beam_off, ext_off, precision="fp64_fast", fast_diff=False):
        self.grid = grid
        self.beam_off = np.ascontiguousarray(beam_off, dtype=np.int64)
        self.ext_off = np.ascontiguousarray(ext_off, dtype=np.int64)
        self.norient = len(self.beam_off) - 1
        if len(self.ext_off) != self.norient + 1:
            raise AbtError("beam_off and ext_off must have the same length")
        h = ctypes.c_void_p()
        _check(load().abt_batch_plan_create(grid.handle, self.norient, _ptr(self.beam_off),
                                            _ptr(self.ext_off), PRECISIONS[precision],
                                            int(fast_diff), ctypes.byref(h)))
        self.handle = h

    def workspace_bytes(self):
        n = ctypes.c_size_t()
        _check(load().abt_batch_workspace(self.handle, ctypes.byref(n)))
        return n.value

    def set_tiles(self, beams_per_tile, bins_per_thread):
        _check(load().abt_batch_set_tiles(self.handle, beams_per_tile, bins_per_thread))

    def tiles(self):
        v = [ctypes.c_int32() for _ in range(4)]
        _check(load().abt_batch_get_tiles(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("beams_per_tile", "bins_per_thread", "chunks", "workgroups"),
                        (x.value for x in v)))

    def autotune(self, recs_b, recs_e, vrecs_b=None, vrecs_e=None, sample=0, stream=None):
        """Time every tile shape on the first `sample` orientations; keep the fastest.
        Returns {(beams_per_tile, bins_per_thread): ms}."""
        ms = np.zeros(6)
        _check(load().abt_batch_autotune(self.handle, _ptr(recs_b), _ptr(vrecs_b), _ptr(recs_e),
                                         _ptr(vrecs_e), int(sample), _stream_ptr(stream), _ptr(ms)))
        return dict(zip(self.TILE_SHAPES, ms.tolist()))

    def run(self, recs_b, recs_e, mueller, integrals, mueller_1d=None, vrecs_b=None, vrecs_e=None,
            ampl=None, stream=None):
        """ampl (optional, 8*nbins*norient): per-orientation total amplitudes x k."""
        _check(load().abt_batch_run(self.handle, _ptr(recs_b), _ptr(vrecs_b), _ptr(recs_e),
                                    _ptr(vrecs_e), _ptr(mueller), _ptr(mueller_1d), _ptr(integrals),
                                    _ptr(ampl), _stream_ptr(stream)))

    def capture(self, recs_b, recs_e, mueller, integrals, mueller_1d=None, vrecs_b=None,
                vrecs_e=None, ampl=None, stream=None):
        """Capture run() with these buffers into a HIP graph held by the plan."""
        _check(load().abt_batch_capture(self.handle, _ptr(recs_b), _ptr(vrecs_b), _ptr(recs_e),
                                        _ptr(vrecs_e), _ptr(mueller), _ptr(mueller_1d),
                                        _ptr(integrals), _ptr(ampl), _stream_ptr(stream)))

    def launch(self, stream=None):
        """Replay the captured graph."""
        _check(load().abt_batch_launch(self.handle, _stream_ptr(stream)))

    def set_scaling(self, beam_energy_out=None, ext_energy_out=None):
        """-scaling (outputs_mod.f90:146-184): float64 device tensors of norient energies,
        kept referenced by the plan and read by every run; None, None turns it off."""
        if (beam_energy_out is None) != (ext_energy_out is None):
            raise ValueError("give both energy arrays, or neither")
        if beam_energy_out is not None:
            for t in (beam_energy_out, ext_energy_out):
                if not t.is_cuda or str(t.dtype) != "torch.float64" or t.numel() != self.norient \
                        or not t.is_contiguous():
                    raise ValueError("energies: contiguous float64 device tensors of norient values")
            self._energies = (beam_energy_out, ext_energy_out)
            _check(load().abt_batch_set_scaling(self.handle, beam_energy_out.data_ptr(),
                                                 ext_energy_out.data_ptr()))
        else:
            _check(load().abt_batch_set_scaling(self.handle, None, None))
            self._energies = None

    def refined_bins(self, stream=None):
        """mixed_gated plans: bins the last run recomputed in FP64 (synchronises)."""
        n = ctypes.c_int32()
        _check(load().abt_batch_gate_stats(self.handle, ctypes.byref(n), _stream_ptr(stream)))
        return n.value

    def close(self):
        if self.handle is not None and self.handle.value:
            load().abt_batch_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def integrals_to_params(integrals):
    """Per-orientation scattering cross sections and asymmetry parameters from
    BatchPlan integrals (outputs_mod.f90:133-135, 198-200)."""
    I = np.asarray(integrals).reshape(-1, 6)
    return {"scatt": I[:, 0], "asymmetry": I[:, 1] / I[:, 0],
            "scatt_beam": I[:, 2], "asymmetry_beam": I[:, 3] / I[:, 2],
            "scatt_ext_diff": I[:, 4], "asymmetry_ext_diff": I[:, 5] / I[:, 4]}


class AdjointPlan:
    """Gradients of a loss of the batched forward's per-orientation amplitudes
    (abt_adjoint_*): run(recs_b, recs_e, dldx) -> per-beam vertex (9) and amplitude (8)
    gradients of both trees."""

    def __init__(self, grid: DeviceGrid, beam_off, ext_off, precision="mixed"):
        self.grid = grid
        self.beam_off = np.ascontiguousarray(beam_off, dtype=np.int64)
        self.ext_off = np.ascontiguousarray(ext_off, dtype=np.int64)
        self.norient = len(self.beam_off) - 1
        h = ctypes.c_void_p()
        _check(load().abt_adjoint_plan_create(grid.handle, self.norient, _ptr(self.beam_off),
                                              _ptr(self.ext_off), PRECISIONS[precision],
                                              ctypes.byref(h)))
        self.handle = h

    def run(self, recs_b, recs_e, dldx, gv_b=None, ga_b=None, gv_e=None, ga_e=None, stream=None):
        _check(load().abt_adjoint_run(self.handle, _ptr(recs_b), _ptr(recs_e), _ptr(dldx), _ptr(gv_b),
                                      _ptr(ga_b), _ptr(gv_e), _ptr(ga_e), _stream_ptr(stream)))

    def close(self):
        if self.handle is not None and self.handle.value:
            load().abt_adjoint_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
