"""``AcousticCoEvolver`` — the instrument co-evolver on the MI355X engine.

Mirrors the reference's pybind11 class ``_spectral_kernels.AcousticCoEvolver``
(csrc/bindings/pybind_kernels.cpp:5704-6498, registered at :9024-9190):
the constructor takes the same lists of string / pickup / mic dicts and the
body dict, with the same keys and defaults; methods, properties and
``RuntimeError`` messages are the reference's.  Every call goes through the
``coevolver_*`` C ABI of libafdtd.so (include/afdtd_coevolver.h), whose
per-sample loop runs on the device.

Both bodies are provided: the uniform grid (the constructor,
``coevolver_create``) and the AMR stepper (``from_amr_desc``,
``coevolver_create_amr``, with the reference's descriptor dict).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib

PICKUP_SINGLE_COIL, PICKUP_HUMBUCKER, PICKUP_PIEZO = 0, 1, 2
CE_OK, CE_ERR_NULL, CE_ERR_DIM, CE_ERR_UNSTABLE, CE_ERR_PARAM = 0, -1, -2, -3, -4
CE_ERR_STRING_IDX, CE_ERR_BUSY, CE_ERR_FDTD = -5, -6, -7
CE_ERR_FDTD_BRIDGE, CE_ERR_FDTD_STEP, CE_ERR_FDTD_MIC_P, CE_ERR_FDTD_MIC_V = -8, -9, -10, -11
_STATUS = {0: "idle", 1: "running", 2: "done", 3: "cancelled", 4: "error"}


class CoEvolverStringDef(C.Structure):          # acoustic_coevolver.h:175-199
    _fields_ = [("n_segs", C.c_int), ("path_xyz", C.c_void_p), ("tension_N", C.c_float),
                ("linear_mass_kgm", C.c_float), ("damping", C.c_float),
                ("stiffness_EI", C.c_float), ("axial_stiffness_N", C.c_float),
                ("neck_freq_hz", C.c_float), ("neck_mass_kg", C.c_float), ("neck_Q", C.c_float)]


class CoEvolverPickupDef(C.Structure):          # :217-225
    _fields_ = [("type", C.c_int), ("pos", C.c_float * 3), ("axis", C.c_float * 3),
                ("pole_sigma", C.c_float), ("coil_spacing", C.c_float),
                ("sensitivity", C.c_float), ("string_mask", C.c_uint32)]


class CoEvolverMicDef(C.Structure):             # :246-252
    _fields_ = [("pos", C.c_float * 3), ("axis", C.c_float * 3), ("polar_a", C.c_float),
                ("polar_b", C.c_float), ("gain", C.c_float)]


class CoEvolverBodyDef(C.Structure):            # :264-282
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("dx", C.c_float),
                ("origin", C.c_float * 3), ("c", C.c_float), ("rho_air", C.c_float),
                ("cell_type", C.c_void_p), ("plate_iz", C.c_int), ("plate_active", C.c_void_p),
                ("plate_mass_density", C.c_float), ("plate_stiffness_D", C.c_float),
                ("plate_alpha_M", C.c_float), ("plate_beta_K", C.c_float), ("n_pml", C.c_int),
                ("n_bridge_src", C.c_int), ("bridge_src_xyz", C.c_void_p),
                ("bridge_sigma", C.c_float)]


class CoEvolverProgress(C.Structure):           # :146-151
    _fields_ = [("samples_done", C.c_int), ("samples_total", C.c_int), ("status", C.c_int),
                ("error_code", C.c_int)]


_CE_MSG = {                                      # pybind_kernels.cpp:5668-5702
    CE_ERR_NULL: "null pointer (CE_ERR_NULL -1)",
    CE_ERR_DIM: "invalid dimension (CE_ERR_DIM -2)",
    CE_ERR_UNSTABLE: "body FDTD diverged — CFL violated or drive too large (CE_ERR_UNSTABLE -3)",
    CE_ERR_PARAM: "invalid parameter (CE_ERR_PARAM -4)",
    CE_ERR_STRING_IDX: "string index out of range (CE_ERR_STRING_IDX -5)",
    CE_ERR_BUSY: "async job is already running (CE_ERR_BUSY -6)",
    CE_ERR_FDTD: "FDTD subsystem error (CE_ERR_FDTD -7)",
    CE_ERR_FDTD_BRIDGE: ("FDTD bridge injection failed — fdtd_inject_bridge_drive returned "
                         "non-OK; check that st->fdtd is initialised and bridge cell count "
                         "matches (CE_ERR_FDTD_BRIDGE -8)"),
    CE_ERR_FDTD_STEP: ("fdtd_step returned a non-unstable error — likely NULL FDTD state or "
                       "bad grid dims (CE_ERR_FDTD_STEP -9)"),
    CE_ERR_FDTD_MIC_P: ("FDTD pressure sampling failed — check mic index/weight arrays and "
                        "FDTD grid size (CE_ERR_FDTD_MIC_P -10)"),
    CE_ERR_FDTD_MIC_V: ("FDTD velocity sampling failed — check mic velocity index/weight "
                        "arrays (CE_ERR_FDTD_MIC_V -11)"),
}


def _ce_throw(rc: int, ctx: str = ""):
    prefix = f"{ctx}: " if ctx else ""
    raise RuntimeError(prefix + _CE_MSG.get(rc, f"unknown error code {rc}"))


def _v(a):
    return a.ctypes.data_as(C.c_void_p)


def _f3(x):
    a = np.asarray(x, dtype=np.float32).ravel()
    return (C.c_float * 3)(float(a[0]), float(a[1]), float(a[2]))


def build_defs(string_defs, pickup_defs, mic_defs, body):
    """The descriptor structs the pybind constructor builds from its dicts
    (pybind_kernels.cpp:5756-5875), and the arrays they point into (keep
    them alive while the structs are used)."""
    keep = []
    ns = len(string_defs)
    sd = (CoEvolverStringDef * max(ns, 1))()
    for i, d in enumerate(string_defs):
        path = np.ascontiguousarray(np.asarray(d["path_xyz"], dtype=np.float32).ravel())
        keep.append(path)
        sd[i].path_xyz = _v(path)
        sd[i].n_segs = path.size // 3 - 1
        sd[i].tension_N = float(d["tension_N"])
        sd[i].linear_mass_kgm = float(d["linear_mass_kgm"])
        sd[i].damping = float(d["damping"])
        sd[i].stiffness_EI = float(d.get("stiffness_EI", 0.0))
        sd[i].axial_stiffness_N = float(d.get("axial_stiffness_N", 0.0))
        sd[i].neck_freq_hz = float(d.get("neck_freq_hz", 0.0))
        sd[i].neck_mass_kg = float(d.get("neck_mass_kg", 0.0))
        sd[i].neck_Q = float(d.get("neck_Q", 0.0))
    npk = len(pickup_defs)
    pd = (CoEvolverPickupDef * max(npk, 1))()
    for i, d in enumerate(pickup_defs):
        pd[i].type = int(d["type"])
        pd[i].pole_sigma = float(d["pole_sigma"])
        pd[i].coil_spacing = float(d.get("coil_spacing", 0.018))
        pd[i].sensitivity = float(d.get("sensitivity", 1.0))
        pd[i].string_mask = int(d.get("string_mask", -1)) & 0xFFFFFFFF
        pd[i].pos = _f3(d["pos"])
        pd[i].axis = _f3(d["axis"])
    nm = len(mic_defs)
    md = (CoEvolverMicDef * max(nm, 1))()
    for i, d in enumerate(mic_defs):
        md[i].pos = _f3(d["pos"])
        md[i].gain = float(d.get("gain", 1.0))
        md[i].polar_a = float(d.get("polar_a", 1.0))
        md[i].polar_b = float(d.get("polar_b", 0.0))
        md[i].axis = _f3(d["axis"]) if "axis" in d else _f3((0.0, 0.0, 1.0))
    ct = np.ascontiguousarray(np.asarray(body["cell_type"], dtype=np.uint8).ravel())
    pa = np.ascontiguousarray(np.asarray(body["plate_active"], dtype=np.uint8).ravel())
    bs = np.ascontiguousarray(np.asarray(body["bridge_src_xyz"], dtype=np.float32).ravel())
    keep += [ct, pa, bs]
    b = CoEvolverBodyDef()
    b.Nx, b.Ny, b.Nz = int(body["Nx"]), int(body["Ny"]), int(body["Nz"])
    b.dx = float(body["dx"])
    b.origin = _f3((body.get("gx_min", 0.0), body.get("gy_min", 0.0), body.get("gz_min", 0.0)))
    b.c = float(body.get("c", 343.0))
    b.rho_air = float(body.get("rho_air", 1.21))
    b.cell_type = _v(ct)
    b.plate_iz = int(body["plate_iz"])
    b.plate_active = _v(pa)
    b.plate_mass_density = float(body.get("plate_mass_density", 7.0))
    b.plate_stiffness_D = float(body.get("plate_stiffness_D", 0.45))
    b.plate_alpha_M = float(body.get("plate_alpha_M", 2.0))
    b.plate_beta_K = float(body.get("plate_beta_K", 1e-5))
    b.n_pml = int(body.get("n_pml", 10))
    b.n_bridge_src = bs.size // 3
    b.bridge_src_xyz = _v(bs)
    b.bridge_sigma = float(body.get("bridge_sigma", 0.01))
    # the pybind layer hands NULL for an empty list (sdefs.empty() ? nullptr)
    ptrs = (C.cast(sd, C.c_void_p) if ns else None, C.cast(pd, C.c_void_p) if npk else None,
            C.cast(md, C.c_void_p) if nm else None)
    return (ns, npk, nm), ptrs, C.pointer(b), keep + [sd, pd, md, b]


class AMRCoevolverDescriptor(C.Structure):      # acoustic_coevolver.h:331-397
    _fields_ = [
        ("n_cells", C.c_int), ("cell_centers", C.c_void_p), ("cell_volumes", C.c_void_p),
        ("open_volume_frac", C.c_void_p), ("cell_types", C.c_void_p),
        ("cell_levels", C.c_void_p), ("n_faces", C.c_int), ("face_cell_neg", C.c_void_p),
        ("face_cell_pos", C.c_void_p), ("face_area", C.c_void_p),
        ("face_open_frac", C.c_void_p), ("face_distance", C.c_void_p), ("c", C.c_double),
        ("rho_air", C.c_double), ("min_dx", C.c_double), ("bounds_min", C.c_double * 3),
        ("bounds_max", C.c_double * 3), ("plate_Nx", C.c_int), ("plate_Ny", C.c_int),
        ("plate_dx", C.c_float), ("plate_origin", C.c_float * 3), ("plate_active", C.c_void_p),
        ("plate_mass_density", C.c_float), ("plate_stiffness_D", C.c_float),
        ("plate_alpha_M", C.c_float), ("plate_beta_K", C.c_float), ("n_active_plate", C.c_int),
        ("plate_active_flat_idx", C.c_void_p), ("plate_face_above_starts", C.c_void_p),
        ("plate_face_above_idx", C.c_void_p), ("plate_face_above_wgt", C.c_void_p),
        ("plate_face_below_starts", C.c_void_p), ("plate_face_below_idx", C.c_void_p),
        ("plate_face_below_wgt", C.c_void_p), ("plate_cell_above", C.c_void_p),
        ("plate_cell_below", C.c_void_p), ("n_bridge_plate", C.c_int),
        ("bridge_plate_idx", C.c_void_p), ("bridge_plate_wgt", C.c_void_p),
        ("n_neck_plate", C.c_int), ("neck_plate_idx", C.c_void_p), ("neck_plate_wgt", C.c_void_p),
        ("soundhole_cx", C.c_double), ("soundhole_cy", C.c_double),
        ("soundhole_radius", C.c_double), ("border_mode", C.c_int),
        ("border_sigma_order", C.c_float), ("border_R_reflection", C.c_float),
        ("border_Z_match", C.c_float), ("n_pml", C.c_int), ("gradient_order", C.c_int)]


# (key, dtype) of the descriptor's arrays as the pybind factory reads them
# (pybind_kernels.cpp:5961-5988)
_AMR_ARRAYS = (
    ("cell_centers", np.float64), ("cell_volumes", np.float64),
    ("open_volume_frac", np.float64), ("cell_types", np.uint8), ("cell_levels", np.int16),
    ("face_cell_neg", np.int32), ("face_cell_pos", np.int32), ("face_area", np.float64),
    ("face_open_frac", np.float64), ("face_distance", np.float64),
    ("plate_active", np.uint8), ("plate_active_flat_idx", np.int32),
    ("plate_face_above_starts", np.int32), ("plate_face_above_idx", np.int32),
    ("plate_face_above_wgt", np.float32), ("plate_face_below_starts", np.int32),
    ("plate_face_below_idx", np.int32), ("plate_face_below_wgt", np.float32),
    ("plate_cell_above", np.int32), ("plate_cell_below", np.int32),
    ("bridge_plate_idx", np.int32), ("bridge_plate_wgt", np.float32),
    ("neck_plate_idx", np.int32), ("neck_plate_wgt", np.float32))


def build_amr_desc(amr_dict):
    """The AMRCoevolverDescriptor the pybind factory fills from the dict of
    ``build_amr_coevolver_descriptor()`` (pybind_kernels.cpp:5959-6037): the
    same keys, dtypes and optional-key defaults.  Returns (struct, keep)."""
    keep = []
    d = AMRCoevolverDescriptor()
    for key, dt in _AMR_ARRAYS:
        a = np.ascontiguousarray(np.asarray(amr_dict[key], dtype=dt).ravel())
        keep.append(a)
        setattr(d, key, _v(a))
    d.n_cells = int(amr_dict["n_cells"])
    d.n_faces = int(amr_dict["n_faces"])
    d.c = float(amr_dict["c"])
    d.rho_air = float(amr_dict["rho_air"])
    d.min_dx = float(amr_dict["min_dx"])
    bmin = np.asarray(amr_dict["bounds_min"], np.float64).ravel()
    bmax = np.asarray(amr_dict["bounds_max"], np.float64).ravel()
    d.bounds_min = (C.c_double * 3)(*[float(x) for x in bmin[:3]])
    d.bounds_max = (C.c_double * 3)(*[float(x) for x in bmax[:3]])
    d.plate_Nx = int(amr_dict["plate_Nx"])
    d.plate_Ny = int(amr_dict["plate_Ny"])
    d.plate_dx = float(amr_dict["plate_dx"])
    d.plate_origin = _f3(amr_dict["plate_origin"])
    d.plate_mass_density = float(amr_dict["plate_mass_density"])
    d.plate_stiffness_D = float(amr_dict["plate_stiffness_D"])
    d.plate_alpha_M = float(amr_dict["plate_alpha_M"])
    d.plate_beta_K = float(amr_dict["plate_beta_K"])
    d.n_active_plate = int(amr_dict["n_active_plate"])
    d.n_bridge_plate = int(amr_dict["n_bridge_plate"])
    d.n_neck_plate = int(amr_dict["n_neck_plate"])
    d.soundhole_cx = float(amr_dict["soundhole_cx"])
    d.soundhole_cy = float(amr_dict["soundhole_cy"])
    d.soundhole_radius = float(amr_dict["soundhole_radius"])
    d.border_mode = int(amr_dict.get("border_mode", 0))
    d.border_sigma_order = float(amr_dict.get("border_sigma_order", 3.0))
    d.border_R_reflection = float(amr_dict.get("border_R_reflection", 0.0))
    d.border_Z_match = float(amr_dict.get("border_Z_match", 0.0))
    d.n_pml = int(amr_dict.get("n_pml", 0))
    d.gradient_order = int(amr_dict["gradient_order"])
    keep.append(d)
    return d, keep


class AcousticCoEvolver:
    """Strings, bridge / neck impedances, pickups and microphones coupled to the
    body FDTD, advanced per audio sample on the device."""

    def __init__(self, string_defs, pickup_defs, mic_defs, body, sample_rate: float,
                 modal_stride: int = 16, force_scale: float = 1.0):
        self._L = lib()
        self._h = None
        (ns, npk, nm), (sp, pp, mp), bp, self._keep = build_defs(
            list(string_defs), list(pickup_defs), list(mic_defs), body)
        self._n_strings, self._n_pickups, self._n_mics = ns, npk, nm
        self._n_nodes = [np.asarray(d["path_xyz"]).size // 3 for d in string_defs]
        self._Nx, self._Ny, self._Nz = int(body["Nx"]), int(body["Ny"]), int(body["Nz"])
        self._n_cells_amr = 0
        h = self._L.coevolver_create(ns, sp, npk, pp, nm, mp, C.cast(bp, C.c_void_p),
                                     float(sample_rate), int(modal_stride), float(force_scale))
        if not h:
            raise RuntimeError("coevolver_create: allocation failed")
        self._h = h

    @classmethod
    def from_amr_desc(cls, amr_desc, string_defs, pickup_defs, mic_defs, sample_rate: float,
                      modal_stride: int = 16):
        """AMR-backed co-evolver (pybind_kernels.cpp:5887-6098): the body is the
        AMR stepper built from ``amr_desc``; get_pressure_field scatters the
        cell pressures onto a box grid of spacing max(4 min_dx, 5 mm)."""
        self = cls.__new__(cls)
        self._L = lib()
        self._h = None
        dummy = {"cell_type": np.zeros(1, np.uint8), "plate_active": np.zeros(1, np.uint8),
                 "bridge_src_xyz": np.zeros(0, np.float32), "Nx": 1, "Ny": 1, "Nz": 1,
                 "dx": 1.0, "plate_iz": 0}
        (ns, npk, nm), (sp, pp, mp), _bp, keep = build_defs(
            list(string_defs), list(pickup_defs), list(mic_defs), dummy)
        d, dkeep = build_amr_desc(amr_desc)
        self._keep = keep + dkeep
        self._n_strings, self._n_pickups, self._n_mics = ns, npk, nm
        self._n_nodes = [np.asarray(x["path_xyz"]).size // 3 for x in string_defs]
        h = self._L.coevolver_create_amr(ns, sp, npk, pp, nm, mp, C.cast(C.pointer(d), C.c_void_p),
                                         float(sample_rate), int(modal_stride))
        if not h:
            code = (self._L.coevolver_amr_backend_error_code() or b"").decode()
            msg = (self._L.coevolver_amr_backend_error_message() or b"").decode()
            stage = (self._L.coevolver_create_amr_last_error() or b"").decode()
            detail = "coevolver_create_amr failed"
            if code:
                detail += f" [{code}]"
            if msg:
                detail += f": {msg}"
            if stage:
                detail += f" (stage: {stage})"
            raise RuntimeError(detail)
        self._h = h
        viz = max(float(d.min_dx) * 4.0, 0.005)
        n = [max(1, min(int(round((d.bounds_max[q] - d.bounds_min[q]) / viz)) + 1, 256))
             for q in range(3)]
        self._Nx, self._Ny, self._Nz = n
        self._n_cells_amr = int(d.n_cells)
        self._bmin = (C.c_double * 3)(*d.bounds_min)
        self._bmax = (C.c_double * 3)(*d.bounds_max)
        return self

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def close(self):
        if getattr(self, "_h", None):
            self._L.coevolver_destroy(self._h)
            self._h = None

    # ── stepping ────────────────────────────────────────────────────────
    def step(self, n_samples: int = 1) -> int:
        rc = self._L.coevolver_step(self._h, int(n_samples))
        if rc != CE_OK:
            _ce_throw(rc, "step")
        return rc

    def step_block_with_drive(self, drive_blocks, force_pos_norm: float = 0.15,
                              n_samples: int = 0) -> np.ndarray:
        arrays = []
        if drive_blocks is not None:
            arrays = [np.ascontiguousarray(np.asarray(b, dtype=np.float32).ravel())
                      for b in drive_blocks]
        n_str = len(arrays)
        ptrs = (C.c_void_p * max(n_str, 1))(*[_v(a) for a in arrays])
        out = np.zeros(int(n_samples), np.float32)
        rc = self._L.coevolver_step_block_with_drive(
            self._h, C.cast(ptrs, C.c_void_p) if n_str else None,
            n_str if n_str > 0 else self._n_strings, int(n_samples), float(force_pos_norm),
            _v(out))
        if rc == CE_ERR_PARAM:
            raise RuntimeError(f"step_block_with_drive: n_strings mismatch — got {n_str}, "
                               f"expected {self._n_strings} (CE_ERR_PARAM -4)")
        if rc != CE_OK:
            _ce_throw(rc, "step_block_with_drive")
        return out

    def step_async(self, n_samples: int):
        rc = self._L.coevolver_step_async(self._h, int(n_samples))
        if rc == CE_ERR_PARAM:
            raise RuntimeError("AcousticCoEvolver: async job already running")
        if rc != CE_OK:
            raise RuntimeError(f"coevolver_step_async error: {rc}")

    def get_progress(self) -> dict:
        p = CoEvolverProgress()
        self._L.coevolver_get_progress(self._h, C.byref(p))
        return {"samples_done": p.samples_done, "samples_total": p.samples_total,
                "pct": 100.0 * p.samples_done / p.samples_total if p.samples_total > 0 else 0.0,
                "status": p.status, "status_str": _STATUS.get(p.status, "unknown"),
                "error_code": p.error_code}

    def wait(self) -> int:
        return self._L.coevolver_wait(self._h)

    def cancel(self):
        self._L.coevolver_cancel(self._h)

    def is_running(self) -> int:
        return self._L.coevolver_is_running(self._h)

    # ── excitation ──────────────────────────────────────────────────────
    def pluck_string(self, string_idx: int, position_norm: float, amplitude: float,
                     n_duration_samples: int = 0):
        rc = self._L.coevolver_pluck_string(self._h, int(string_idx), float(position_norm),
                                            float(amplitude), int(n_duration_samples))
        if rc != CE_OK:
            raise RuntimeError(f"coevolver_pluck_string error: {rc}")

    def schedule_pluck(self, onset_sample: int, string_idx: int, position_norm: float,
                       amplitude: float, n_duration_samples: int = 0):
        rc = self._L.coevolver_schedule_pluck(self._h, int(onset_sample), int(string_idx),
                                              float(position_norm), float(amplitude),
                                              int(n_duration_samples))
        if rc != CE_OK:
            raise RuntimeError(f"coevolver_schedule_pluck error: {rc}")

    def clear_pluck_schedule(self):
        self._L.coevolver_clear_pluck_schedule(self._h)

    def inject_string_force(self, string_idx: int, force_buf, force_pos_norm: float = 0.1):
        f = np.ascontiguousarray(np.asarray(force_buf, dtype=np.float32).ravel())
        rc = self._L.coevolver_inject_string_force(self._h, int(string_idx), f.size, _v(f),
                                                   float(force_pos_norm))
        if rc != CE_OK:
            raise RuntimeError(f"coevolver_inject_string_force error: {rc}")

    # ── outputs ─────────────────────────────────────────────────────────
    def get_pickup_output(self, pickup_idx: int, n_samples: int) -> np.ndarray:
        out = np.zeros(int(n_samples), np.float32)
        rc = self._L.coevolver_get_pickup_output(self._h, int(pickup_idx), _v(out), int(n_samples))
        if rc != CE_OK:
            raise RuntimeError(f"get_pickup_output error: {rc}")
        return out

    def get_mic_output(self, mic_idx: int, n_samples: int) -> np.ndarray:
        out = np.zeros(int(n_samples), np.float32)
        rc = self._L.coevolver_get_mic_output(self._h, int(mic_idx), _v(out), int(n_samples))
        if rc != CE_OK:
            raise RuntimeError(f"get_mic_output error: {rc}")
        return out

    def get_pressure_field(self) -> np.ndarray:
        out = np.zeros((self._Nx, self._Ny, self._Nz), np.float32)
        if self._n_cells_amr > 0:                    # pybind_kernels.cpp:6227-6236
            rc = self._L.coevolver_get_pressure_field_uniform(
                self._h, self._Nx, self._Ny, self._Nz, C.cast(self._bmin, C.c_void_p),
                C.cast(self._bmax, C.c_void_p), _v(out), out.size)
        else:
            rc = self._L.coevolver_get_pressure_field(self._h, _v(out), out.size)
        if rc != CE_OK:
            raise RuntimeError(f"get_pressure_field error: {rc}")
        return out

    def get_amr_cell_centers(self):
        if self._n_cells_amr <= 0:
            raise RuntimeError("get_amr_cell_centers: not an AMR coevolver")
        out = np.zeros((self._n_cells_amr, 3), np.float32)
        rc = self._L.coevolver_get_amr_cell_centers(self._h, _v(out), out.size)
        if rc != CE_OK:
            raise RuntimeError(f"get_amr_cell_centers error: {rc}")
        return out

    def get_amr_pressure(self) -> np.ndarray:
        """Extension: the AMR body's flat cell pressures (the C ABI's
        coevolver_get_pressure_field on an AMR handle)."""
        out = np.zeros(self._n_cells_amr, np.float32)
        rc = self._L.coevolver_get_pressure_field(self._h, _v(out), out.size)
        if rc != CE_OK:
            raise RuntimeError(f"get_pressure_field error: {rc}")
        return out

    def get_plate_dims(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        rc = self._L.coevolver_get_plate_dims(self._h, C.byref(a), C.byref(b), C.byref(c))
        if rc != CE_OK:
            raise RuntimeError(f"get_plate_dims error: {rc}")
        return a.value, b.value, c.value

    def get_plate_displacement(self) -> np.ndarray:
        nx, ny, cnt = self.get_plate_dims()
        if nx <= 0 or ny <= 0 or cnt != nx * ny:
            raise RuntimeError(f"get_plate_dims returned invalid dimensions: {nx}x{ny} count={cnt}")
        out = np.zeros((nx, ny), np.float32)
        rc = self._L.coevolver_get_plate_displacement(self._h, _v(out), cnt)
        if rc != CE_OK:
            raise RuntimeError(f"get_plate_displacement error: {rc}")
        return out

    def get_modal_amplitudes(self, string_idx: int, n_modes: int = 32):
        re = np.zeros(int(n_modes), np.float32)
        im = np.zeros(int(n_modes), np.float32)
        rc = self._L.coevolver_get_modal_amplitudes(self._h, int(string_idx), _v(re), _v(im),
                                                    int(n_modes))
        if rc != CE_OK:
            raise RuntimeError(f"get_modal_amplitudes error: {rc}")
        return re, im

    def sample_pressure(self, rec_xyz) -> np.ndarray:
        r = np.ascontiguousarray(np.asarray(rec_xyz, dtype=np.float32).ravel())
        out = np.zeros(r.size // 3, np.float32)
        rc = self._L.coevolver_sample_pressure(self._h, r.size // 3, _v(r), _v(out))
        if rc != CE_OK:
            raise RuntimeError(f"sample_pressure error: {rc}")
        return out

    def _string_field(self, fn, what, string_idx, n, cols=3):
        out = np.zeros((int(n), cols), np.float32)
        rc = fn(self._h, int(string_idx), _v(out), out.size)
        if rc != CE_OK:
            raise RuntimeError(f"{what} error: {rc}")
        return out

    def get_string_velocity(self, string_idx: int) -> np.ndarray:
        if string_idx < 0 or string_idx >= self._n_strings:
            raise RuntimeError("string index out of range")
        out = np.zeros((512, 3), np.float32)
        rc = self._L.coevolver_get_string_velocity(self._h, int(string_idx), _v(out), out.size)
        if rc == CE_ERR_DIM:
            raise RuntimeError("get_string_velocity: out_len mismatch. "
                               "Pass n_segs*3 explicitly via get_string_velocity_n.")
        if rc != CE_OK:
            raise RuntimeError(f"get_string_velocity error: {rc}")
        return out

    def get_string_velocity_n(self, string_idx: int, n_segs: int) -> np.ndarray:
        return self._string_field(self._L.coevolver_get_string_velocity, "get_string_velocity",
                                  string_idx, n_segs)

    def get_string_displacement_n(self, string_idx: int, n_segs: int) -> np.ndarray:
        return self._string_field(self._L.coevolver_get_string_displacement,
                                  "get_string_displacement", string_idx, n_segs)

    def get_string_position_n(self, string_idx: int, n_nodes: int) -> np.ndarray:
        return self._string_field(self._L.coevolver_get_string_position, "get_string_position",
                                  string_idx, n_nodes)

    def get_surface_emission(self, xyz, normals):
        x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float32).ravel())
        nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float32).ravel())
        n = x.size // 3
        P = np.zeros(n, np.float32)
        vn = np.zeros(n, np.float32)
        rc = self._L.coevolver_get_surface_emission(self._h, n, _v(x), _v(nrm), _v(P), _v(vn))
        if rc != CE_OK:
            raise RuntimeError(f"get_surface_emission error: {rc}")
        return P, vn

    def sample_velocity(self, rec_xyz):
        r = np.ascontiguousarray(np.asarray(rec_xyz, dtype=np.float32).ravel())
        n = r.size // 3
        vx, vy, vz = (np.zeros(n, np.float32) for _ in range(3))
        rc = self._L.coevolver_sample_velocity(self._h, n, _v(r), _v(vx), _v(vy), _v(vz))
        if rc != CE_OK:
            raise RuntimeError(f"sample_velocity error: {rc}")
        return vx, vy, vz

    def reset(self):
        self._L.coevolver_reset(self._h)

    def set_face_fractions(self, vx_frac, vy_frac, vz_frac):
        a = [np.ascontiguousarray(np.asarray(f, dtype=np.float32).ravel())
             for f in (vx_frac, vy_frac, vz_frac)]
        rc = self._L.coevolver_set_face_fractions(self._h, _v(a[0]), _v(a[1]), _v(a[2]))
        if rc != CE_OK:
            raise RuntimeError(f"coevolver_set_face_fractions error: {rc}")

    def set_damping_scale(self, scale: float):
        self._L.coevolver_set_damping_scale(self._h, float(scale))

    @property
    def n_strings(self) -> int:
        return self._L.coevolver_get_n_strings(self._h)

    @property
    def n_pickups(self) -> int:
        return self._L.coevolver_get_n_pickups(self._h)

    @property
    def n_mics(self) -> int:
        return self._L.coevolver_get_n_mics(self._h)

    @property
    def dt_audio(self) -> float:
        return self._L.coevolver_get_dt_audio(self._h)

    @property
    def dt_fdtd(self) -> float:
        return self._L.coevolver_get_dt_fdtd(self._h)

    # ── extensions ──────────────────────────────────────────────────────
    @property
    def substeps(self) -> int:
        return self._L.coevolver_get_substeps(self._h)

    def string_substeps(self, string_idx: int) -> int:
        return self._L.coevolver_get_string_substeps(self._h, int(string_idx))

    @property
    def samples_written(self) -> int:
        return self._L.coevolver_get_samples_written(self._h)

    def step_timed(self, n_samples: int) -> float:
        """coevolver_step with the device time of the whole loop (ms)."""
        ms = C.c_float()
        rc = self._L.coevolver_step_timed(self._h, int(n_samples), C.byref(ms))
        if rc != CE_OK:
            _ce_throw(rc, "step")
        return ms.value

    def string_state(self, string_idx: int) -> dict:
        """Raw state of one string (coevolver_get_string_state)."""
        N = self._n_nodes[string_idx]
        out = np.zeros(4 * N + 6, np.float32)
        rc = self._L.coevolver_get_string_state(self._h, int(string_idx), _v(out), out.size)
        if rc != CE_OK:
            raise RuntimeError(f"string_state error: {rc}")
        return {"u0": out[:N], "u1": out[N:2 * N], "up0": out[2 * N:3 * N],
                "up1": out[3 * N:4 * N], "T_eff": out[4 * N], "saddle_w": out[4 * N + 1],
                "saddle_w_prev": out[4 * N + 2], "bridge_w": out[4 * N + 3],
                "bridge_w_prev": out[4 * N + 4], "n_sub": int(out[4 * N + 5])}

    def neck_state(self) -> np.ndarray:
        out = np.zeros(3, np.float32)
        rc = self._L.coevolver_get_neck_state(self._h, _v(out))
        if rc != CE_OK:
            raise RuntimeError(f"neck_state error: {rc}")
        return out

    def fdtd_handle(self):
        """The owned body FDTD handle (for the fdtd_* ABI)."""
        return self._L.coevolver_get_fdtd(self._h)
