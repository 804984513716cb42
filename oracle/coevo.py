"""ctypes front end of the co-evolver oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, as the checker.  It wraps libcoevo_oracle.so,
a plain-C restatement of the reference's uniform co-evolver
(csrc/kernels/acoustic_coevolver.cpp) built on the float32 FDTD oracle.
The constructor takes the same string / pickup / mic / body dicts as the
reference's pybind class (pybind_kernels.cpp:5756-5875); the descriptor
structs are built here, independently of the product's own marshalling.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcoevo_oracle.so")
_lib = None


class _Str(C.Structure):
    _fields_ = [("n_segs", C.c_int), ("path_xyz", C.c_void_p), ("tension_N", C.c_float),
                ("linear_mass_kgm", C.c_float), ("damping", C.c_float),
                ("stiffness_EI", C.c_float), ("axial_stiffness_N", C.c_float),
                ("neck_freq_hz", C.c_float), ("neck_mass_kg", C.c_float), ("neck_Q", C.c_float)]


class _Pk(C.Structure):
    _fields_ = [("type", C.c_int), ("pos", C.c_float * 3), ("axis", C.c_float * 3),
                ("pole_sigma", C.c_float), ("coil_spacing", C.c_float),
                ("sensitivity", C.c_float), ("string_mask", C.c_uint32)]


class _Mic(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("axis", C.c_float * 3), ("polar_a", C.c_float),
                ("polar_b", C.c_float), ("gain", C.c_float)]


class _Body(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("dx", C.c_float),
                ("origin", C.c_float * 3), ("c", C.c_float), ("rho_air", C.c_float),
                ("cell_type", C.c_void_p), ("plate_iz", C.c_int), ("plate_active", C.c_void_p),
                ("plate_mass_density", C.c_float), ("plate_stiffness_D", C.c_float),
                ("plate_alpha_M", C.c_float), ("plate_beta_K", C.c_float), ("n_pml", C.c_int),
                ("n_bridge_src", C.c_int), ("bridge_src_xyz", C.c_void_p),
                ("bridge_sigma", C.c_float)]


# the AMR body descriptor, field for field (acoustic_coevolver.h:331-397)
_AMR_FIELDS = [
    ("n_cells", "i"), ("cell_centers", "p"), ("cell_volumes", "p"), ("open_volume_frac", "p"),
    ("cell_types", "p"), ("cell_levels", "p"), ("n_faces", "i"), ("face_cell_neg", "p"),
    ("face_cell_pos", "p"), ("face_area", "p"), ("face_open_frac", "p"), ("face_distance", "p"),
    ("c", "d"), ("rho_air", "d"), ("min_dx", "d"), ("bounds_min", "d3"), ("bounds_max", "d3"),
    ("plate_Nx", "i"), ("plate_Ny", "i"), ("plate_dx", "f"), ("plate_origin", "f3"),
    ("plate_active", "p"), ("plate_mass_density", "f"), ("plate_stiffness_D", "f"),
    ("plate_alpha_M", "f"), ("plate_beta_K", "f"), ("n_active_plate", "i"),
    ("plate_active_flat_idx", "p"), ("plate_face_above_starts", "p"),
    ("plate_face_above_idx", "p"), ("plate_face_above_wgt", "p"),
    ("plate_face_below_starts", "p"), ("plate_face_below_idx", "p"),
    ("plate_face_below_wgt", "p"), ("plate_cell_above", "p"), ("plate_cell_below", "p"),
    ("n_bridge_plate", "i"), ("bridge_plate_idx", "p"), ("bridge_plate_wgt", "p"),
    ("n_neck_plate", "i"), ("neck_plate_idx", "p"), ("neck_plate_wgt", "p"),
    ("soundhole_cx", "d"), ("soundhole_cy", "d"), ("soundhole_radius", "d"),
    ("border_mode", "i"), ("border_sigma_order", "f"), ("border_R_reflection", "f"),
    ("border_Z_match", "f"), ("n_pml", "i"), ("gradient_order", "i")]
_CT = {"i": C.c_int, "p": C.c_void_p, "d": C.c_double, "f": C.c_float,
       "d3": C.c_double * 3, "f3": C.c_float * 3}
# array fields and their dtypes (pybind_kernels.cpp:5961-5988)
_AMR_ARRAYS = {
    "cell_centers": np.float64, "cell_volumes": np.float64, "open_volume_frac": np.float64,
    "cell_types": np.uint8, "cell_levels": np.int16, "face_cell_neg": np.int32,
    "face_cell_pos": np.int32, "face_area": np.float64, "face_open_frac": np.float64,
    "face_distance": np.float64, "plate_active": np.uint8, "plate_active_flat_idx": np.int32,
    "plate_face_above_starts": np.int32, "plate_face_above_idx": np.int32,
    "plate_face_above_wgt": np.float32, "plate_face_below_starts": np.int32,
    "plate_face_below_idx": np.int32, "plate_face_below_wgt": np.float32,
    "plate_cell_above": np.int32, "plate_cell_below": np.int32, "bridge_plate_idx": np.int32,
    "bridge_plate_wgt": np.float32, "neck_plate_idx": np.int32, "neck_plate_wgt": np.float32}


class _AmrDesc(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t in _AMR_FIELDS]


def _amr_desc(d, keep):
    """The descriptor struct from the reference's amr_dict keys."""
    s = _AmrDesc()
    for n, t in _AMR_FIELDS:
        if t == "p":
            a = np.ascontiguousarray(np.asarray(d[n], _AMR_ARRAYS[n]).ravel())
            keep.append(a)
            setattr(s, n, _v(a))
        elif t in ("d3", "f3"):
            a = np.asarray(d[n], np.float64).ravel()
            setattr(s, n, _CT[t](*[float(x) for x in a[:3]]))
        elif n in ("border_mode", "n_pml"):
            setattr(s, n, int(d.get(n, 0)))
        elif n in ("border_sigma_order",):
            setattr(s, n, float(d.get(n, 3.0)))
        elif n in ("border_R_reflection", "border_Z_match"):
            setattr(s, n, float(d.get(n, 0.0)))
        elif t == "i":
            setattr(s, n, int(d[n]))
        else:
            setattr(s, n, float(d[n]))
    keep.append(s)
    return s


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        vp, i, f = C.c_void_p, C.c_int, C.c_float
        sig = {
            "orc_coevo_create": (vp, [i, vp, i, vp, i, vp, vp, f, i, f]),
            "orc_coevo_destroy": (None, [vp]),
            "orc_coevo_step": (i, [vp, i]),
            "orc_coevo_step_worker": (i, [vp, i]),
            "orc_coevo_step_block_with_drive": (i, [vp, vp, i, i, f, vp]),
            "orc_coevo_pluck": (i, [vp, i, f, f, i]),
            "orc_coevo_inject_force": (i, [vp, i, i, vp, f]),
            "orc_coevo_schedule_pluck": (i, [vp, i, i, f, f, i]),
            "orc_coevo_clear_schedule": (None, [vp]),
            "orc_coevo_reset": (i, [vp]),
            "orc_coevo_set_damping_scale": (None, [vp, f]),
            "orc_coevo_set_face_fractions": (i, [vp, vp, vp, vp]),
            "orc_coevo_get_pickup_output": (i, [vp, i, vp, i]),
            "orc_coevo_get_mic_output": (i, [vp, i, vp, i]),
            "orc_coevo_fdtd": (vp, [vp]),
            "orc_coevo_substeps": (i, [vp]),
            "orc_coevo_written": (i, [vp]),
            "orc_coevo_dt_fdtd": (f, [vp]),
            "orc_coevo_n_bridge": (i, [vp]),
            "orc_coevo_n_neck": (i, [vp]),
            "orc_coevo_string_state": (i, [vp, i, vp, i]),
            "orc_coevo_neck_state": (None, [vp, vp]),
            "orc_coevo_modal": (i, [vp, i, vp, vp, i]),
            "orc_coevo_string_frame": (i, [vp, i, vp, i]),
            "orc_get_fields_f32": (i, [vp, vp, vp, vp, vp]),
            "orc_get_plate_f32": (i, [vp, vp]),
            "orc_get_step_count_f32": (i, [vp]),
            "orc_coevo_set_string": (i, [vp, i, vp, vp, vp, vp]),
            "orc_coevo_string_params": (None, [vp, i, vp]),
            "orc_coevo_tables": (None, [vp] * 8),
            "orc_get_cell_type_f32": (i, [vp, vp]),
            "orc_coevo_create_amr": (vp, [i, vp, i, vp, i, vp, vp, f, i, C.POINTER(C.c_char_p)]),
            "orc_coevo_amr": (vp, [vp]),
            "amo_get_pressure": (i, [vp, vp, i]),
            "amo_get_velocity": (i, [vp, vp, i]),
            "amo_get_plate": (i, [vp, vp, i]),
            "amo_get_step_count": (i, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _v(a):
    return a.ctypes.data_as(C.c_void_p)


def _f3(x):
    a = np.asarray(x, dtype=np.float32).ravel()
    return (C.c_float * 3)(float(a[0]), float(a[1]), float(a[2]))


class OracleCoEvolver:
    """The co-evolver oracle with the reference constructor's dict inputs."""

    def __init__(self, string_defs, pickup_defs, mic_defs, body, sample_rate,
                 modal_stride=16, force_scale=1.0):
        L = self._L = lib()
        self._h = None
        self._amr_desc = None
        self._keep = []
        sd, pd, md = self._defs(string_defs, pickup_defs, mic_defs)
        ns, npk, nm = len(string_defs), len(pickup_defs), len(mic_defs)
        ct = np.ascontiguousarray(np.asarray(body["cell_type"], np.uint8).ravel())
        pa = np.ascontiguousarray(np.asarray(body["plate_active"], np.uint8).ravel())
        bs = np.ascontiguousarray(np.asarray(body["bridge_src_xyz"], np.float32).ravel())
        self._keep += [ct, pa, bs]
        b = _Body()
        b.Nx, b.Ny, b.Nz = body["Nx"], body["Ny"], body["Nz"]
        b.dx = body["dx"]
        b.origin = _f3((body.get("gx_min", 0.0), body.get("gy_min", 0.0), body.get("gz_min", 0.0)))
        b.c = body.get("c", 343.0)
        b.rho_air = body.get("rho_air", 1.21)
        b.cell_type = _v(ct)
        b.plate_iz = body["plate_iz"]
        b.plate_active = _v(pa)
        b.plate_mass_density = body.get("plate_mass_density", 7.0)
        b.plate_stiffness_D = body.get("plate_stiffness_D", 0.45)
        b.plate_alpha_M = body.get("plate_alpha_M", 2.0)
        b.plate_beta_K = body.get("plate_beta_K", 1e-5)
        b.n_pml = body.get("n_pml", 10)
        b.n_bridge_src = bs.size // 3
        b.bridge_src_xyz = _v(bs)
        b.bridge_sigma = body.get("bridge_sigma", 0.01)
        self._keep.append(b)
        self.Nx, self.Ny, self.Nz = b.Nx, b.Ny, b.Nz
        self.n_strings, self.n_pickups, self.n_mics = ns, npk, nm
        h = L.orc_coevo_create(ns, C.cast(sd, C.c_void_p) if ns else None, npk,
                               C.cast(pd, C.c_void_p) if npk else None, nm,
                               C.cast(md, C.c_void_p) if nm else None,
                               C.cast(C.pointer(b), C.c_void_p), float(sample_rate),
                               int(modal_stride), float(force_scale))
        if not h:
            raise RuntimeError("orc_coevo_create failed")
        self._h = h

    def _defs(self, string_defs, pickup_defs, mic_defs):
        ns, npk, nm = len(string_defs), len(pickup_defs), len(mic_defs)
        sd = (_Str * max(ns, 1))()
        self.n_nodes = []
        for q, d in enumerate(string_defs):
            p = np.ascontiguousarray(np.asarray(d["path_xyz"], np.float32).ravel())
            self._keep.append(p)
            sd[q].n_segs = p.size // 3 - 1
            self.n_nodes.append(p.size // 3)
            sd[q].path_xyz = _v(p)
            sd[q].tension_N = d["tension_N"]
            sd[q].linear_mass_kgm = d["linear_mass_kgm"]
            sd[q].damping = d["damping"]
            sd[q].stiffness_EI = d.get("stiffness_EI", 0.0)
            sd[q].axial_stiffness_N = d.get("axial_stiffness_N", 0.0)
            sd[q].neck_freq_hz = d.get("neck_freq_hz", 0.0)
            sd[q].neck_mass_kg = d.get("neck_mass_kg", 0.0)
            sd[q].neck_Q = d.get("neck_Q", 0.0)
        pd = (_Pk * max(npk, 1))()
        for q, d in enumerate(pickup_defs):
            pd[q].type = d["type"]
            pd[q].pos = _f3(d["pos"])
            pd[q].axis = _f3(d["axis"])
            pd[q].pole_sigma = d["pole_sigma"]
            pd[q].coil_spacing = d.get("coil_spacing", 0.018)
            pd[q].sensitivity = d.get("sensitivity", 1.0)
            pd[q].string_mask = int(d.get("string_mask", -1)) & 0xFFFFFFFF
        md = (_Mic * max(nm, 1))()
        for q, d in enumerate(mic_defs):
            md[q].pos = _f3(d["pos"])
            md[q].axis = _f3(d.get("axis", (0.0, 0.0, 1.0)))
            md[q].polar_a = d.get("polar_a", 1.0)
            md[q].polar_b = d.get("polar_b", 0.0)
            md[q].gain = d.get("gain", 1.0)
        self._keep += [sd, pd, md]
        return sd, pd, md

    @classmethod
    def from_amr(cls, amr_desc, string_defs, pickup_defs, mic_defs, sample_rate,
                 modal_stride=16):
        """coevolver_create_amr (acoustic_coevolver.cpp:1387-1746) with the
        reference's amr_dict keys (pybind_kernels.cpp:5887-6060)."""
        self = cls.__new__(cls)
        L = self._L = lib()
        self._h = None
        self._keep = []
        sd, pd, md = self._defs(string_defs, pickup_defs, mic_defs)
        ns, npk, nm = len(string_defs), len(pickup_defs), len(mic_defs)
        d = _amr_desc(amr_desc, self._keep)
        self._amr_desc = amr_desc
        self.Nx, self.Ny, self.Nz = int(amr_desc["plate_Nx"]), int(amr_desc["plate_Ny"]), 1
        self.n_cells, self.n_faces = int(amr_desc["n_cells"]), int(amr_desc["n_faces"])
        self.n_strings, self.n_pickups, self.n_mics = ns, npk, nm
        stage = C.c_char_p()
        h = L.orc_coevo_create_amr(ns, C.cast(sd, C.c_void_p) if ns else None, npk,
                                   C.cast(pd, C.c_void_p) if npk else None, nm,
                                   C.cast(md, C.c_void_p) if nm else None,
                                   C.cast(C.pointer(d), C.c_void_p), float(sample_rate),
                                   int(modal_stride), C.byref(stage))
        if not h:
            raise RuntimeError(f"orc_coevo_create_amr failed at {stage.value.decode()}")
        self._h = h
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_coevo_destroy(self._h)
            self._h = None

    def step(self, n):
        return self._L.orc_coevo_step(self._h, int(n))

    def step_worker(self, n):
        return self._L.orc_coevo_step_worker(self._h, int(n))

    def step_block_with_drive(self, blocks, pos, n):
        arrs = [] if blocks is None else [np.ascontiguousarray(np.asarray(x, np.float32))
                                          for x in blocks]
        ptrs = (C.c_void_p * max(len(arrs), 1))(*[_v(a) for a in arrs])
        out = np.zeros(int(n), np.float32)
        rc = self._L.orc_coevo_step_block_with_drive(
            self._h, C.cast(ptrs, C.c_void_p) if arrs else None,
            len(arrs) if arrs else self.n_strings, int(n), float(pos), _v(out))
        return rc, out

    def pluck_string(self, si, pos, amp, ndur=0):
        return self._L.orc_coevo_pluck(self._h, si, pos, amp, ndur)

    def inject_string_force(self, si, f, pos):
        f = np.ascontiguousarray(np.asarray(f, np.float32))
        return self._L.orc_coevo_inject_force(self._h, si, f.size, _v(f), pos)

    def schedule_pluck(self, onset, si, pos, amp, ndur=0):
        return self._L.orc_coevo_schedule_pluck(self._h, onset, si, pos, amp, ndur)

    def clear_pluck_schedule(self):
        self._L.orc_coevo_clear_schedule(self._h)

    def reset(self):
        return self._L.orc_coevo_reset(self._h)

    def set_damping_scale(self, s):
        self._L.orc_coevo_set_damping_scale(self._h, s)

    def set_face_fractions(self, fx, fy, fz):
        a = [np.ascontiguousarray(np.asarray(x, np.float32)) for x in (fx, fy, fz)]
        return self._L.orc_coevo_set_face_fractions(self._h, _v(a[0]), _v(a[1]), _v(a[2]))

    def get_pickup_output(self, p, n):
        out = np.zeros(int(n), np.float32)
        rc = self._L.orc_coevo_get_pickup_output(self._h, p, _v(out), int(n))
        assert rc == 0
        return out

    def get_mic_output(self, m, n):
        out = np.zeros(int(n), np.float32)
        rc = self._L.orc_coevo_get_mic_output(self._h, m, _v(out), int(n))
        assert rc == 0
        return out

    @property
    def substeps(self):
        return self._L.orc_coevo_substeps(self._h)

    @property
    def samples_written(self):
        return self._L.orc_coevo_written(self._h)

    @property
    def dt_fdtd(self):
        return self._L.orc_coevo_dt_fdtd(self._h)

    @property
    def n_bridge(self):
        return self._L.orc_coevo_n_bridge(self._h)

    @property
    def n_neck(self):
        return self._L.orc_coevo_n_neck(self._h)

    def fields(self):
        if self._amr_desc is not None:                   # AMR body: cells, faces
            a = self._L.orc_coevo_amr(self._h)
            P = np.zeros(self.n_cells, np.float32)
            V = np.zeros(self.n_faces, np.float32)
            self._L.amo_get_pressure(a, _v(P), P.size)
            self._L.amo_get_velocity(a, _v(V), V.size)
            return P, V
        f = self._L.orc_coevo_fdtd(self._h)
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        P = np.zeros(Nx * Ny * Nz, np.float32)
        vx = np.zeros((Nx - 1) * Ny * Nz, np.float32)
        vy = np.zeros(Nx * (Ny - 1) * Nz, np.float32)
        vz = np.zeros(Nx * Ny * (Nz - 1), np.float32)
        self._L.orc_get_fields_f32(f, _v(P), _v(vx), _v(vy), _v(vz))
        return P, vx, vy, vz

    def plate(self):
        if self._amr_desc is not None:
            w = np.zeros(self.Nx * self.Ny, np.float32)
            self._L.amo_get_plate(self._L.orc_coevo_amr(self._h), _v(w), w.size)
            return w
        f = self._L.orc_coevo_fdtd(self._h)
        w = np.zeros(self.Nx * self.Ny, np.float32)
        self._L.orc_get_plate_f32(f, _v(w))
        return w

    def fdtd_step_count(self):
        return self._L.orc_get_step_count_f32(self._L.orc_coevo_fdtd(self._h))

    def string_state(self, si):
        N = self.n_nodes[si]
        out = np.zeros(4 * N + 6, np.float32)
        rc = self._L.orc_coevo_string_state(self._h, si, _v(out), out.size)
        assert rc == 0
        return {"u0": out[:N], "u1": out[N:2 * N], "up0": out[2 * N:3 * N],
                "up1": out[3 * N:4 * N], "T_eff": out[4 * N], "saddle_w": out[4 * N + 1],
                "saddle_w_prev": out[4 * N + 2], "bridge_w": out[4 * N + 3],
                "bridge_w_prev": out[4 * N + 4], "n_sub": int(out[4 * N + 5])}

    def neck_state(self):
        out = np.zeros(3, np.float32)
        self._L.orc_coevo_neck_state(self._h, _v(out))
        return out

    def modal(self, si, n=32):
        re = np.zeros(n, np.float32)
        im = np.zeros(n, np.float32)
        self._L.orc_coevo_modal(self._h, si, _v(re), _v(im), n)
        return re, im

    def string_frame(self, si):
        S = self.n_nodes[si] - 1
        out = np.zeros(9 * S + 4, np.float32)
        rc = self._L.orc_coevo_string_frame(self._h, si, _v(out), out.size)
        assert rc == 0
        return {"tan": out[:3 * S].reshape(S, 3), "nrm": out[3 * S:6 * S].reshape(S, 3),
                "bin": out[6 * S:9 * S].reshape(S, 3), "ds": out[9 * S], "L": out[9 * S + 1],
                "dt_s": out[9 * S + 2], "biharm": out[9 * S + 3]}

    def set_string(self, si, u0=None, u1=None, up0=None, up1=None):
        arrs = [None if x is None else np.ascontiguousarray(np.asarray(x, np.float32))
                for x in (u0, u1, up0, up1)]
        rc = self._L.orc_coevo_set_string(self._h, si, *[None if x is None else _v(x) for x in arrs])
        assert rc == 0

    def string_params(self, si):
        out = np.zeros(4, np.float32)
        self._L.orc_coevo_string_params(self._h, si, _v(out))
        return {"n_sub": int(out[0]), "ds": float(out[1]), "dt_s": float(out[2]), "c2": float(out[3])}

    def tables(self):
        nb, nn, ns = self.n_bridge, self.n_neck, self.n_strings
        bi, bw = np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1), np.float32)
        sbk = np.zeros(max(nb * ns, 1), np.float32)
        ni, nw = np.zeros(max(nn, 1), np.int32), np.zeros(max(nn, 1), np.float32)
        si, sw = np.zeros(max(4 * ns, 1), np.int32), np.zeros(max(4 * ns, 1), np.float32)
        self._L.orc_coevo_tables(self._h, _v(bi), _v(bw), _v(sbk), _v(ni), _v(nw), _v(si), _v(sw))
        return {"bidx": bi[:nb], "bwgt": bw[:nb], "sbk": sbk[:nb * ns].reshape(ns, nb),
                "nidx": ni[:nn], "nwgt": nw[:nn], "sidx": si[:4 * ns].reshape(ns, 4),
                "swgt": sw[:4 * ns].reshape(ns, 4)}

    def cell_type(self):
        out = np.zeros(self.Nx * self.Ny * self.Nz, np.uint8)
        self._L.orc_get_cell_type_f32(self._L.orc_coevo_fdtd(self._h), _v(out))
        return out
