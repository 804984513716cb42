"""Pins of the co-evolver oracle (oracle/coevo_oracle.c) by what the
mathematics and the reference's stated physics fix, independently of the
oracle's own expressions (acoustic_coevolver.cpp):

  * an ideal string mode evolves by the scalar recurrence of the centred,
    damped leapfrog (discrete sine modes are exact eigenvectors);
  * one step of a stiff string from a displaced node: the hand-derived
    5-point answer (3-point wave + 1,-4,6,-4,1 biharmonic, (1+h) division);
  * the geometric tension of a triangular shape from its closed-form arc
    length, T0 + EA (arc - L) / L;
  * the validity limits: a 0.28 slope kink fails, 0.24 passes; a 0.06 slope
    triangle (stretch > 0.0015) fails, 0.05 passes;
  * modal projection of sine modes (discrete orthogonality, scale 2/N);
  * single-coil and humbucker pickups brute-forced from the geometry, piezo
    from the saddle slope and tension;
  * mics brute-forced from the body fields (trilinear, half-cell shifts);
  * a slow pluck settles at the static amplitude of its per-step force
    (amp / (K n_sub): the reference spreads each sample's force over the
    sample's string steps, :1926-1930);
  * the bridge and neck modes satisfy their second-order ODEs
    m x'' + c (x' - v_p) + k (x - w_p) = f (:2009-2046, :1934-1983).
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def co():
    from oracle import coevo
    coevo.build()
    return coevo


def _free_body(n=6, dx=0.01):
    """A small air box, no active plate, no bridge sources: strings run
    decoupled from the body with rigid ends."""
    return {"Nx": n, "Ny": n, "Nz": n, "dx": dx, "gx_min": 0.0, "gy_min": 0.0, "gz_min": 0.0,
            "cell_type": np.zeros(n ** 3, np.uint8), "plate_iz": n // 2,
            "plate_active": np.zeros(n * n, np.uint8), "n_pml": 0,
            "bridge_src_xyz": np.zeros((0, 3), np.float32)}


def _string(n_segs=60, L=0.6, T=60.0, mu=0.005, gamma=3.0, EI=0.0, EA=1e-3, y0=0.0):
    path = np.zeros((n_segs + 1, 3), np.float32)
    path[:, 1] = np.linspace(y0, y0 + L, n_segs + 1)
    path[:, 2] = 0.05
    return {"path_xyz": path, "tension_N": T, "linear_mass_kgm": mu, "damping": gamma,
            "stiffness_EI": EI, "axial_stiffness_N": EA}


_PK = [{"type": 0, "pos": np.array([0.0, 0.3, 0.06], np.float32),
        "axis": np.array([0.0, 0.0, 1.0], np.float32), "pole_sigma": 0.01}]
_MIC = [{"pos": np.array([0.03, 0.03, 0.03], np.float32)}]


def _make(co, strings, sr=44100.0, pickups=_PK, mics=_MIC, body=None, modal_stride=16):
    return co.OracleCoEvolver(strings, pickups, mics, body or _free_body(), sr, modal_stride)


def test_ideal_string_mode_recurrence(co):
    o = _make(co, [_string()])
    K = o.substeps
    p = o.string_params(0)
    assert p["n_sub"] == 1
    N = 60
    i = np.arange(N + 1)
    A, B = 1e-4, 4e-5
    m0, m1 = 3, 5
    s0, s1 = np.sin(m0 * np.pi * i / N), np.sin(m1 * np.pi * i / N)
    o.set_string(0, u0=A * s0, u1=B * s1, up0=0.999 * A * s0, up1=1.001 * B * s1)
    n = 40
    assert o.step(n) == 0
    h = 0.5 * 3.0 * p["dt_s"]
    c2 = p["c2"]

    def amp(a, ap, m, steps):
        lam = 4.0 * math.sin(m * math.pi / (2 * N)) ** 2
        for _ in range(steps):
            a, ap = ((2.0 - c2 * lam) * a - (1.0 - h) * ap) / (1.0 + h), a
        return a

    st = o.string_state(0)
    steps = n * K * p["n_sub"]
    for u, s, a0, m in ((st["u0"], s0, A, m0), (st["u1"], s1, B, m1)):
        ap0 = 0.999 * A if m == m0 else 1.001 * B
        exp = amp(a0, ap0, m, steps) * s
        assert np.max(np.abs(u - exp)) < 1e-4 * a0, m
    assert abs(st["T_eff"] - 60.0) < 1e-4


def test_stiff_string_one_step_known_answer(co):
    EI, mu, gamma = 1e-4, 0.005, 3.0
    o = _make(co, [_string(EI=EI, mu=mu, gamma=gamma)], sr=80000.0)
    assert o.substeps == 1
    p = o.string_params(0)
    assert p["n_sub"] == 1
    ds, dt = p["ds"], p["dt_s"]
    c2 = (math.sqrt(60.0 / mu) * dt / ds) ** 2
    b = (EI / mu) * dt * dt / ds ** 4
    h = 0.5 * gamma * dt
    i0, a = 30, 1e-5
    u = np.zeros(61)
    u[i0] = a
    o.set_string(0, u0=u, up0=u)
    assert o.step(1) == 0
    got = o.string_state(0)["u0"].astype(np.float64)
    exp = np.zeros(61)
    exp[i0] = a * (1.0 + h - 2.0 * c2 - 6.0 * b) / (1.0 + h)
    exp[i0 - 1] = exp[i0 + 1] = a * (c2 + 4.0 * b) / (1.0 + h)
    exp[i0 - 2] = exp[i0 + 2] = -a * b / (1.0 + h)
    assert b > 1e-3 * c2                   # the stiffness term is not negligible
    assert np.max(np.abs(got - exp)) < 1e-6 * a


def _triangle(n, ip, height):
    i = np.arange(n + 1, dtype=np.float64)
    return np.where(i <= ip, height * i / ip, height * (n - i) / (n - ip))


def test_geometric_tension_closed_form(co):
    EA, T0, N, ip, a = 1e4, 60.0, 60, 20, 0.01
    o = _make(co, [_string(mu=1e3, gamma=0.0, EA=EA)], sr=80000.0)   # frozen string
    u = _triangle(N, ip, a)
    o.set_string(0, u0=u, up0=u)
    assert o.step(1) == 0
    ds, L = 0.01, 0.6
    arc = ip * math.sqrt(ds * ds + (a / ip) ** 2) + (N - ip) * math.sqrt(ds * ds + (a / (N - ip)) ** 2)
    T = T0 + EA * (arc - L) / L
    got = float(o.string_state(0)["T_eff"])
    assert abs(got - T) < 0.02 * (T - T0)


@pytest.mark.parametrize("shape,slope,ok", [("kink", 0.28, False), ("kink", 0.24, True),
                                           ("triangle", 0.06, False), ("triangle", 0.05, True)])
def test_validity_limits(co, shape, slope, ok):
    o = _make(co, [_string(mu=1e3, gamma=0.0)], sr=80000.0)
    N = 60
    if shape == "kink":
        u = np.zeros(N + 1)
        u[25:35] = slope * 0.01                     # two segments of the given slope
    else:
        u = _triangle(N, 30, slope * 30 * 0.01)
    o.set_string(0, u0=u, up0=u)
    rc = o.step(1)
    assert (rc == 0) == ok
    assert o.samples_written == (1 if ok else 0)


def test_modal_projection_orthogonality(co):
    o = _make(co, [_string(mu=1e3, gamma=0.0)], sr=80000.0, modal_stride=1)
    N = 60
    i = np.arange(N + 1)
    A, B = 2e-4, -7e-5
    u = A * np.sin(3 * np.pi * i / N) + B * np.sin(7 * np.pi * i / N)
    o.set_string(0, u0=u, up0=u)
    assert o.step(1) == 0
    re, im = o.modal(0)
    exp = np.zeros(32)
    exp[2], exp[6] = A, B
    assert np.max(np.abs(re - exp)) < 1e-6 * A
    assert np.max(np.abs(im - exp)) < 1e-6 * A


def _b_single(mid, pos, axis, sigma):
    ax = np.asarray(axis, np.float64)
    ax /= np.linalg.norm(ax)
    d = mid - np.asarray(pos, np.float64)
    r = d - np.outer(d @ ax, ax)
    return np.exp(-np.sum(r * r, axis=1) / (2 * sigma ** 2)), ax


def test_pickups_brute_force(co):
    axis = np.array([0.6, 0.0, 0.8], np.float32)
    pks = [{"type": 0, "pos": np.array([0.004, 0.31, 0.058], np.float32), "axis": axis,
            "pole_sigma": 0.012},
           {"type": 1, "pos": np.array([0.0, 0.2, 0.06], np.float32), "axis": axis,
            "pole_sigma": 0.012, "coil_spacing": 0.018},
           {"type": 2, "pos": np.zeros(3, np.float32), "axis": np.zeros(3, np.float32),
            "pole_sigma": 0.0, "sensitivity": 0.7, "string_mask": 0b10}]
    strs = [_string(mu=1e3, gamma=0.0), _string(mu=1e3, gamma=0.0, y0=0.01)]
    o = _make(co, strs, sr=80000.0, pickups=pks)
    N = 60
    i = np.arange(N + 1)
    for si, (m, a, v) in enumerate(((2, 3e-4, 0.02), (4, 1e-4, -0.03))):
        u = a * np.sin(m * np.pi * i / N)
        w = np.sin(np.pi * i / N) * 5e-5
        dt = o.string_params(si)["dt_s"]
        o.set_string(si, u0=u, up0=u - v * dt * np.sin(m * np.pi * i / N),
                     u1=w, up1=w + 0.5 * v * dt * np.sin(np.pi * i / N))
    assert o.step(1) == 0
    single, hum, piezo = 0.0, 0.0, 0.0
    for si, d in enumerate(strs):
        st = o.string_state(si)
        path = np.asarray(d["path_xyz"], np.float64)
        mid = 0.5 * (path[:-1] + path[1:])
        dt = o.string_params(si)["dt_s"]
        v0 = (st["u0"][:-1].astype(np.float64) - st["up0"][:-1]) / dt
        v1 = (st["u1"][:-1].astype(np.float64) - st["up1"][:-1]) / dt
        # frame of a string along +y: normal +z, binormal y x z = +x
        Bs, ax = _b_single(mid, pks[0]["pos"], axis, 0.012)
        ds = 0.01
        single += np.sum(Bs * (ax[2] * v0 + ax[0] * v1)) * ds
        B1, _ = _b_single(mid, pks[1]["pos"] + 0.009 * ax, axis, 0.012)
        B2, _ = _b_single(mid, pks[1]["pos"] - 0.009 * ax, axis, 0.012)
        hum += np.sum((B1 - B2) * (ax[2] * v0 + ax[0] * v1)) * ds
        if si == 1:
            s0 = (float(st["u0"][-1]) - float(st["u0"][-2])) / ds
            s1 = (float(st["u1"][-1]) - float(st["u1"][-2])) / ds
            piezo += 0.7 * float(st["T_eff"]) * math.hypot(s0, s1)
    out = [float(o.get_pickup_output(p, 1)[0]) for p in range(3)]
    assert abs(out[0] - single) < 1e-5 * abs(single) and abs(single) > 0
    # the two coils sit on the same axis line, so their lobes cancel
    assert abs(out[1]) <= 1e-6 * abs(single) and abs(hum) <= 1e-6 * abs(single)
    assert abs(out[2] - piezo) < 1e-5 * abs(piezo) and piezo > 0


def _trilinear(F, g):
    n = F.shape
    x = [min(max(g[a], 0.0), n[a] - 1.0) for a in range(3)]
    i0 = [int(x[a]) for a in range(3)]
    i1 = [min(i0[a] + 1, n[a] - 1) for a in range(3)]
    f = [x[a] - i0[a] for a in range(3)]
    v = 0.0
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                w = ((f[0] if cx else 1 - f[0]) * (f[1] if cy else 1 - f[1]) *
                     (f[2] if cz else 1 - f[2]))
                v += w * F[(i1 if cx else i0)[0], (i1 if cy else i0)[1], (i1 if cz else i0)[2]]
    return v


def test_mics_brute_force(co):
    import sys
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_gpu_coevolver import _instrument
    S, P, M, B = _instrument(n_segs=60, n_strings=2)
    M = [{"pos": np.array([0.013, -0.02, 0.131], np.float32), "axis": np.array([0.3, -0.4, -1.0],
                                                                            np.float32),
          "polar_a": 0.3, "polar_b": 0.7, "gain": 1.7},
         {"pos": np.array([-0.05, 0.07, 0.05], np.float32), "gain": 0.5}]
    o = co.OracleCoEvolver(S, P, M, B, 44100.0, 16)
    o.pluck_string(0, 0.3, 0.003, 8)
    assert o.step(80) == 0
    Pf, vx, vy, vz = o.fields()
    Nx, Ny, Nz = B["Nx"], B["Ny"], B["Nz"]
    Pf = Pf.reshape(Nx, Ny, Nz).astype(np.float64)
    V = [vx.reshape(Nx - 1, Ny, Nz), vy.reshape(Nx, Ny - 1, Nz), vz.reshape(Nx, Ny, Nz - 1)]
    org = np.array([B["gx_min"], B["gy_min"], B["gz_min"]])
    for mi, d in enumerate(M):
        g = (np.asarray(d["pos"], np.float64) - org) / B["dx"]
        p = _trilinear(Pf, g)
        v = np.array([_trilinear(V[a].astype(np.float64), g - 0.5 * np.eye(3)[a])
                      for a in range(3)])
        axis = np.asarray(d.get("axis", (0, 0, 1)), np.float64)
        exp = d.get("gain", 1.0) * (d.get("polar_a", 1.0) * p +
                                    d.get("polar_b", 0.0) * 1.21 * 343.0 * (v @ axis))
        got = float(o.get_mic_output(mi, 1)[0])
        assert abs(p) > 0 and abs(got - exp) < 1e-5 * max(abs(exp), abs(p)), mi


def test_slow_pluck_static_amplitude(co):
    o = _make(co, [_string()])
    o.set_damping_scale(100.0)
    K, n_sub = o.substeps, o.string_params(0)["n_sub"]
    amp, nd = 0.002, 4000
    o.pluck_string(0, 0.3, amp, nd)
    assert o.step(nd) == 0
    u = o.string_state(0)["u0"]
    ip = int(round(0.3 * 60))
    assert int(np.argmax(u)) == ip
    assert abs(u[ip] - amp / (K * n_sub)) < 0.02 * amp / (K * n_sub)


def test_bridge_and_neck_modes_obey_their_odes(co):
    import sys
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_gpu_coevolver import _instrument
    S, P, M, B = _instrument(dx=0.016, pad=3, n_segs=50, n_strings=2, n_pml=0)
    o = co.OracleCoEvolver(S, P, M, B, 50000.0, 16)
    assert o.substeps == 1 and o.n_bridge > 0 and o.n_neck > 0
    for si in range(2):
        assert o.string_params(si)["n_sub"] == 1
    o.pluck_string(0, 0.4, 0.003, 6)
    o.pluck_string(1, 0.6, -0.002, 10)
    hist = []
    for _ in range(60):
        assert o.step(1) == 0
        hist.append(([o.string_state(si) for si in range(2)], o.neck_state().astype(np.float64)))
    dt = o.dt_fdtd
    m = 0.00035
    wb = 2 * math.pi * 450.0
    kb, cb = m * wb * wb, m * wb / 2.5
    mn = 0.18
    wn, wj = 2 * math.pi * 65.0, 2 * math.pi * 95.0
    kn, cn, kj, cj = mn * wn * wn, mn * wn / 35.0, mn * wj * wj, mn * wj / 2.0
    worst_b = worst_n = 0.0
    for n in range(2, 60):
        (sa, na), (sb, nb_) = hist[n - 1], hist[n]
        for si in range(2):
            a, b = sa[si], sb[si]
            ds = float(np.linalg.norm(np.diff(np.asarray(S[si]["path_xyz"], np.float64), axis=0),
                                      axis=1).sum() / 50)
            x, xm, xp = float(a["bridge_w"]), float(a["bridge_w_prev"]), float(b["bridge_w"])
            s, sm = float(a["saddle_w"]), float(a["saddle_w_prev"])
            assert float(b["saddle_w_prev"]) == s
            f = float(b["T_eff"]) * (float(b["u0"][-2]) - x) / ds
            terms = (m * (xp - 2 * x + xm) / dt ** 2, cb * (xp - xm) / (2 * dt),
                     -cb * (s - sm) / dt, kb * (x - s), -f)
            worst_b = max(worst_b, abs(sum(terms)) / max(max(abs(t) for t in terms), 1e-300))
        x, xm, xp = na[0], na[1], nb_[0]
        pw, pwm = nb_[2], na[2]
        fs = sum(float(a["T_eff"]) * (float(a["u0"][1]) - x) /
                 float(np.linalg.norm(np.diff(np.asarray(S[si]["path_xyz"], np.float64), axis=0),
                                      axis=1).sum() / 50)
                 for si, a in enumerate(sa))
        terms = (mn * (xp - 2 * x + xm) / dt ** 2, (cn + cj) * (xp - xm) / (2 * dt),
                 -cj * (pw - pwm) / dt, (kn + kj) * x, -kj * pw, -fs)
        worst_n = max(worst_n, abs(sum(terms)) / max(max(abs(t) for t in terms), 1e-300))
    # float32 roundoff through the inertia terms; every term reaches >= 1% of the largest
    assert worst_b < 1e-3 and worst_n < 4e-3, (worst_b, worst_n)
    assert max(abs(float(h[0][0]["bridge_w"])) for h in hist) > 0
    assert max(abs(h[1][0]) for h in hist) > 0


def test_setup_tables_brute_force(co):
    """Bridge cells (Gaussian disc of radius 3 sigma under the plate, weights
    >= 1e-4, merged, normalised; :799-861), per-string bridge kernels (:1322-
    1375), the neck-joint kernel (centre 12 mm below the highest active plate
    row, :868-940), the saddle bilinear taps (:1160-1183) and the canonical
    soundhole carved from the plate (:946-981), recomputed in float64 from
    those rules."""
    import sys
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_gpu_coevolver import _instrument
    S, P, M, B = _instrument(dx=0.008, pad=4, n_segs=40, n_strings=3, n_pml=0)
    o = co.OracleCoEvolver(S, P, M, B, 44100.0, 16)
    t = o.tables()
    Nx, Ny, Nz, dx = B["Nx"], B["Ny"], B["Nz"], B["dx"]
    org = np.array([B["gx_min"], B["gy_min"]])
    sig = B["bridge_sigma"]
    rc = int(math.ceil(3 * sig / dx)) + 1
    cells = {}
    for sx, sy, _ in np.asarray(B["bridge_src_xyz"], np.float64):
        c0 = np.rint((np.array([sx, sy]) - org) / dx).astype(int)
        for di in range(-rc, rc + 1):
            for dj in range(-rc, rc + 1):
                i, j = c0[0] + di, c0[1] + dj
                if not (0 <= i < Nx and 0 <= j < Ny):
                    continue
                w = math.exp(-((i * dx + org[0] - sx) ** 2 + (j * dx + org[1] - sy) ** 2) /
                             (2 * sig * sig))
                if w < 1e-4:
                    continue
                flat = (B["plate_iz"] - 1) + Nz * (j + Ny * i)
                cells[flat] = cells.get(flat, 0.0) + w
    assert list(t["bidx"]) == list(cells)                  # first-seen order
    tot = sum(cells.values())
    assert np.allclose(t["bwgt"], np.array(list(cells.values())) / tot, rtol=1e-5, atol=0)
    ij = np.array([((f // Nz) // Ny, (f // Nz) % Ny) for f in cells], np.float64)
    xy = ij * dx + org
    for si, d in enumerate(S):
        sx, sy = np.asarray(d["path_xyz"], np.float64)[-1, :2]
        k = np.exp(-((xy[:, 0] - sx) ** 2 + (xy[:, 1] - sy) ** 2) / (2 * sig * sig))
        assert np.allclose(t["sbk"][si], k / k.sum(), rtol=1e-5, atol=1e-9)
        g = np.clip((np.array([sx, sy]) - org) / dx, 0, [Nx - 1, Ny - 1])
        i0, j0 = int(g[0]), int(g[1])
        i1, j1 = min(i0 + 1, Nx - 1), min(j0 + 1, Ny - 1)
        fi, fj = g[0] - i0, g[1] - j0
        assert list(t["sidx"][si]) == [j0 + Ny * i0, j0 + Ny * i1, j1 + Ny * i0, j1 + Ny * i1]
        assert np.allclose(t["swgt"][si], [(1 - fi) * (1 - fj), fi * (1 - fj), (1 - fi) * fj,
                                           fi * fj], atol=1e-6)
    pa = np.asarray(B["plate_active"]).reshape(Nx, Ny).astype(bool)
    ys = np.arange(Ny) * dx + org[1]
    xs = np.arange(Nx) * dx + org[0]
    ymax = ys[np.where(pa.any(axis=0))[0].max()]
    cy, s2 = ymax - 0.012, 0.030
    rcn = int(math.ceil(3 * s2 / dx)) + 1
    ci, cj = int(math.floor((0.0 - org[0]) / dx)), int(math.floor((cy - org[1]) / dx))
    nk = [(j + Ny * i, math.exp(-(xs[i] ** 2 + (ys[j] - cy) ** 2) / (2 * s2 * s2)))
          for i in range(max(ci - rcn, 0), min(ci + rcn + 1, Nx))
          for j in range(max(cj - rcn, 0), min(cj + rcn + 1, Ny)) if pa[i, j]]
    assert list(t["nidx"]) == [p for p, _ in nk]
    wn = np.array([w for _, w in nk])
    assert np.allclose(t["nwgt"], wn / wn.sum(), rtol=1e-5, atol=0)
    # soundhole: plate cells of the body FDTD = active minus the disc
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    hole = X ** 2 + (Y - 0.050) ** 2 < 0.028 ** 2
    plate = o.cell_type().reshape(Nx, Ny, Nz)[:, :, B["plate_iz"]] == 2
    assert hole.any() and (pa & hole).any()
    assert np.array_equal(plate, pa & ~hole)
