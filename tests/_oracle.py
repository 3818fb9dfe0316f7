"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and the reference
harness (oracle/_ref/ref_harness).  Test infrastructure: used only as the
checker for the HIP library.
"""
import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_HARNESS = os.path.join(ORACLE_DIR, "_ref", "ref_harness")
REF_HARNESS_FAST = os.path.join(ORACLE_DIR, "_ref", "ref_harness_fast")


class OrParams(C.Structure):
    _fields_ = [("N", C.c_int), ("a", C.c_double), ("mass_mode", C.c_int), ("m2", C.c_double),
                ("lambda_mode", C.c_int), ("kappa", C.c_double)]


class OrIcParams(C.Structure):
    _fields_ = [("QS", C.c_double), ("scale2", C.c_double), ("amplitude", C.c_double),
                ("noise", C.c_double), ("m2_long", C.c_double), ("justnoise", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                           capture_output=True)
        L = C.CDLL(ORACLE_LIB)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        P = C.POINTER
        sig = {
            "or_time_replay": (None, [C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp]),
            "or_exchange": (None, [C.c_int, dp]),
            "or_update_momenta": (None, [P(OrParams), dp, dp, C.c_double]),
            "or_update_fields": (None, [C.c_int, dp, dp, C.c_double]),
            "or_evolve": (None, [P(OrParams), dp, dp, C.c_int, dp, dp]),
            "or_volume_averages": (None, [P(OrParams), dp, dp, dp, dp]),
            "or_extrema": (None, [C.c_int, dp, dp, ip]),
            "or_bin_count": (C.c_int, [C.c_int, C.c_double, dp]),
            "or_bin_table": (C.c_int, [C.c_int, C.c_double, ip]),
            "or_spectrum_bins": (None, [C.c_int, C.c_double, dp, dp, dp, dp, dp]),
            "or_print_fp_rows": (None, [C.c_int, C.c_double, C.c_int, ip, dp, dp, dp, ip]),
            "or_srand48": (None, [C.c_long]),
            "or_drand48": (C.c_double, []),
            "or_gauss": (C.c_double, []),
            "or_ic_modes": (None, [C.c_int, C.c_double, P(OrIcParams), C.c_long, dp, dp]),
            "or_ensure_real": (None, [C.c_int, dp]),
            "or_ic_fill": (None, [C.c_int, C.c_double, dp, dp, C.c_double, C.c_double, dp, dp]),
            "or_islands_bfs": (C.c_int, [P(OrParams), C.c_int, C.c_double, dp, dp, dp, ip, dp, C.c_int]),
            "or_evolve_f32": (None, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp, fp,
                                     C.c_int, fp, fp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# cfg0 model in reference units (Lambda = 6*lambda, kappa^2 = Lambda^2/(120 g))
def model(N, lam=1.0, g=0.2, a=0.5):
    L = 6.0 * lam
    kappa = float(np.sqrt(L * L / (120.0 * g)))
    return dict(N=N, a=a, mass_mode=1, m2=1.0, lambda_mode=-1, kappa=kappa, Lambda=L)


def params(m):
    return OrParams(m["N"], m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"])


def to_ref(dense, N):
    """dense N^3 interior -> reference pos() layout with halo planes."""
    r = np.empty(N * N * (N + 2))
    P = N * N
    r[P:P * (N + 1)] = dense
    r[P * (N + 1):] = dense[:P]
    r[:P] = dense[P * (N - 1):]
    return r


def from_ref(ref, N):
    P = N * N
    return ref[P:P * (N + 1)].copy()


def time_factors(nsteps, T0=0.0, dtbase=0.1, a=0.5):
    dtk = np.empty(nsteps)
    dtd = np.empty(nsteps)
    tau = np.empty(nsteps)
    lib().or_time_replay(T0, dtbase, a, nsteps, dptr(dtk), dptr(dtd), dptr(tau))
    return dtk, dtd, tau


def evolve(m, phi_dense, pi_dense, nsteps, T0=0.0, dtbase=0.1):
    N = m["N"]
    phi = to_ref(phi_dense, N)
    pi = to_ref(pi_dense, N)
    dtk, dtd, _ = time_factors(nsteps, T0, dtbase, m["a"])
    p = params(m)
    lib().or_evolve(C.byref(p), dptr(phi), dptr(pi), nsteps, dptr(dtk), dptr(dtd))
    return from_ref(phi, N), from_ref(pi, N)


def evolve_tables(m, phi_dense, pi_dense, dtk, dtd):
    """evolve with given per-step factors (a slice of one continuous replay)."""
    N = m["N"]
    phi = to_ref(phi_dense, N)
    pi = to_ref(pi_dense, N)
    dtk = np.ascontiguousarray(dtk, dtype=np.float64)
    dtd = np.ascontiguousarray(dtd, dtype=np.float64)
    p = params(m)
    lib().or_evolve(C.byref(p), dptr(phi), dptr(pi), int(dtk.size), dptr(dtk), dptr(dtd))
    return from_ref(phi, N), from_ref(pi, N)


def volume_averages(m, phi_dense, pi_dense, want_delta=False):
    N = m["N"]
    phi = to_ref(phi_dense, N)
    pi = to_ref(pi_dense, N)
    pi[N * N * (N + 1):] = pi[N * N:2 * N * N]
    out = np.empty(10)
    delta = np.zeros(N * N * (N + 2)) if want_delta else None
    p = params(m)
    lib().or_volume_averages(C.byref(p), dptr(phi), dptr(pi), dptr(out),
                             dptr(delta) if want_delta else None)
    if want_delta:
        return out, from_ref(delta, N)
    return out


def extrema(N, phi_dense):
    phi = to_ref(phi_dense, N)
    vals = np.empty(3)
    idx = np.empty(3, dtype=np.int32)
    lib().or_extrema(N, dptr(phi), dptr(vals), iptr(idx))
    return vals, idx


def bin_table(N, a):
    nb = lib().or_bin_count(N, a, None)
    counts = np.zeros(nb, dtype=np.int32)
    r = lib().or_bin_table(N, a, iptr(counts))
    assert r == nb
    return counts


def print_fp_rows(N, a, counts, F, K):
    """Print_fp's seven columns per shell and which shells it writes
    (momentumspace.cpp:318-369)."""
    nb = counts.size
    rows = np.zeros(7 * nb)
    valid = np.zeros(nb, dtype=np.int32)
    c = np.ascontiguousarray(counts, dtype=np.int32)
    lib().or_print_fp_rows(N, a, nb, iptr(c), dptr(np.ascontiguousarray(F, dtype=np.float64)),
                           dptr(np.ascontiguousarray(K, dtype=np.float64)), dptr(rows), iptr(valid))
    return rows.reshape(nb, 7), valid.astype(bool)


def ref_modes(field_dense, N, a):
    """a^3-scaled r2c in the reference p-space order [pz][px][py] (clusterfft.cpp:185-225)."""
    f = field_dense.reshape(N, N, N)            # [z][y][x]
    arr = np.ascontiguousarray(f.transpose(0, 2, 1))  # [z][x][y]
    return (a * a * a) * np.fft.rfftn(arr)      # [pz][px][py]


def spectrum(N, a, phi_dense, pi_dense):
    nb = lib().or_bin_count(N, a, None)
    pk = np.ascontiguousarray(ref_modes(phi_dense, N, a)).view(np.float64).ravel()
    kk = np.ascontiguousarray(ref_modes(pi_dense, N, a)).view(np.float64).ravel()
    F = np.zeros(nb)
    K = np.zeros(nb)
    Q = np.zeros(nb)
    lib().or_spectrum_bins(N, a, dptr(pk), dptr(kk), dptr(F), dptr(K), dptr(Q))
    return F, K, Q


def islands(m, mode, thr, phi_dense, pi_dense, cap=100000):
    N = m["N"]
    phi = to_ref(phi_dense, N)
    pi = to_ref(pi_dense, N)
    delta = None
    if mode == 1:
        _, dd = volume_averages(m, phi_dense, pi_dense, want_delta=True)
        delta = to_ref(dd, N)
    b = np.empty(N * N * (N + 2), dtype=np.int32)
    cat = np.zeros(10 * cap)
    p = params(m)
    n = lib().or_islands_bfs(C.byref(p), mode, thr, dptr(phi), dptr(pi),
                             dptr(delta) if delta is not None else None, iptr(b), dptr(cat), cap)
    assert n >= 0
    return cat[:10 * n].reshape(n, 10), from_ref(b.astype(np.float64), N).astype(np.int32)


def ic_modes(N, a, seed, m2_long=1.0, Lambda=6.0, noise=0.5, QS=None, scale2=None, amplitude=0.0):
    """initialize==3 modes, with ensureRealResult applied."""
    if QS is None:
        QS = np.sqrt(3.0 * (4.0 / (a * a))) / 2.0
    if scale2 is None:
        scale2 = QS
    ic = OrIcParams(QS, scale2, amplitude, noise * Lambda, m2_long, 0)
    H = N // 2 + 1
    A = np.zeros(2 * N * N * H)
    E = np.zeros(2 * N * N * H)
    lib().or_ic_modes(N, a, C.byref(ic), seed, dptr(A), dptr(E))
    lib().or_ensure_real(N, dptr(A))
    lib().or_ensure_real(N, dptr(E))
    return A.view(np.complex128).reshape(N, N, H), E.view(np.complex128).reshape(N, N, H)


def c2r_unnormalised(F, N):
    """FFTW-style unnormalised c2r of [pz][px][py] modes -> Out[x + y*N + z*N*N]."""
    out_zxy = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2)) * float(N) ** 3   # [z][x][y]
    return np.ascontiguousarray(out_zxy.transpose(0, 2, 1)).ravel()   # [z][y][x]


def initial_conditions(m, seed, phi0, pi0=0.0, noise=0.5):
    """SetInitialConditions, initialize==3 (initialconditions.cpp:45-723) with numpy's c2r."""
    N, a = m["N"], m["a"]
    A, E = ic_modes(N, a, seed, m2_long=m["m2"], Lambda=m["Lambda"], noise=noise)
    AX = c2r_unnormalised(A, N)
    EX = c2r_unnormalised(E, N)
    phi = np.empty(N * N * (N + 2))
    pi = np.empty(N * N * (N + 2))
    lib().or_ic_fill(N, a, dptr(AX), dptr(EX), phi0, pi0, dptr(phi), dptr(pi))
    return from_ref(phi, N), from_ref(pi, N)


def evolve_f32(m, phi_dense, pi_dense, nsteps, T0=0.0, dtbase=0.1):
    N, a = m["N"], m["a"]
    inv_a2 = np.float32(1.0 / (a * a))
    c2 = np.float32(m["mass_mode"] * m["m2"])
    c4 = np.float32(m["lambda_mode"] / 6.0)
    c6 = np.float32(1.0 / (120.0 * m["kappa"] * m["kappa"]))
    phi = to_ref(phi_dense.astype(np.float32).astype(np.float64), N).astype(np.float32)
    pi = to_ref(pi_dense.astype(np.float32).astype(np.float64), N).astype(np.float32)
    dtk, dtd, _ = time_factors(nsteps, T0, dtbase, a)
    dtk = dtk.astype(np.float32)
    dtd = dtd.astype(np.float32)
    lib().or_evolve_f32(N, inv_a2, c2, c4, c6, fptr(phi), fptr(pi), nsteps, fptr(dtk), fptr(dtd))
    P = N * N
    return phi[P:P * (N + 1)].copy(), pi[P:P * (N + 1)].copy()


# ---------------------------------------------------------------------------
# reference harness (reference sources compiled unmodified, strict flags)
# ---------------------------------------------------------------------------
def have_ref(fast=False):
    return os.path.exists(REF_HARNESS_FAST if fast else REF_HARNESS)


def run_ref(m, phi_dense, pi_dense, nsteps, islands=False, thr=-1.0, T0=0.0, dtbase=0.1,
            fast=False, save_config=None):
    N = m["N"]
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<iid", nsteps, int(islands), thr))
            f.write(np.ascontiguousarray(phi_dense, dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(pi_dense, dtype=np.float64).tobytes())
        args = [REF_HARNESS_FAST if fast else REF_HARNESS, fin, fout,
                "-Nt", str(N), "-at", repr(m["a"]), "-M", repr(m["m2"]),
                "-Mmode", str(m["mass_mode"]), "-L", repr(m.get("Lambda", 6.0)),
                "-Lmode", str(m["lambda_mode"]), "-Kappa", repr(m["kappa"]),
                "-T0", repr(T0), "-dt", repr(dtbase)]
        env = dict(os.environ)
        if save_config is not None:
            env["REF_SAVECONFIG"] = str(int(save_config))
        r = subprocess.run(args, capture_output=True, text=True, cwd=td, timeout=600, env=env)
        if r.returncode != 0:
            raise RuntimeError("ref harness failed: %s %s" % (r.stdout, r.stderr))
        b = open(fout, "rb").read()
        files = {}
        if save_config is not None:
            for fn in ("field-%d-0.out" % save_config, "moment-%d.out" % save_config):
                files[fn] = open(os.path.join(td, fn)).read()
        odir = os.path.join(td, "output", "0")
        if os.path.isdir(odir):
            for fn in sorted(os.listdir(odir)):
                files[fn] = open(os.path.join(odir, fn)).read()
    n3 = N ** 3
    o = 0
    phi = np.frombuffer(b, np.float64, n3, o).copy(); o += 8 * n3
    pi = np.frombuffer(b, np.float64, n3, o).copy(); o += 8 * n3
    avg = np.frombuffer(b, np.float64, 10, o).copy(); o += 80
    ext = np.frombuffer(b, np.float64, 3, o).copy(); o += 24
    extpos = np.frombuffer(b, np.int32, 3, o).copy(); o += 12
    thr_used = struct.unpack_from("<d", b, o)[0]; o += 8
    nisl = struct.unpack_from("<i", b, o)[0]; o += 4
    cat = np.frombuffer(b, np.float64, 10 * nisl, o).reshape(nisl, 10).copy(); o += 80 * nisl
    labels = np.frombuffer(b, np.int32, n3, o).copy(); o += 4 * n3
    track = None
    if int(islands) >= 2:
        nt = struct.unpack_from("<i", b, o)[0]; o += 4
        track = []
        for _ in range(nt):
            index = struct.unpack_from("<q", b, o)[0]; o += 8
            status, disp, nh = struct.unpack_from("<3i", b, o); o += 12
            e = np.frombuffer(b, np.float64, 8 * nh, o).reshape(nh, 8); o += 64 * nh
            hist = [(r[0], int(r[1]), r[2], r[3], r[4], r[5], r[6], r[7]) for r in e]
            track.append((index, status, disp, hist))
    return dict(phi=phi, pi=pi, avg=avg, ext=ext, extpos=extpos, thr=thr_used, cat=cat,
                labels=labels, track=track, files=files)


# ---------------------------------------------------------------------------
# differentiable leapfrog checker (oracle/oracle_adjoint.c)
# ---------------------------------------------------------------------------
class OrAdjParams(C.Structure):
    _fields_ = [("N", C.c_int), ("a2", C.c_double), ("c2", C.c_double), ("c4", C.c_double),
                ("c6", C.c_double), ("t", C.c_double), ("w", C.c_double)]


def _adj_lib():
    L = lib()
    if not getattr(L, "_adj_ready", False):
        dp = C.POINTER(C.c_double)
        P = C.POINTER(OrAdjParams)
        for name, res, args in [
            ("or_poly_evolve", None, [P, dp, dp, C.c_int, dp, dp]),
            ("or_observable", C.c_double, [P, dp, dp, dp, dp]),
            ("or_forward_F", C.c_double, [P, dp, dp, C.c_int, dp, dp]),
            ("or_tlm", C.c_double, [P, dp, dp, C.c_int, dp, dp, dp, dp, dp]),
            ("or_adjoint", C.c_double, [P, dp, dp, C.c_int, dp, dp, dp, dp, dp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._adj_ready = True
    return L


def adj_params(N, c2=1.0, c4=-1.0, c6=0.2, a=0.5, t=4.0, w=0.1):
    """Physical-units theory 0: V = c2 phi^2/2 + c4 phi^4/4 + c6 phi^6/6 (c4 = -lambda, c6 = g)."""
    return OrAdjParams(N, a * a, c2, c4, c6, t, w)


def adj_forward_F(p, phi0, pi0, dtk, dtd):
    return _adj_lib().or_forward_F(C.byref(p), dptr(np.ascontiguousarray(phi0)),
                                   dptr(np.ascontiguousarray(pi0)), len(dtk),
                                   dptr(np.ascontiguousarray(dtk)), dptr(np.ascontiguousarray(dtd)))


def adj_observable(p, phi, pi):
    n3 = p.N ** 3
    rho = np.empty(n3)
    g = np.empty(n3)
    F = _adj_lib().or_observable(C.byref(p), dptr(np.ascontiguousarray(phi)),
                                 dptr(np.ascontiguousarray(pi)), dptr(rho), dptr(g))
    return F, rho, g


def adj_tlm(p, phi0, pi0, dtk, dtd, dphi0, dpi0, dc):
    dc = np.ascontiguousarray(dc, dtype=np.float64)
    return _adj_lib().or_tlm(C.byref(p), dptr(np.ascontiguousarray(phi0)),
                             dptr(np.ascontiguousarray(pi0)), len(dtk),
                             dptr(np.ascontiguousarray(dtk)), dptr(np.ascontiguousarray(dtd)),
                             dptr(np.ascontiguousarray(dphi0)), dptr(np.ascontiguousarray(dpi0)),
                             dptr(dc))


def adj_reverse(p, phi0, pi0, dtk, dtd):
    n3 = p.N ** 3
    gphi = np.empty(n3)
    gpi = np.empty(n3)
    gc = np.empty(3)
    F = _adj_lib().or_adjoint(C.byref(p), dptr(np.ascontiguousarray(phi0)),
                              dptr(np.ascontiguousarray(pi0)), len(dtk),
                              dptr(np.ascontiguousarray(dtk)), dptr(np.ascontiguousarray(dtd)),
                              dptr(gphi), dptr(gpi), dptr(gc))
    return F, gphi, gpi, gc


# ---------------------------------------------------------------------------
# island tracking checker (oracle/oracle_track.cpp)
# ---------------------------------------------------------------------------
ORACLE_TRACK_LIB = os.path.join(ORACLE_DIR, "liboracle_track.so")


class TrackEntry(C.Structure):
    _fields_ = [("at_time", C.c_double), ("center_pos", C.c_longlong), ("peak_value", C.c_double),
                ("avg_phi", C.c_double), ("volume", C.c_double), ("pot_en", C.c_double),
                ("kin_en", C.c_double), ("grad_en", C.c_double)]


class TrackedIsland(C.Structure):
    _fields_ = [("index", C.c_longlong), ("status", C.c_int), ("disp", C.c_int), ("nhist", C.c_int),
                ("hist", TrackEntry * 5)]


_track_lib = None


def _tlib():
    global _track_lib
    if _track_lib is None:
        if not os.path.exists(ORACLE_TRACK_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle_track.so"], check=True,
                           capture_output=True)
        L = C.CDLL(ORACLE_TRACK_LIB)
        dp = C.POINTER(C.c_double)
        L.or_track_create.restype = C.c_void_p
        L.or_track_create.argtypes = [C.POINTER(OrParams)]
        L.or_track_destroy.argtypes = [C.c_void_p]
        L.or_track_count.restype = C.c_int
        L.or_track_count.argtypes = [C.c_void_p]
        L.or_track_get.argtypes = [C.c_void_p, C.c_int, C.POINTER(TrackedIsland)]
        L.or_track_update.argtypes = [C.c_void_p, C.c_int, C.c_double, dp, dp, dp, C.c_double,
                                      C.c_char_p]
        _track_lib = L
    return _track_lib


def tracked_list(get, n):
    """[(index, status, disp, [entry tuples])] from a getter filling TrackedIsland."""
    out = []
    for i in range(n):
        t = TrackedIsland()
        get(i, t)
        hist = [(e.at_time, int(e.center_pos), e.peak_value, e.avg_phi, e.volume, e.pot_en,
                 e.kin_en, e.grad_en) for e in t.hist[:t.nhist]]
        out.append((int(t.index), int(t.status), int(t.disp), hist))
    return out


class Tracker:
    """Full_Analysis called repeatedly on the CPU (islands.cpp:788-835)."""

    def __init__(self, m):
        self.m = dict(m)
        self._p = params(m)
        self.h = _tlib().or_track_create(C.byref(self._p))

    def update(self, mode, thr, phi_dense, pi_dense, tau, outdir=None):
        N = self.m["N"]
        phi = to_ref(phi_dense, N)
        pi = to_ref(pi_dense, N)
        delta = None
        if mode == 1:
            _, dd = volume_averages(self.m, phi_dense, pi_dense, want_delta=True)
            delta = to_ref(dd, N)
        _tlib().or_track_update(self.h, mode, thr, dptr(phi), dptr(pi),
                                dptr(delta) if delta is not None else None, tau,
                                outdir.encode() if outdir else None)

    def islands(self):
        L = _tlib()
        return tracked_list(lambda i, t: L.or_track_get(self.h, i, C.byref(t)),
                            L.or_track_count(self.h))

    def close(self):
        if self.h:
            _tlib().or_track_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def track_replay(m, phi, pi, analyses, steps, thr, mode=0):
    """The tracking run of the reference harness on the CPU oracle: an
    analysis at tau = T0 = 0 and after each block of `steps` steps (one
    continuous time replay).  Returns (tracked list, {file name: text})."""
    dtk, dtd, taus = time_factors(steps * (analyses - 1))
    T = Tracker(m)
    f, p = np.array(phi, dtype=np.float64), np.array(pi, dtype=np.float64)
    with tempfile.TemporaryDirectory() as td:
        for k in range(analyses):
            if k > 0:
                sl = slice(steps * (k - 1), steps * k)
                f, p = evolve_tables(m, f, p, dtk[sl], dtd[sl])
            T.update(mode, thr, f, p, 0.0 if k == 0 else taus[steps * k - 1], td)
        files = {fn: open(os.path.join(td, fn)).read() for fn in sorted(os.listdir(td))}
    out = T.islands()
    T.close()
    return out, files


def load_track_fixture(path):
    z = np.load(path)
    m = dict(N=int(z["N"]), a=float(z["a"]), mass_mode=int(z["mass_mode"]), m2=float(z["m2"]),
             lambda_mode=int(z["lambda_mode"]), kappa=float(z["kappa"]), Lambda=float(z["Lambda"]))
    track = []
    for i in range(z["t_index"].size):
        nh = int(z["t_nhist"][i])
        hist = [(r[0], int(r[1]), r[2], r[3], r[4], r[5], r[6], r[7]) for r in z["t_hist"][i, :nh]]
        track.append((int(z["t_index"][i]), int(z["t_status"][i]), int(z["t_disp"][i]), hist))
    files = {str(n): str(z["file_" + str(n).replace(".", "_")]) for n in z["file_names"]}
    return m, z, track, files


def config_text(phi_dense, pi_dense, fp32=False):
    """SaveConfig's field file (fieldanalysis.cpp:252-275): one "Phi Pi" line
    per interior site in pos() order, printf %.17e (%.9e for FP32 fields)."""
    fmt = "%.9e %.9e\n" if fp32 else "%.17e %.17e\n"
    return "".join(fmt % (float(a), float(b)) for a, b in zip(phi_dense, pi_dense))


def moment_text(tau):
    """SaveConfig's moment file: `fout << tau << endl` (six significant digits)."""
    return "%g\n" % tau
