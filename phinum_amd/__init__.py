"""phinum_amd -- MI355X (gfx950) lattice hot path of PhiNum.

Python side of the C ABI declared in include/phinum_gpu.h (ctypes).  The
shared library libphinum_gpu.so is built in-tree by phinum_amd/build.py; if
it is missing, importing the bindings raises instead of falling back to any
CPU path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphinum_gpu.so")

PN_OK = 0
PN_ERR_ARG = -1
PN_ERR_HIP = -2
PN_ERR_FFT = -3
PN_ERR_STATE = -4
PN_ERR_NOMEM = -5
PN_ERR_CAPACITY = -6
PN_FP64, PN_FP32 = 0, 1
PN_LAYOUT_REF, PN_LAYOUT_DENSE = 0, 1
PN_OPT_TWO_STEP = 1
PN_OPT_IC_ON_DEVICE = 2
PN_OPT_STEP2_ZBLOCK = 3
PN_OPT_THREE_STEP = 4
PN_OPT_STEP3_ZBLOCK = 5
PN_OPT_KERNEL_TIMED = 6
PN_VOLUME_OUT = 16
PN_NOALLOC_MAX_N = 2048


class PnParams(C.Structure):
    _fields_ = [("N", C.c_int), ("a", C.c_double), ("mass_mode", C.c_int), ("m2", C.c_double),
                ("lambda_mode", C.c_int), ("kappa", C.c_double), ("precision", C.c_int),
                ("device", C.c_int)]


class PnAverages(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("phi", "pi", "phi2", "pi2", "phipi", "phi4", "phi6",
                                          "energy", "interaction", "charge")]

    def as_array(self):
        return np.array([getattr(self, n) for n, _ in self._fields_])


class PnExtrema(C.Structure):
    _fields_ = [("min", C.c_double), ("max", C.c_double), ("maxabs", C.c_double),
                ("minpos", C.c_longlong), ("maxpos", C.c_longlong), ("maxabspos", C.c_longlong)]


class PnSpectrum(C.Structure):
    _fields_ = [("nbins", C.c_int), ("binsize", C.c_double), ("counts", C.POINTER(C.c_int)),
                ("corr_f", C.POINTER(C.c_double)), ("corr_k", C.POINTER(C.c_double)),
                ("corr_q", C.POINTER(C.c_double))]


class PnIsland(C.Structure):
    _fields_ = [("index", C.c_longlong), ("nn", C.c_longlong), ("first_pos", C.c_longlong),
                ("peak_pos", C.c_longlong), ("volume", C.c_double), ("peak_value", C.c_double),
                ("avg_phi", C.c_double), ("pot_en", C.c_double), ("kin_en", C.c_double),
                ("grad_en", C.c_double), ("sum_dx", C.c_longlong), ("sum_dy", C.c_longlong),
                ("sum_dz", C.c_longlong), ("cx", C.c_double), ("cy", C.c_double),
                ("cz", C.c_double), ("peak_ties", C.c_longlong)]


class PnIcParams(C.Structure):
    _fields_ = [("seed", C.c_longlong), ("QS", C.c_double), ("scale2", C.c_double),
                ("amplitude", C.c_double), ("noise", C.c_double), ("m2_long", C.c_double),
                ("justnoise", C.c_int), ("phi0", C.c_double), ("pi0", C.c_double)]


class PnRealization(C.Structure):
    _fields_ = [("N", C.c_int), ("a", C.c_double), ("mass_mode", C.c_int), ("m2", C.c_double),
                ("lambda_mode", C.c_int), ("kappa", C.c_double)]


class PnAdjParams(C.Structure):
    _fields_ = [("N", C.c_int), ("a", C.c_double), ("c2", C.c_double), ("c4", C.c_double),
                ("c6", C.c_double), ("threshold", C.c_double), ("width", C.c_double),
                ("checkpoint_every", C.c_int), ("device", C.c_int)]


class PnAdjResult(C.Structure):
    _fields_ = [("F", C.c_double), ("grad_c", C.c_double * 3), ("forward_ms", C.c_double),
                ("backward_ms", C.c_double), ("peak_bytes", C.c_longlong)]


class PnCatalog(C.Structure):
    _fields_ = [("capacity", C.c_int), ("count", C.c_int), ("threshold", C.c_double),
                ("islands", C.POINTER(PnIsland)), ("labels", C.POINTER(C.c_int))]


class PnFieldDiff(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("max_dphi", "max_dpi", "max_phi", "max_pi",
                                          "sum_dphi2", "sum_phi2", "sum_dpi2", "sum_pi2")]


class PnTrackEntry(C.Structure):
    _fields_ = [("at_time", C.c_double), ("center_pos", C.c_longlong), ("peak_value", C.c_double),
                ("avg_phi", C.c_double), ("volume", C.c_double), ("pot_en", C.c_double),
                ("kin_en", C.c_double), ("grad_en", C.c_double)]


class PnTrackedIsland(C.Structure):
    _fields_ = [("index", C.c_longlong), ("status", C.c_int), ("disp", C.c_int), ("nhist", C.c_int),
                ("hist", PnTrackEntry * 5)]


_lib = None


def lib():
    """Loads libphinum_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("phinum_amd: %s is missing; run phinum_amd/build.py (or "
                           "__graft_entry__.build()) first" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp = P(C.c_double)
    sig = {
        "pn_create": (C.c_int, [P(PnParams), P(C.c_void_p)]),
        "pn_destroy": (C.c_int, [C.c_void_p]),
        "pn_last_error": (C.c_char_p, [C.c_void_p]),
        "pn_abi_version": (C.c_int, []),
        "pn_set_fields": (C.c_int, [C.c_void_p, dp, dp, C.c_int]),
        "pn_get_fields": (C.c_int, [C.c_void_p, dp, dp, C.c_int]),
        "pn_kick": (C.c_int, [C.c_void_p, C.c_double]),
        "pn_drift": (C.c_int, [C.c_void_p, C.c_double]),
        "pn_evolve": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
        "pn_sync": (C.c_int, [C.c_void_p]),
        "pn_evolve_timed": (C.c_int, [C.c_void_p, C.c_int, dp, dp, dp]),
        "pn_volume_averages": (C.c_int, [C.c_void_p, P(PnAverages)]),
        "pn_extrema_get": (C.c_int, [C.c_void_p, P(PnExtrema)]),
        "pn_volume_analysis": (C.c_int, [C.c_void_p, P(PnAverages), P(PnExtrema)]),
        "pn_volume_averages_synced": (C.c_int, [C.c_void_p, C.c_double, P(PnAverages)]),
        "pn_volume_analysis_noalloc": (C.c_int, [P(PnParams), C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]),
        "pn_spectrum_nbins": (C.c_int, [C.c_int, C.c_double]),
        "pn_spectrum_get": (C.c_int, [C.c_void_p, P(PnSpectrum)]),
        "pn_fft_modes": (C.c_int, [C.c_void_p, C.c_int, dp]),
        "pn_islands": (C.c_int, [C.c_void_p, C.c_int, C.c_double, P(PnCatalog)]),
        "pn_islands_lookup": (C.c_int, [C.c_void_p, C.c_int, P(C.c_longlong), P(C.c_int)]),
        "pn_host_alloc": (C.c_int, [C.c_size_t, P(C.c_void_p)]),
        "pn_host_free": (C.c_int, [C.c_void_p]),
        "pn_tracker_create": (C.c_int, [P(C.c_void_p)]),
        "pn_tracker_destroy": (C.c_int, [C.c_void_p]),
        "pn_tracker_last_error": (C.c_char_p, [C.c_void_p]),
        "pn_tracker_update": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                        C.c_char_p]),
        "pn_tracker_count": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "pn_tracker_catalog": (C.c_int, [C.c_void_p, P(PnCatalog)]),
        "pn_tracker_get": (C.c_int, [C.c_void_p, C.c_int, P(PnTrackedIsland)]),
        "pn_energy_density": (C.c_int, [C.c_void_p, dp]),
        "pn_field_diff_get": (C.c_int, [C.c_void_p, C.c_void_p, P(PnFieldDiff)]),
        "pn_ic_modes": (C.c_int, [C.c_int, C.c_double, P(PnIcParams), dp, dp]),
        "pn_initial_conditions": (C.c_int, [C.c_void_p, P(PnIcParams)]),
        "pn_ic_modes_device": (C.c_int, [C.c_void_p, P(PnIcParams), dp, dp]),
        "pn_set_couplings": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double]),
        "pn_batch_create": (C.c_int, [C.c_int, P(PnRealization), C.c_int, C.c_int, P(C.c_void_p)]),
        "pn_batch_destroy": (C.c_int, [C.c_void_p]),
        "pn_batch_last_error": (C.c_char_p, [C.c_void_p]),
        "pn_batch_view": (C.c_int, [C.c_void_p, C.c_int, P(C.c_void_p)]),
        "pn_batch_evolve": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_int]),
        "pn_batch_evolve_timed": (C.c_int, [C.c_void_p, C.c_int, dp, dp, C.c_int, dp]),
        "pn_batch_sync": (C.c_int, [C.c_void_p]),
        "pn_batch_set_decayed": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "pn_batch_set_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
        "pn_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int)]),
        "pn_batch_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "pn_batch_get_option": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int)]),
        "pn_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "pn_get_option": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int)]),
        "pn_forward_observable": (C.c_int, [P(PnAdjParams), dp, dp, C.c_int, dp, dp, dp]),
        "pn_adjoint": (C.c_int, [P(PnAdjParams), dp, dp, C.c_int, dp, dp, dp, dp, P(PnAdjResult)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class PhiNumError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("phinum_gpu error %d: %s" % (code, msg))
        self.code = code


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class Lattice:
    """One lattice context on one GPU (pn_create ... pn_destroy).

    Mirrors LocalLattice (lattice.cpp:8-118) plus the hot-path functions of
    dynamics.cpp, fieldanalysis.cpp, momentumspace.cpp and islands.cpp.
    """

    def __init__(self, N, a=0.5, mass_mode=1, m2=1.0, lambda_mode=-1, kappa=2.0,
                 precision=PN_FP64, device=0):
        self.L = lib()
        self.N = int(N)
        self.a = float(a)
        self.params = PnParams(int(N), float(a), int(mass_mode), float(m2), int(lambda_mode),
                               float(kappa), int(precision), int(device))
        h = C.c_void_p()
        rc = self.L.pn_create(C.byref(self.params), C.byref(h))
        if rc != PN_OK:
            raise PhiNumError(rc, self.L.pn_last_error(None).decode())
        self.h = h
        self._owned = True

    @classmethod
    def _borrowed(cls, handle, N, a):
        """Wraps a context owned elsewhere (a batch view); close() leaves it."""
        obj = cls.__new__(cls)
        obj.L = lib()
        obj.N = int(N)
        obj.a = float(a)
        obj.h = handle
        obj._owned = False
        return obj

    def _chk(self, rc):
        if rc != PN_OK:
            raise PhiNumError(rc, self.L.pn_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "_hrows", None):
            self.L.pn_host_free(self._hrows)
            self._hrows = None
            self._hrows_cap = 0
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.L.pn_destroy(self.h)
            self.h = None

    def _host_rows(self, capacity):
        """A page-locked row array (pn_host_alloc) kept across islands calls."""
        if getattr(self, "_hrows_cap", 0) < capacity:
            if getattr(self, "_hrows", None):
                self.L.pn_host_free(self._hrows)
                self._hrows = None
                self._hrows_cap = 0
            p = C.c_void_p()
            self._chk(self.L.pn_host_alloc(C.sizeof(PnIsland) * capacity, C.byref(p)))
            self._hrows, self._hrows_cap = p, capacity
        return C.cast(self._hrows, C.POINTER(PnIsland))

    def two_step(self, on=None):
        """Gets or sets PN_OPT_TWO_STEP (pairs of steps per pass over HBM)."""
        if on is not None:
            self._chk(self.L.pn_set_option(self.h, PN_OPT_TWO_STEP, int(bool(on))))
        v = C.c_int()
        self._chk(self.L.pn_get_option(self.h, PN_OPT_TWO_STEP, C.byref(v)))
        return bool(v.value)

    def set_couplings(self, mass_mode, m2, lambda_mode, kappa):
        self._chk(self.L.pn_set_couplings(self.h, int(mass_mode), float(m2), int(lambda_mode),
                                          float(kappa)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # fields ---------------------------------------------------------------
    def set_fields(self, phi=None, pi=None, layout=PN_LAYOUT_DENSE):
        phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        pi = None if pi is None else np.ascontiguousarray(pi, dtype=np.float64)
        n = self.N ** 3 if layout == PN_LAYOUT_DENSE else self.N ** 2 * (self.N + 2)
        for arr in (phi, pi):
            if arr is not None and arr.size != n:
                raise ValueError("field has %d values, expected %d" % (arr.size, n))
        self._chk(self.L.pn_set_fields(self.h, _dptr(phi), _dptr(pi), layout))

    def get_fields(self, layout=PN_LAYOUT_DENSE):
        n = self.N ** 3 if layout == PN_LAYOUT_DENSE else self.N ** 2 * (self.N + 2)
        phi = np.empty(n)
        pi = np.empty(n)
        self._chk(self.L.pn_get_fields(self.h, _dptr(phi), _dptr(pi), layout))
        return phi, pi

    # dynamics -------------------------------------------------------------
    def kick(self, dtk):
        self._chk(self.L.pn_kick(self.h, float(dtk)))

    def drift(self, dtd):
        self._chk(self.L.pn_drift(self.h, float(dtd)))

    def evolve(self, dtk, dtd, sync=True):
        dtk = np.ascontiguousarray(dtk, dtype=np.float64)
        dtd = np.ascontiguousarray(dtd, dtype=np.float64)
        if dtk.size != dtd.size:
            raise ValueError("dtk and dtd differ in length")
        self._chk(self.L.pn_evolve(self.h, int(dtk.size), _dptr(dtk), _dptr(dtd)))
        if sync:
            self.sync()

    def evolve_timed(self, dtk, dtd):
        """Runs the steps between two HIP events on the context's stream; returns ms."""
        dtk = np.ascontiguousarray(dtk, dtype=np.float64)
        dtd = np.ascontiguousarray(dtd, dtype=np.float64)
        ms = C.c_double()
        self._chk(self.L.pn_evolve_timed(self.h, int(dtk.size), _dptr(dtk), _dptr(dtd), C.byref(ms)))
        return ms.value

    def sync(self):
        self._chk(self.L.pn_sync(self.h))

    # analysis -------------------------------------------------------------
    def volume_averages(self):
        out = PnAverages()
        self._chk(self.L.pn_volume_averages(self.h, C.byref(out)))
        return out

    def extrema(self):
        out = PnExtrema()
        self._chk(self.L.pn_extrema_get(self.h, C.byref(out)))
        return out

    def volume_analysis(self):
        """(averages, extrema) from one pass over the fields, as
        Print_AveragesExtrema calls the two back to back
        (fieldanalysis.cpp:319-397,399-699)."""
        avg, ext = PnAverages(), PnExtrema()
        self._chk(self.L.pn_volume_analysis(self.h, C.byref(avg), C.byref(ext)))
        return avg, ext

    def volume_averages_synced(self, half_dt):
        """Averages with Pi advanced by a kick of half_dt to Phi's time (the
        time-synchronised energy; pn_volume_averages_synced)."""
        out = PnAverages()
        self._chk(self.L.pn_volume_averages_synced(self.h, float(half_dt), C.byref(out)))
        return out

    def spectrum(self):
        nb = self.L.pn_spectrum_nbins(self.N, self.a)
        counts = np.zeros(nb, dtype=np.int32)
        f = np.zeros(nb)
        k = np.zeros(nb)
        q = np.zeros(nb)
        s = PnSpectrum(nb, 0.0, counts.ctypes.data_as(C.POINTER(C.c_int)), _dptr(f), _dptr(k), _dptr(q))
        self._chk(self.L.pn_spectrum_get(self.h, C.byref(s)))
        return {"nbins": s.nbins, "binsize": s.binsize, "counts": counts, "F": f, "K": k, "Q": q}

    def fft_modes(self, field):
        H = self.N // 2 + 1
        out = np.empty(2 * self.N * self.N * H)
        self._chk(self.L.pn_fft_modes(self.h, int(field), _dptr(out)))
        return out.view(np.complex128).reshape(self.N, self.N, H)

    def islands(self, mode, thr, capacity=None, labels=False, grow=True, as_array=False, view=False):
        """Island catalog (rows of PnIsland, or with as_array one numpy
        structured array), the label array if asked for, and the threshold
        used.  The rows arrive in a page-locked array kept by the lattice;
        they are returned as copies, or with as_array and view as a numpy
        view of that array, valid until the next islands call.  capacity
        defaults to the last catalog's size plus a quarter (at least 65536).
        With grow, a catalog larger than `capacity` is fetched again with
        room for all of it (the library reports the count with
        PN_ERR_CAPACITY); without, that error is raised."""
        if capacity is None:
            capacity = getattr(self, "_isl_cap", 65536)
        while True:
            arr = self._host_rows(max(capacity, 1))
            lab = np.empty(self.N ** 3, dtype=np.int32) if labels else None
            cat = PnCatalog(capacity, 0, 0.0, arr,
                            lab.ctypes.data_as(C.POINTER(C.c_int)) if labels else None)
            rc = self.L.pn_islands(self.h, int(mode), float(thr), C.byref(cat))
            if rc == PN_ERR_CAPACITY and grow and cat.count > capacity:
                capacity = cat.count
                continue
            self._chk(rc)
            self._isl_cap = max(65536, cat.count + cat.count // 4)
            if as_array:
                rows = np.ctypeslib.as_array(arr, shape=(max(cat.count, 1),))[:cat.count]
                if not view:
                    rows = rows.copy()
            else:
                rows = [PnIsland.from_buffer_copy(arr[i]) for i in range(cat.count)]
            return rows, lab, cat.threshold

    def initial_conditions(self, ic):
        """SetInitialConditions (initialize==3) with a PnIcParams."""
        self._chk(self.L.pn_initial_conditions(self.h, C.byref(ic)))

    def three_step(self, on=None, zblock=None):
        """Gets or sets PN_OPT_THREE_STEP (triples of steps per pass over HBM):
        True, False, or "auto" (chosen by timing, the default); optionally its
        z block (PN_OPT_STEP3_ZBLOCK, 0 = by timing).  Returns whether
        evolve advances triples."""
        if zblock is not None:
            self._chk(self.L.pn_set_option(self.h, PN_OPT_STEP3_ZBLOCK, int(zblock)))
        if on is not None:
            v = -1 if on == "auto" else int(bool(on))
            self._chk(self.L.pn_set_option(self.h, PN_OPT_THREE_STEP, v))
        v = C.c_int()
        self._chk(self.L.pn_get_option(self.h, PN_OPT_THREE_STEP, C.byref(v)))
        return bool(v.value)

    def kernel_timed(self):
        """True once the timed choice of the multi-step kernel has been made
        (PN_OPT_KERNEL_TIMED)."""
        v = C.c_int()
        self._chk(self.L.pn_get_option(self.h, PN_OPT_KERNEL_TIMED, C.byref(v)))
        return bool(v.value)

    def step2_zblock(self, zb=None):
        """Gets or sets PN_OPT_STEP2_ZBLOCK (planes per two-step tile; 0: chosen
        by timing on the next evolve calls of at least 12 steps)."""
        if zb is not None:
            self._chk(self.L.pn_set_option(self.h, PN_OPT_STEP2_ZBLOCK, int(zb)))
        v = C.c_int()
        self._chk(self.L.pn_get_option(self.h, PN_OPT_STEP2_ZBLOCK, C.byref(v)))
        return v.value

    def ic_on_device(self, on=None):
        """Gets or sets PN_OPT_IC_ON_DEVICE (modes drawn on the device, the
        default, or on the host)."""
        if on is not None:
            self._chk(self.L.pn_set_option(self.h, PN_OPT_IC_ON_DEVICE, int(bool(on))))
        v = C.c_int()
        self._chk(self.L.pn_get_option(self.h, PN_OPT_IC_ON_DEVICE, C.byref(v)))
        return bool(v.value)

    def ic_modes_device(self, ic):
        """(A, E) modes of ic_modes drawn on this lattice's device."""
        N = self.N
        H = N // 2 + 1
        A = np.empty(2 * N * N * H)
        E = np.empty(2 * N * N * H)
        self._chk(self.L.pn_ic_modes_device(self.h, C.byref(ic), _dptr(A), _dptr(E)))
        return A.view(np.complex128).reshape(N, N, H), E.view(np.complex128).reshape(N, N, H)

    def field_diff(self, other):
        """Field error of this lattice against `other` (same N, either
        precision): maxima, and relative L2 errors sqrt(sum d^2 / sum b^2)."""
        d = PnFieldDiff()
        self._chk(self.L.pn_field_diff_get(self.h, other.h, C.byref(d)))
        out = {n: getattr(d, n) for n, _ in PnFieldDiff._fields_}
        out["l2_dphi"] = float(np.sqrt(d.sum_dphi2 / d.sum_phi2)) if d.sum_phi2 > 0 else float(d.sum_dphi2 > 0)
        out["l2_dpi"] = float(np.sqrt(d.sum_dpi2 / d.sum_pi2)) if d.sum_pi2 > 0 else float(d.sum_dpi2 > 0)
        return out

    def energy_density(self):
        rho = np.empty(self.N ** 3)
        self._chk(self.L.pn_energy_density(self.h, _dptr(rho)))
        return rho


def ic_params(seed, a=0.5, Lambda=6.0, noise=0.5, m2=1.0, phi0=0.0, pi0=0.0, QS=None,
              scale2=None, amplitude=0.0, justnoise=0):
    """initialize==3 parameters; QS and scale2 default to half the lattice
    cutoff pTmax/2 = sqrt(3)/a (momentumspace.cpp:41); noise is multiplied by
    Lambda as CompleteParameters does (parameters.cpp:408-410)."""
    if QS is None:
        QS = float(np.sqrt(3.0 * (4.0 / (a * a)))) / 2.0
    if scale2 is None:
        scale2 = QS
    return PnIcParams(int(seed), float(QS), float(scale2), float(amplitude), float(noise) * float(Lambda),
                      float(m2), int(justnoise), float(phi0), float(pi0))


def ic_modes(N, a, ic):
    """Host part of SetInitialConditions: (A, E) modes in [pz][px][py] order."""
    H = N // 2 + 1
    A = np.empty(2 * N * N * H)
    E = np.empty(2 * N * N * H)
    rc = lib().pn_ic_modes(int(N), float(a), C.byref(ic), _dptr(A), _dptr(E))
    if rc != PN_OK:
        raise PhiNumError(rc, lib().pn_last_error(None).decode())
    return A.view(np.complex128).reshape(N, N, H), E.view(np.complex128).reshape(N, N, H)


def time_factors(T0, dtbase, a, nsteps):
    """Host replay of the leapfrog factors (reset/evolve, dynamics.cpp:28-50,196-237).

    Returns (dtk, dtd, tau_after) with dtk[n] = tau - Last_Momenta_update at the
    kick of step n and dtd[n] = tau - Last_Field_update at its drift.
    """
    tau = float(T0)
    dtau = float(dtbase) * float(a)
    last_f = tau
    last_m = tau
    dtk = np.empty(nsteps)
    dtd = np.empty(nsteps)
    taus = np.empty(nsteps)
    half = 0.5 * dtau
    for n in range(nsteps):
        tau = tau + half
        dtk[n] = tau - last_m
        last_m = tau
        tau = tau + half
        dtd[n] = tau - last_f
        last_f = tau
        taus[n] = tau
    return dtk, dtd, taus


class Tracker:
    """Island tracking over time (Full_Analysis called repeatedly,
    islands.cpp:788-835) on the GPU catalog."""

    STATUS = {0: "island", 1: "expanding bubble", 2: "oscillon"}

    def __init__(self):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.pn_tracker_create(C.byref(h))
        if rc != PN_OK:
            raise PhiNumError(rc, "pn_tracker_create failed")
        self.h = h

    def update(self, lattice, mode, thr, tau, outdir=None):
        rc = self.L.pn_tracker_update(self.h, lattice.h, int(mode), float(thr), float(tau),
                                      outdir.encode() if outdir else None)
        if rc != PN_OK:
            raise PhiNumError(rc, self.L.pn_tracker_last_error(self.h).decode())

    def catalog(self):
        """The components of the last update as pn_islands numbers them
        (pn_tracker_catalog), a numpy structured array of PnIsland."""
        cat = PnCatalog(0, 0, 0.0, None, None)
        self.L.pn_tracker_catalog(self.h, C.byref(cat))  # the count
        arr = (PnIsland * max(cat.count, 1))()
        cat = PnCatalog(cat.count, 0, 0.0, arr, None)
        rc = self.L.pn_tracker_catalog(self.h, C.byref(cat))
        if rc != PN_OK:
            raise PhiNumError(rc, self.L.pn_tracker_last_error(self.h).decode())
        return np.ctypeslib.as_array(arr)[:cat.count]

    def islands(self):
        """[(index, status, disp, [(at_time, peak pos, peak value, avg phi, volume, PotEn,
        KinEn, GradEn), ...])] in list order."""
        n = C.c_int()
        self.L.pn_tracker_count(self.h, C.byref(n))
        out = []
        for i in range(n.value):
            t = PnTrackedIsland()
            self.L.pn_tracker_get(self.h, i, C.byref(t))
            hist = [(e.at_time, int(e.center_pos), e.peak_value, e.avg_phi, e.volume, e.pot_en,
                     e.kin_en, e.grad_en) for e in t.hist[:t.nhist]]
            out.append((int(t.index), int(t.status), int(t.disp), hist))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.pn_tracker_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many lattices stepped together (pn_batch_*); views give per-realization
    access through the Lattice interface."""

    def __init__(self, realizations, precision=PN_FP64, device=0):
        self.L = lib()
        self.reals = [dict(r) for r in realizations]
        arr = (PnRealization * len(self.reals))()
        for i, r in enumerate(self.reals):
            arr[i] = PnRealization(int(r["N"]), float(r.get("a", 0.5)), int(r.get("mass_mode", 1)),
                                   float(r.get("m2", 1.0)), int(r.get("lambda_mode", -1)),
                                   float(r["kappa"]))
        h = C.c_void_p()
        rc = self.L.pn_batch_create(len(self.reals), arr, int(precision), int(device), C.byref(h))
        if rc != PN_OK:
            raise PhiNumError(rc, self.L.pn_batch_last_error(None).decode())
        self.h = h
        self.R = len(self.reals)

    def _chk(self, rc):
        if rc != PN_OK:
            raise PhiNumError(rc, self.L.pn_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.pn_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def view(self, r):
        v = C.c_void_p()
        self._chk(self.L.pn_batch_view(self.h, int(r), C.byref(v)))
        rr = self.reals[r]
        return Lattice._borrowed(v, rr["N"], rr.get("a", 0.5))

    def _tables(self, dtk, dtd):
        dtk = np.ascontiguousarray(dtk, dtype=np.float64)
        dtd = np.ascontiguousarray(dtd, dtype=np.float64)
        if dtk.ndim == 1:
            dtk = np.ascontiguousarray(np.broadcast_to(dtk, (self.R, dtk.size)))
            dtd = np.ascontiguousarray(np.broadcast_to(dtd, (self.R, dtd.size)))
        if dtk.shape != dtd.shape or dtk.shape[0] != self.R:
            raise ValueError("dtk/dtd must be [R][nsteps]")
        return dtk, dtd

    def evolve(self, dtk, dtd, use_graph=True, sync=True):
        dtk, dtd = self._tables(dtk, dtd)
        self._chk(self.L.pn_batch_evolve(self.h, int(dtk.shape[1]), _dptr(dtk), _dptr(dtd),
                                         int(bool(use_graph))))
        if sync:
            self._chk(self.L.pn_batch_sync(self.h))

    def evolve_timed(self, dtk, dtd, use_graph=True):
        dtk, dtd = self._tables(dtk, dtd)
        ms = C.c_double()
        self._chk(self.L.pn_batch_evolve_timed(self.h, int(dtk.shape[1]), _dptr(dtk), _dptr(dtd),
                                               int(bool(use_graph)), C.byref(ms)))
        return ms.value

    def set_decayed(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        self._chk(self.L.pn_batch_set_decayed(self.h, f.ctypes.data_as(C.POINTER(C.c_int))))

    def set_tiles(self, N, rows, zblock):
        self._chk(self.L.pn_batch_set_tiles(self.h, int(N), int(rows), int(zblock)))

    def two_step(self, on=None):
        """Gets or sets PN_OPT_TWO_STEP (pairs of steps per launch) for the batch."""
        if on is not None:
            self._chk(self.L.pn_batch_set_option(self.h, PN_OPT_TWO_STEP, int(bool(on))))
        v = C.c_int()
        self._chk(self.L.pn_batch_get_option(self.h, PN_OPT_TWO_STEP, C.byref(v)))
        return bool(v.value)

    def three_step(self, on=None):
        """Gets or sets PN_OPT_THREE_STEP (three steps per launch; on by default
        when every realization fits a three-step tile) for the batch."""
        if on is not None:
            self._chk(self.L.pn_batch_set_option(self.h, PN_OPT_THREE_STEP, int(bool(on))))
        v = C.c_int()
        self._chk(self.L.pn_batch_get_option(self.h, PN_OPT_THREE_STEP, C.byref(v)))
        return bool(v.value)

    def autotune(self, trial_steps=10):
        out = np.zeros(3 * 64, dtype=np.int32)
        n = C.c_int()
        self._chk(self.L.pn_batch_autotune(self.h, int(trial_steps),
                                           out.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n)))
        return [tuple(int(v) for v in out[3 * i:3 * i + 3]) for i in range(n.value)]


def adj_params(N, c2=1.0, c4=-1.0, c6=0.2, a=0.5, threshold=4.0, width=0.1, checkpoint_every=100,
               device=0):
    """Couplings of the differentiable leapfrog: V = c2 phi^2/2 + c4 phi^4/4 + c6 phi^6/6
    (physical units of V = phi^2/2 - lambda phi^4/4 + g phi^6/6: c4 = -lambda, c6 = g)."""
    return PnAdjParams(int(N), float(a), float(c2), float(c4), float(c6), float(threshold),
                       float(width), int(checkpoint_every), int(device))


def forward_observable(p, phi0, pi0, dtk, dtd):
    phi0 = np.ascontiguousarray(phi0, dtype=np.float64)
    pi0 = np.ascontiguousarray(pi0, dtype=np.float64)
    dtk = np.ascontiguousarray(dtk, dtype=np.float64)
    dtd = np.ascontiguousarray(dtd, dtype=np.float64)
    F = C.c_double()
    rc = lib().pn_forward_observable(C.byref(p), _dptr(phi0), _dptr(pi0), int(dtk.size), _dptr(dtk),
                                     _dptr(dtd), C.byref(F))
    if rc != PN_OK:
        raise PhiNumError(rc, lib().pn_last_error(None).decode())
    return F.value


def adjoint(p, phi0, pi0, dtk, dtd):
    """Reverse entry point: returns (F, dF/dphi0, dF/dpi0, dF/dc (3,), result struct)."""
    phi0 = np.ascontiguousarray(phi0, dtype=np.float64)
    pi0 = np.ascontiguousarray(pi0, dtype=np.float64)
    dtk = np.ascontiguousarray(dtk, dtype=np.float64)
    dtd = np.ascontiguousarray(dtd, dtype=np.float64)
    gphi = np.empty_like(phi0)
    gpi = np.empty_like(pi0)
    res = PnAdjResult()
    rc = lib().pn_adjoint(C.byref(p), _dptr(phi0), _dptr(pi0), int(dtk.size), _dptr(dtk), _dptr(dtd),
                          _dptr(gphi), _dptr(gpi), C.byref(res))
    if rc != PN_OK:
        raise PhiNumError(rc, lib().pn_last_error(None).decode())
    return res.F, gphi, gpi, np.array(list(res.grad_c)), res
