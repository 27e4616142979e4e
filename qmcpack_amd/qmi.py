"""ctypes binding of libqmcmi.so, the MI355X-native batched walker-move path.

The method names mirror the reference's multi-walker API (TrialWaveFunction / ParticleSet
mw_* functions, QMCPACK src/QMCWaveFunctions/TrialWaveFunction.cpp and
src/Particle/ParticleSet.cpp) so tests read like the reference's own.  There is no CPU
fallback: loading fails loudly when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqmcmi.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "qmcmi.h")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DmcParams(C.Structure):
    """qmi_dmc_params (include/qmcmi.h)."""
    _fields_ = [("tau", C.c_double), ("feedback", C.c_double), ("sigma_bound", C.c_double),
                ("warmup_steps", C.c_int), ("refenergy_history", C.c_int), ("cutoff_scheme", C.c_int),
                ("capacity", C.c_int), ("target_walkers", C.c_int), ("recompute_period", C.c_int),
                ("seed", C.c_uint32), ("nonlocal_moves", C.c_int)]


class QmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qmcmi error {code}: {msg}")
        self.code = code


class SystemDesc(C.Structure):
    """Mirror of qmi_system_desc (include/qmcmi.h)."""
    _fields_ = [
        ("L", C.c_double * 3),
        ("n_up", C.c_int), ("n_dn", C.c_int),
        ("n_ion", C.c_int), ("ions", _dp),
        ("M", C.c_int * 3), ("norb", C.c_int), ("coefs", _dp),
        ("j1_np", C.c_int), ("j1_params", _dp), ("j1_rcut", C.c_double), ("j1_cusp", C.c_double),
        ("j2_np", C.c_int), ("j2_uu", _dp), ("j2_ud", _dp), ("j2_rcut", C.c_double),
        ("j2_cusp_uu", C.c_double), ("j2_cusp_ud", C.c_double),
        ("delay_rank", C.c_int),
        ("complex_orbitals", C.c_int), ("kpoints", _dp),
    ]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build the HIP library first (python -c "
                          f"'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    L.qmi_last_error.restype = C.c_char_p
    L.qmi_last_error.argtypes = [C.c_void_p]
    L.qmi_batch_last_error.restype = C.c_char_p
    L.qmi_batch_last_error.argtypes = [C.c_void_p]
    L.qmi_checkpoint_save.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.qmi_checkpoint_load.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
    _lib = L
    return L


LR_HANDLERS = {"ewald": 0, "opt_breakup": 1}


def lr_breakup(L, dim_cutoff=15.0):
    """The optimized breakup of an orthorhombic cell (host computation, no device): the 45 basis
    coefficients and {r_c, v_l(r=0), v_s(k=0), chi^2} (LRHandlerTemp::InitBreakup)."""
    lib = load_library()
    Lv = np.ascontiguousarray(L, dtype=np.float64)
    t = np.zeros(45)
    info = np.zeros(4)
    code = lib.qmi_lr_breakup(_p(Lv), C.c_double(dim_cutoff), _p(t), C.c_int(45), _p(info))
    if code != 0:
        raise QmiError(code, "qmi_lr_breakup: invalid cell or cutoff")
    return t, dict(rc=info[0], vlr_r0=info[1], vsr_k0=info[2], chi2=info[3])


def exported_symbols_from_header(header: str = HEADER) -> list[str]:
    import re
    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(qmi_\w+)\s*\(", txt, flags=re.M)))


# ---------------------------------------------------------------- restart files
CKPT_MAGIC = b"QMICKPT1"
_CKPT_TYPES = {1: np.dtype("<f8"), 2: np.dtype("<i8"), 3: np.dtype("<u4")}


def read_checkpoint(data: bytes) -> dict:
    """The datasets of a checkpoint (qmi_checkpoint_save, include/qmcmi.h) by name, as numpy arrays:
    state_0/walkers [nw][N][3], state_0/walker_weights, random/std::mt19937 [streams][625], ..."""
    import struct
    mv = memoryview(data)
    if len(mv) < 16 or bytes(mv[:8]) != CKPT_MAGIC:
        raise ValueError("not a walker checkpoint")
    version, count = struct.unpack_from("<II", mv, 8)
    if version != 1:
        raise ValueError(f"unsupported checkpoint format version {version}")
    off, out = 16, {}
    for _ in range(count):
        (n,) = struct.unpack_from("<I", mv, off)
        name = bytes(mv[off + 4:off + 4 + n]).decode()
        off += 4 + n
        t, rank = struct.unpack_from("<II", mv, off)
        off += 8
        dims = struct.unpack_from(f"<{rank}Q", mv, off)
        off += 8 * rank
        dt = _CKPT_TYPES[t]
        size = int(np.prod(dims, dtype=np.int64)) * dt.itemsize
        if off + size > len(mv):
            raise ValueError("truncated checkpoint")
        out[name] = np.frombuffer(mv[off:off + size], dtype=dt).reshape(dims).copy()
        off += size
    if off != len(mv):
        raise ValueError("trailing bytes after the checkpoint's datasets")
    return out


def write_checkpoint(datasets: dict) -> bytes:
    """Inverse of read_checkpoint (float64, int64 and uint32 arrays)."""
    import struct
    code = {np.dtype("float64"): 1, np.dtype("int64"): 2, np.dtype("uint32"): 3}
    parts = [CKPT_MAGIC, struct.pack("<II", 1, len(datasets))]
    for name, a in datasets.items():
        a = np.ascontiguousarray(a)
        nb = name.encode()
        parts += [struct.pack("<I", len(nb)), nb, struct.pack("<II", code[a.dtype], a.ndim),
                  struct.pack(f"<{a.ndim}Q", *a.shape), a.astype(a.dtype.newbyteorder("<")).tobytes()]
    return b"".join(parts)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def make_desc(system, keep: list) -> SystemDesc:
    def arr(x):
        if x is None:
            return None
        a = np.ascontiguousarray(x, dtype=np.float64)
        keep.append(a)
        return _p(a)

    d = SystemDesc()
    for i in range(3):
        d.L[i] = float(system.L[i])
    d.n_up, d.n_dn = int(system.n_up), int(system.n_dn)
    ions = np.asarray(system.ions, dtype=np.float64).reshape(-1, 3)
    d.n_ion = ions.shape[0]
    d.ions = arr(ions) if d.n_ion else None
    has_det = system.coefs is not None and system.norb > 0
    for i in range(3):
        d.M[i] = int(system.M[i]) if has_det else 0
    d.norb = int(system.norb) if has_det else 0
    d.coefs = arr(system.coefs) if has_det else None
    d.j1_np = len(system.j1_params) if system.j1_params is not None else 0
    d.j1_params = arr(system.j1_params)
    d.j1_rcut, d.j1_cusp = float(system.j1_rcut), float(system.j1_cusp)
    d.j2_np = len(system.j2_uu) if system.j2_uu is not None else 0
    d.j2_uu, d.j2_ud = arr(system.j2_uu), arr(system.j2_ud)
    d.j2_rcut = float(system.j2_rcut)
    d.j2_cusp_uu, d.j2_cusp_ud = float(system.j2_cusp_uu), float(system.j2_cusp_ud)
    d.delay_rank = int(system.delay_rank)
    kp = getattr(system, "kpoints", None)
    d.complex_orbitals = int(kp is not None and d.norb > 0)
    d.kpoints = arr(kp) if d.complex_orbitals else None
    return d


# storage precisions (include/qmcmi.h)
PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16, PREC_TABLE_FP32, PREC_TABLE_BF16 = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_MIXED_FP32, "bf16": PREC_MIXED_BF16, "t32": PREC_TABLE_FP32,
              "tbf16": PREC_TABLE_BF16}


class Context:
    """One qmcmi context: one device stream, one system, one walker population.

    precision: PREC_FP64 (default), PREC_MIXED_FP32 or PREC_MIXED_BF16, or their names."""

    def __init__(self, device: int = 0, precision=PREC_FP64):
        self.L = load_library()
        self.h = C.c_void_p()
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        st = self.L.qmi_ctx_create(C.byref(self.h), device, self.precision)
        if st != 0:
            raise QmiError(st, "qmi_ctx_create failed (no HIP device?)")
        self.nw = 0
        self.N = 0
        self.cplx = False

    def _vals(self, *shape):
        """Host buffer for value-typed outputs: (re, im) pairs for complex orbitals."""
        return np.empty(shape + ((2,) if self.cplx else ()))

    def _out(self, a):
        return a[..., 0] + 1j * a[..., 1] if self.cplx else a

    def _vin(self, a):
        """Value-typed input: complex arrays as (re, im) pairs."""
        if self.cplx:
            a = np.asarray(a, dtype=np.complex128)
            return np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1))
        return np.ascontiguousarray(a, dtype=np.float64)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.qmi_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise QmiError(st, self.L.qmi_last_error(self.h).decode())

    # ---------------------------------------------------------------- setup
    def system_create(self, system):
        keep = []
        d = make_desc(system, keep)
        self._check(self.L.qmi_system_create(self.h, C.byref(d)))
        self.N = system.n_up + system.n_dn
        self.n_up = system.n_up
        self.cplx = getattr(system, "kpoints", None) is not None and system.norb > 0

    def walkers_load(self, R, walkers_per_crowd: int = 1):
        R = np.ascontiguousarray(R, dtype=np.float64)
        self.nw = R.shape[0]
        self._check(self.L.qmi_walkers_load(self.h, self.nw, _p(R), int(walkers_per_crowd)))

    def walkers_get(self):
        R = np.empty((self.nw, self.N, 3))
        self._check(self.L.qmi_walkers_get(self.h, _p(R)))
        return R

    # ---------------------------------------------------------------- multi-walker API
    def mw_evaluateLog(self):
        """log psi (log|psi|, or log|psi| + i phase for complex orbitals), G, L (value type)."""
        lp, G, L = self._vals(self.nw), self._vals(self.nw, self.N, 3), self._vals(self.nw, self.N)
        self._check(self.L.qmi_mw_evaluate_log(self.h, _p(lp), _p(G), _p(L)))
        return self._out(lp), self._out(G), self._out(L)

    def mw_evalGrad(self, iat):
        g = self._vals(self.nw, 3)
        self._check(self.L.qmi_mw_eval_grad(self.h, int(iat), _p(g)))
        return self._out(g)

    def mw_makeMove(self, iat, displ):
        displ = np.ascontiguousarray(displ, dtype=np.float64).reshape(self.nw, 3)
        self._check(self.L.qmi_mw_make_move(self.h, int(iat), _p(displ)))

    def mw_calcRatioGrad(self, iat):
        r, g = self._vals(self.nw), self._vals(self.nw, 3)
        self._check(self.L.qmi_mw_ratio_grad(self.h, int(iat), _p(r), _p(g)))
        return self._out(r), self._out(g)

    def mw_ratioGrad(self, iat, grad_new):
        """WaveFunctionComponent::mw_ratioGrad of this context's components: returns the ratios and
        adds their gradient at the proposed position to grad_new [nw][3] in place."""
        r = self._vals(self.nw)
        g = self._vin(grad_new)
        self._check(self.L.qmi_mw_component_ratio_grad(self.h, int(iat), _p(r), _p(g)))
        grad_new[...] = self._out(g)
        return self._out(r)

    def mw_evaluateVGLandDetRatioGrads(self, pos, inv_rows, with_vgl=True):
        """SPOSet::mw_evaluateVGLandDetRatioGrads with caller inverse rows inv_rows [nw][n]:
        (phi_vgl [5][nw][n] or None, ratios [nw], grads [nw][3])."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nw, 3)
        n = np.asarray(inv_rows).shape[1]
        inv = self._vin(inv_rows)
        phi = self._vals(5, self.nw, n) if with_vgl else None
        r, g = self._vals(self.nw), self._vals(self.nw, 3)
        self._check(self.L.qmi_mw_spo_vgl_det_ratio_grads(self.h, _p(pos), int(n), _p(inv), _p(phi), _p(r), _p(g)))
        return (self._out(phi) if with_vgl else None), self._out(r), self._out(g)

    def distance_table(self, kind):
        """A stateful DistanceTable over this context's walkers: kind "AA" (electron pairs) or
        "AB" (electrons to ions)."""
        return DistanceTable(self, kind)

    def mw_accept_rejectMove(self, iat, accepted, safe_to_delay=True):
        acc = np.ascontiguousarray(np.asarray(accepted, dtype=np.int32).reshape(self.nw))
        self._check(self.L.qmi_mw_accept_reject(self.h, int(iat), acc.ctypes.data_as(_ip), int(bool(safe_to_delay))))

    def mw_completeUpdates(self):
        self._check(self.L.qmi_mw_complete_updates(self.h))

    def mw_evaluateGL(self, fromscratch=False):
        lp, G, L, ek = self._vals(self.nw), self._vals(self.nw, self.N, 3), self._vals(self.nw, self.N), np.empty(self.nw)
        self._check(self.L.qmi_mw_evaluate_gl(self.h, int(bool(fromscratch)), _p(lp), _p(G), _p(L), _p(ek)))
        return self._out(lp), self._out(G), self._out(L), ek

    def mw_evaluateVGL(self, pos, n):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nw, 3)
        out = self._vals(self.nw, 5, n)
        self._check(self.L.qmi_mw_spo_vgl(self.h, _p(pos), int(n), _p(out)))
        return self._out(out)

    def mw_evaluateDetRatios(self, pos, inv_row):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        nvp = pos.shape[1]
        n = np.asarray(inv_row).shape[1]
        inv_row = self._vin(inv_row)
        out = self._vals(self.nw, nvp)
        self._check(self.L.qmi_mw_det_ratios(self.h, nvp, _p(pos), n, _p(inv_row), _p(out)))
        return self._out(out)

    def mw_evaluateRatios(self, iat, pos):
        """TrialWaveFunction::mw_evaluateRatios over virtual positions pos [nw][nvp][3] of electron
        iat -> ratios [nw][nvp] (complex for complex orbitals)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        nvp = pos.shape[1]
        out = self._vals(self.nw, nvp)
        self._check(self.L.qmi_mw_evaluate_ratios(self.h, int(iat), nvp, _p(pos), _p(out)))
        return self._out(out)

    def mw_invertTranspose(self, a):
        """DiracMatrixInverter::mw_invert_transpose: a [nmat][n][n] (real or complex) -> ((a^-1)^T,
        log det per matrix as log|det| + i phase)."""
        a = np.asarray(a)
        cplx = np.iscomplexobj(a)
        nmat, n = a.shape[0], a.shape[1]
        src = np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1) if cplx else a, dtype=np.float64)
        out = np.empty_like(src)
        lv = np.empty((nmat, 2))
        self._check(self.L.qmi_mw_invert_transpose(self.h, int(cplx), n, nmat, _p(src), _p(out), _p(lv)))
        inv = out[..., 0] + 1j * out[..., 1] if cplx else out
        return inv, lv[:, 0] + 1j * lv[:, 1]

    def mw_invertTranspose_noalloc(self, cplx, n, nmat, a_dev, inv_dev, log_dev, info_dev):
        """qmi_mw_invert_transpose_noalloc on device addresses (integers): enqueued on this context's
        stream, no allocation, no synchronisation; info_dev receives the zero-pivot column per matrix."""
        vp = C.c_void_p
        self._check(self.L.qmi_mw_invert_transpose_noalloc(
            self.h, int(bool(cplx)), int(n), int(nmat), vp(a_dev), vp(inv_dev), vp(log_dev) if log_dev else None,
            vp(info_dev)))

    def mw_getInverse(self, g):
        """DelayedUpdateBatched::mw_transferAinv_D2H: inverse Slater matrix of spin group g of every
        walker, [nw][n][n] (row i = inverse row of electron i of the group)."""
        n = self.n_up if g == 0 else self.N - self.n_up
        out = self._vals(self.nw, n, n)
        self._check(self.L.qmi_mw_get_inverse(self.h, int(g), _p(out)))
        return self._out(out)

    def mw_structure_factor(self, kvecs):
        """StructFact rho_k of the electrons per spin group: [nw][2][nk] complex at kvecs [nk][3]."""
        kv = np.ascontiguousarray(kvecs, dtype=np.float64).reshape(-1, 3)
        out = np.empty((self.nw, 2, kv.shape[0], 2))
        self._check(self.L.qmi_mw_structure_factor(self.h, C.c_int(kv.shape[0]), _p(kv), _p(out)))
        return out[..., 0] + 1j * out[..., 1]

    def system_set_lattice(self, R):
        """Cell vectors (rows of R) of the orbitals' lattice; a skewed cell serves the orbital paths
        only (no Jastrow factors, Coulomb, pseudopotential or distance tables)."""
        Rv = np.ascontiguousarray(R, dtype=np.float64).reshape(3, 3)
        self._check(self.L.qmi_system_set_lattice(self.h, _p(Rv)))

    def coulomb_setup(self, ion_charge, dim_cutoff=15.0, handler="ewald"):
        """Periodic Coulomb tables; handler is the lattice input's LR_handler: "ewald"
        (EwaldHandler3D) or "opt_breakup" (the optimized LPQHI breakup, the reference's default)."""
        code = LR_HANDLERS[handler]
        self._check(self.L.qmi_coulomb_setup_handler(self.h, C.c_double(ion_charge), C.c_double(dim_cutoff),
                                                     C.c_int(code)))

    # ---------------------------------------------------------------- pseudopotential
    def nlpp_setup(self, ls, dr, v, v2, rcut, knots, weights):
        """Channels with angular momenta ls; v, v2 [nch][ngrid] spline values / second derivatives on
        r_i = i dr; quadrature knots [nknot][3] and weights [nknot]."""
        ls = np.ascontiguousarray(ls, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        v2 = np.ascontiguousarray(v2, dtype=np.float64)
        knots = np.ascontiguousarray(knots, dtype=np.float64)
        weights = np.ascontiguousarray(weights, dtype=np.float64)
        self._check(self.L.qmi_nlpp_setup(self.h, len(ls), ls.ctypes.data_as(C.POINTER(C.c_int)), v.shape[1],
                                          C.c_double(dr), _p(v), _p(v2), C.c_double(rcut), len(weights),
                                          _p(knots), _p(weights)))

    def mw_evaluateNLPP(self, rot):
        """Non-local energy [nw]; rot [nw][3][3] rotations of the quadrature grid."""
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(self.nw, 9)
        v = np.empty(self.nw)
        self._check(self.L.qmi_mw_evaluate_nlpp(self.h, _p(rot), _p(v)))
        return v

    # ---------------------------------------------------------------- DMC
    def dmc_init(self, tau, capacity, seed, feedback=1.0, sigma_bound=10.0, warmup_steps=200,
                 refenergy_history=0, cutoff_scheme=0, target_walkers=0, recompute_period=0, nonlocal_moves=0):
        """DMC on the loaded walkers (one walker per crowd; call coulomb_setup first; nonlocal_moves=1:
        T-moves v0, after nlpp_setup)."""
        p = DmcParams(tau, feedback, sigma_bound, warmup_steps, refenergy_history, cutoff_scheme, capacity,
                      target_walkers, recompute_period, seed, nonlocal_moves)
        self._check(self.L.qmi_dmc_init(self.h, C.byref(p)))

    def dmc_run(self, steps):
        """Per step [steps][11]: energy, variance, weight, walkers before / after branching, E_trial,
        E_ref, tau_eff, accepted, rejected, accepted non-local moves."""
        out = np.empty((steps, 11))
        self._check(self.L.qmi_dmc_run(self.h, int(steps), _p(out)))
        self.nw = int(self.L.qmi_num_walkers(self.h))
        return out

    def dmc_walker_data(self):
        w, e = np.empty(self.nw), np.empty(self.nw)
        self._check(self.L.qmi_dmc_walker_data(self.h, _p(w), _p(e)))
        return w, e

    # ---------------------------------------------------------------- restart files
    def checkpoint_save(self) -> bytes:
        """Walkers, random streams and (DMC) branch state; read_checkpoint() decodes it."""
        n = C.c_size_t(0)
        self._check(self.L.qmi_checkpoint_save(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._check(self.L.qmi_checkpoint_save(self.h, buf, n, C.byref(n)))
        return buf.raw[:n.value]

    def checkpoint_load(self, data: bytes):
        """Restart from checkpoint_save's bytes (system, and for DMC coulomb_setup / nlpp_setup, first)."""
        self._check(self.L.qmi_checkpoint_load(self.h, bytes(data), len(data)))
        self.nw = int(self.L.qmi_num_walkers(self.h))

    def coulomb_nk(self):
        """Number of k vectors of the Ewald long-range sum (after coulomb_setup)."""
        return int(self.L.qmi_coulomb_nk(self.h))

    def mw_evaluatePotential(self):
        """[nw][3]: electron-electron, electron-ion, ion-ion Coulomb energy (the handler of coulomb_setup)."""
        v = np.empty((self.nw, 3))
        self._check(self.L.qmi_mw_evaluate_potential(self.h, _p(v)))
        return v

    def mw_distanceRows(self, pos, n_ion):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nw, 3)
        ee_r, ee_dr = np.empty((self.nw, self.N)), np.empty((self.nw, self.N, 3))
        ei_r, ei_dr = np.empty((self.nw, max(n_ion, 1))), np.empty((self.nw, max(n_ion, 1), 3))
        self._check(self.L.qmi_mw_distance_rows(self.h, _p(pos), _p(ee_r), _p(ee_dr), _p(ei_r), _p(ei_dr)))
        return ee_r, ee_dr, ei_r[:, :n_ion], ei_dr[:, :n_ion]

    # ---------------------------------------------------------------- driver
    def vmc_run(self, steps, sub_steps=1, tau=0.3, recompute_period=0, seed=1, trace=False, energies=True):
        E = np.empty((steps, self.nw)) if energies else None
        tr = np.empty((steps, sub_steps, self.N, self.nw, 8)) if trace else None
        self._check(self.L.qmi_vmc_run(self.h, int(steps), int(sub_steps), C.c_double(tau), int(recompute_period),
                                       C.c_uint32(seed), _p(E), _p(tr)))
        return E, tr

    # ---------------------------------------------------------------- learned proposal
    MLP_NPARAM = 42372

    def mlp_params(self, params=None):
        """Upload params (FP32, MLP_NPARAM) when given; return the current parameters."""
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            assert p.size == self.MLP_NPARAM
            self._check(self.L.qmi_mlp_params(self.h, 1, p.ctypes.data_as(C.POINTER(C.c_float))))
        out = np.empty(self.MLP_NPARAM, dtype=np.float32)
        self._check(self.L.qmi_mlp_params(self.h, 0, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty((X.shape[0], 4), dtype=np.float32)
        self._check(self.L.qmi_mlp_forward(self.h, X.shape[0], X.ctypes.data_as(C.POINTER(C.c_float)),
                                           out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def mlp_gradient(self, X, dout):
        X = np.ascontiguousarray(X, dtype=np.float32)
        dout = np.ascontiguousarray(dout, dtype=np.float32)
        g = np.empty(self.MLP_NPARAM, dtype=np.float32)
        self._check(self.L.qmi_mlp_gradient(self.h, X.shape[0], X.ctypes.data_as(C.POINTER(C.c_float)),
                                            dout.ctypes.data_as(C.POINTER(C.c_float)),
                                            g.ctypes.data_as(C.POINTER(C.c_float))))
        return g

    def proposal_config(self, learned=True, train=False, lr=1e-3, init_seed=0):
        self._check(self.L.qmi_proposal_config(self.h, int(learned), int(train), C.c_double(lr),
                                               C.c_uint32(init_seed)))

    def proposal_stats(self):
        out = np.empty(4)
        self._check(self.L.qmi_proposal_stats(self.h, _p(out)))
        return dict(objective=out[0], acceptance=out[1], moves=out[2], train_steps=out[3])

    def proposal_features(self, iat, pos, drift, tau):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nw, 3)
        drift = np.ascontiguousarray(drift, dtype=np.float64).reshape(self.nw, 3)
        out = np.empty((self.nw, 68), dtype=np.float32)
        self._check(self.L.qmi_mw_proposal_features(self.h, int(iat), _p(pos), _p(drift), C.c_double(tau),
                                                    out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def vmc_run_learned(self, steps, sub_steps=1, tau=0.3, recompute_period=0, seed=1, trace=False):
        """vmc_run that also returns the learned-proposal trace [steps][sub][N][nw][15]: mu[3], s,
        mu'[3], s', log q_fwd, log q_rev, acceptance probability, objective, drift at r'[3]."""
        E = np.empty((steps, self.nw))
        tr = np.empty((steps, sub_steps, self.N, self.nw, 8)) if trace else None
        lt = np.empty((steps, sub_steps, self.N, self.nw, 15)) if trace else None
        self._check(self.L.qmi_vmc_run_ex(self.h, int(steps), int(sub_steps), C.c_double(tau), int(recompute_period),
                                          C.c_uint32(seed), _p(E), _p(tr), _p(lt)))
        return E, tr, lt

    def vmc_stats(self):
        s = np.empty(2)
        self._check(self.L.qmi_vmc_stats(self.h, _p(s)))
        return dict(accepted=s[0], rejected=s[1])

    # ---------------------------------------------------------------- tuning and graph replay
    def tune_set(self, delay_rank=0, move_kernel=0, flush_variant=0):
        """delay_rank (0: unchanged), move_kernel (0: launch per move, 1: per delay window),
        flush_variant (0: matrix-core flush for n <= 256, 1: three LDS passes)."""
        self._check(self.L.qmi_tune_set(self.h, int(delay_rank), int(move_kernel), int(flush_variant)))

    def tune_get(self):
        out = np.empty(3, dtype=np.int32)
        self._check(self.L.qmi_tune_get(self.h, out.ctypes.data_as(_ip)))
        return dict(delay_rank=int(out[0]), move_kernel=int(out[1]), flush_variant=int(out[2]))

    def tune_partitions(self, partitions=0):
        """Walker partitions of the fused sweep (partitions > 0 sets them); returns the current count."""
        cur = C.c_int()
        self._check(self.L.qmi_tune_partitions(self.h, int(partitions), C.byref(cur)))
        return cur.value

    def graph_enable(self, on=True):
        self._check(self.L.qmi_graph_enable(self.h, int(bool(on))))

    # ---------------------------------------------------------------- timing
    def timing_enable(self, on=True):
        self._check(self.L.qmi_timing_enable(self.h, int(bool(on))))

    def timing_reset(self):
        self._check(self.L.qmi_timing_reset(self.h))

    def timing(self):
        buf = C.create_string_buffer(4096)
        ms, calls = np.empty(64), np.empty(64)
        nk = C.c_int()
        self._check(self.L.qmi_timing_get(self.h, buf, 4096, _p(ms), _p(calls), 64, C.byref(nk)))
        names = buf.value.decode().split("\n")[: nk.value]
        return {n: dict(ms=float(ms[i]), calls=int(calls[i])) for i, n in enumerate(names)}


class DistanceTable:
    """qmi_dtable: DistanceTable (src/Particle/DistanceTable.h) of one context's walkers.  AA rows
    hold electrons j < i (SoaDistanceTableAA), AB rows every ion (SoaDistanceTableAB);
    displacements are source - target."""

    def __init__(self, ctx: Context, kind: str):
        self.ctx = ctx
        self.kind = {"AA": 0, "AB": 1}[kind]
        self.h = C.c_void_p()
        ctx._check(ctx.L.qmi_dt_create(ctx.h, self.kind, C.byref(self.h)))
        self.nw, self.N = ctx.nw, ctx.N
        self.ns = ctx.N if self.kind == 0 else None

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.ctx.L.qmi_dt_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mw_evaluate(self):
        self.ctx._check(self.ctx.L.qmi_dt_mw_evaluate(self.h))

    def mw_move(self, iat, rnew, prepare_old=True):
        rnew = np.ascontiguousarray(rnew, dtype=np.float64).reshape(self.nw, 3)
        self.ctx._check(self.ctx.L.qmi_dt_mw_move(self.h, int(iat), _p(rnew), int(bool(prepare_old))))

    def _mask(self, accepted):
        return np.ascontiguousarray(np.asarray(accepted, dtype=np.int32).reshape(self.nw))

    def update(self, iat, accepted):
        """DistanceTable::update for the walkers whose move is accepted."""
        acc = self._mask(accepted)
        self.ctx._check(self.ctx.L.qmi_dt_mw_update(self.h, int(iat), acc.ctypes.data_as(_ip), 0))

    def mw_updatePartial(self, iat, from_temp):
        acc = self._mask(from_temp)
        self.ctx._check(self.ctx.L.qmi_dt_mw_update(self.h, int(iat), acc.ctypes.data_as(_ip), 1))

    def get(self, ns):
        """(r [nw][N][ns], dr [nw][N][ns][3], temp_r [nw][ns], temp_dr [nw][ns][3], old_r, old_dr)
        with displacements moved to the last axis (AB: old rows are None)."""
        nw, N = self.nw, self.N
        r, dr = np.empty((nw, N, ns)), np.empty((nw, N, 3, ns))
        tr, tdr = np.empty((nw, ns)), np.empty((nw, 3, ns))
        aa = self.kind == 0
        orr, odr = (np.empty((nw, ns)), np.empty((nw, 3, ns))) if aa else (None, None)
        self.ctx._check(self.ctx.L.qmi_dt_get(self.h, _p(r), _p(dr), _p(tr), _p(tdr), _p(orr), _p(odr)))
        mv = lambda a: None if a is None else np.moveaxis(a, -2, -1)
        return r, mv(dr), tr, mv(tdr), orr, mv(odr)


class Batch:
    """Several contexts advanced together in one captured sub-step graph (uneven batching)."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        self.L = load_library()
        self.h = C.c_void_p()
        arr = (C.c_void_p * len(self.ctxs))(*[c.h.value for c in self.ctxs])
        st = self.L.qmi_batch_create(C.byref(self.h), arr, len(self.ctxs))
        if st != 0:
            raise QmiError(st, "qmi_batch_create failed")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.qmi_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise QmiError(st, self.L.qmi_batch_last_error(self.h).decode())

    def vmc_run(self, steps, sub_steps=1, tau=0.3, recompute_period=0, seeds=None):
        """Returns the kinetic local energies [steps][nw] of each context."""
        seeds = np.ascontiguousarray(seeds if seeds is not None else np.arange(1, len(self.ctxs) + 1),
                                     dtype=np.uint32)
        E = [np.empty((steps, c.nw)) for c in self.ctxs]
        ptrs = (_dp * len(E))(*[_p(e) for e in E])
        self._check(self.L.qmi_batch_vmc_run(self.h, int(steps), int(sub_steps), C.c_double(tau),
                                             int(recompute_period), seeds.ctypes.data_as(C.POINTER(C.c_uint32)),
                                             ptrs))
        return E

    def autotune(self, tau=0.3, trials=3):
        """Returns (choice [nctx][3] = delay rank, move_kernel, flush_variant; ms [nctx + 2])."""
        n = len(self.ctxs)
        choice = np.empty((n, 3), dtype=np.int32)
        ms = np.empty(n + 2)
        self._check(self.L.qmi_batch_autotune(self.h, C.c_double(tau), int(trials), choice.ctypes.data_as(_ip),
                                              _p(ms)))
        return choice, ms

    def set_concurrent(self, on=True):
        self._check(self.L.qmi_batch_set_concurrent(self.h, int(bool(on))))

    def concurrent(self):
        v = C.c_int()
        self._check(self.L.qmi_batch_get_concurrent(self.h, C.byref(v)))
        return bool(v.value)
