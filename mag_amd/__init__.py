"""mag_amd — MI355X-native constrained HMC for SU(3) lattice gauge fields.

Python mirror of the reference's field/HMC entry points (hummingtree/mag:
field-hmc.h, field-matrix.h, field-multi.h, field-wloop.h, field-demon.h),
calling the gfx950 kernels of libmag_amd.so through its C ABI. Device fields
are handles; host arrays use the reference's qlat layout, float64 of shape
(V, 4, 18) = [site][mu][3x3 complex, interleaved re/im].
"""
import ctypes as C
import weakref
from dataclasses import dataclass, field as dfield

import numpy as np

from . import _lib
from ._lib import (ALG8, F32, F64, IWASAKI, LINK12, LINK18, REAL1, WILSON, Action, HmcParams, MagError,
                   TrajResult)

__all__ = [
    "Context", "Field", "Gauge", "Arg_chmc", "Chain", "MagError",
    "get_force", "evolve_gauge_field", "evolve_momentum", "force_gradient_integrator", "leap_frog_integrator",
    "get_hamiltonian", "total_plaq", "total_rectangular", "avg_plaquette", "gauge_action", "kinetic_energy",
    "reunitarize", "init_momentum", "expiQ", "compute_Lambda", "DexpiQ", "chair_staple_dagger",
    "twist_staple_dagger", "stout_smear", "avg_wilson_loop",
    "demon_microcanonical", "run_chmc", "Arg_export", "export_config_nersc", "import_config_nersc",
    "nersc_write", "nersc_read", "total_plaq_async", "total_rectangular_async", "get_hamiltonian_async",
    "avg_wilson_loop_async", "get_hamiltonian_noalloc", "coarsen_links", "update_constrain", "check_constrained_plaquette",
    "lattice_expansion", "GchbArg", "AlgGCtrnHeatBath", "local_metropolis", "ORDER_REFERENCE", "ORDER_CHECKERBOARD",
    "demon_checkerboard",
]

_D = C.POINTER(C.c_double)


def _check(code, ctx=None):
    if code != 0:
        msg = ""
        if ctx is not None:
            m = _lib.lib().mag_last_error(ctx)
            msg = m.decode() if m else ""
        raise MagError(code, msg)


def _dptr(a):
    return a.ctypes.data_as(_D)


class Context:
    """A device context (mag_ctx): one HIP stream plus the workspace it owns."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        self._children = weakref.WeakSet()  # fields, chains, servers created on this context
        _check(_lib.lib().mag_ctx_create(device, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def adopt(self, obj):
        self._children.add(obj)

    def close(self):
        """Destroy the context after every live object created on it (the C ABI
        requires fields, chains and servers to go first)."""
        if self._h:
            for ch in list(self._children):
                ch._release()
            _lib.lib().mag_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(_lib.lib().mag_ctx_synchronize(self._h), self._h)

    def memory(self):
        b = C.c_size_t()
        _check(_lib.lib().mag_ctx_memory(self._h, C.byref(b)), self._h)
        return b.value

    def set_timing(self, on):
        _check(_lib.lib().mag_ctx_set_timing(self._h, 1 if on else 0), self._h)

    def set_timing_family(self, family=None):
        """Time only this kernel family while timing is on (None: every family)."""
        _check(_lib.lib().mag_ctx_set_timing_family(self._h, family.encode() if family else None), self._h)

    def reset_timing(self):
        _check(_lib.lib().mag_ctx_reset_timing(self._h), self._h)

    def kernel_time(self, family):
        ms, n = C.c_double(), C.c_long()
        _check(_lib.lib().mag_ctx_kernel_time(self._h, family.encode(), C.byref(ms), C.byref(n)), self._h)
        return ms.value, n.value


class Field:
    """Device field (qlat::Field<cps::Matrix> of multiplicity 4)."""

    def __init__(self, ctx, dims, kind=LINK18, prec=F64, _handle=None):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.kind, self.prec = kind, prec
        if _handle is not None:
            self._h, self._owned = _handle, False
        else:
            self._h, self._owned = C.c_void_p(), True
            _check(_lib.lib().mag_field_create(ctx.handle, (C.c_int * 4)(*self.dims), kind, prec, C.byref(self._h)),
                   ctx.handle)
            ctx.adopt(self)

    @property
    def handle(self):
        return self._h

    @property
    def volume(self):
        return int(np.prod(self.dims))

    def _release(self):
        if self._owned and self._h:
            _lib.lib().mag_field_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def _host_shape(self):
        # REAL1: one real per link ([site][mu]); otherwise 3x3 complex per link
        return (self.volume, 4) if self.kind == REAL1 else (self.volume, 4, 18)

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64).reshape(self._host_shape())
        _check(_lib.lib().mag_field_upload_aos(self._h, _dptr(a)), self.ctx.handle)
        return self

    def download(self):
        a = np.empty(self._host_shape())
        _check(_lib.lib().mag_field_download_aos(self._h, _dptr(a)), self.ctx.handle)
        return a

    def copy_from(self, other):
        _check(_lib.lib().mag_field_copy(self._h, other.handle), self.ctx.handle)
        return self

    def unit(self):
        _check(_lib.lib().mag_unit_links(self._h), self.ctx.handle)
        return self

    def haar(self, seed):
        _check(_lib.lib().mag_haar_links(self._h, seed), self.ctx.handle)
        return self


# ----------------------------------------------------------------- arguments
@dataclass
class Gauge:
    """Gauge{type, c1} (field-hmc.h:344, 384): WILSON or IWASAKI form (c0 = 1 - 8 c1)."""
    type: int = WILSON
    c1: float = 0.0


@dataclass
class Arg_chmc:
    """Arg_chmc (field-hmc.h:294-307) plus smearing and run controls."""
    mag: int = 0
    trajectory_length: int = 8
    num_trajectory: int = 1
    num_step_between_output: int = 1
    num_forced_accept_step: int = 0
    num_step_before_output: int = 0
    beta: float = 6.0
    dt: float = 0.125
    gauge: Gauge = dfield(default_factory=Gauge)
    n_stout: int = 0
    rho: float = 0.1

    def action(self):
        return Action(self.gauge.type, self.beta, self.gauge.c1, self.mag, self.n_stout, self.rho)


def _act(arg):
    return arg if isinstance(arg, Action) else arg.action()


# ----------------------------------------------------------------- functions
def get_force(F, U, arg):
    """get_force (field-hmc.h:339-425); smeared force when arg.n_stout > 0."""
    a = _act(arg)
    _check(_lib.lib().mag_gauge_force(F.handle, U.handle, C.byref(a)), U.ctx.handle)


def evolve_gauge_field(U, P, dt, arg):
    """evolve_gauge_field (field-hmc.h:442-482)."""
    _check(_lib.lib().mag_evolve_gauge(U.handle, P.handle, dt, _act(arg).mag), U.ctx.handle)


def evolve_momentum(P, F, dt, arg=None, m=1.0):
    """evolve_momentum (field-hmc.h:427-440)."""
    _check(_lib.lib().mag_evolve_momentum(P.handle, F.handle, dt, m), P.ctx.handle)


def force_gradient_integrator(U, P, arg, reunit=True):
    """force_gradient_integrator (field-hmc.h:525-567); returns the reunitarize deviation."""
    a = _act(arg)
    d = C.c_double()
    _check(_lib.lib().mag_fg_integrate(U.handle, P.handle, C.byref(a), arg.trajectory_length, arg.dt,
                                       1 if reunit else 0, C.byref(d)), U.ctx.handle)
    return d.value


def leap_frog_integrator(U, P, arg):
    """leap_frog_integrator (field-hmc.h:569-587)."""
    a = _act(arg)
    _check(_lib.lib().mag_leapfrog_integrate(U.handle, P.handle, C.byref(a), arg.trajectory_length, arg.dt),
           U.ctx.handle)


def get_hamiltonian(U, P, arg):
    """get_hamiltonian (field-hmc.h:589-644): returns (H, [kinetic, potential])."""
    a = _act(arg)
    parts = (C.c_double * 2)()
    h = C.c_double()
    _check(_lib.lib().mag_hamiltonian(U.handle, P.handle, C.byref(a), parts, C.byref(h)), U.ctx.handle)
    return h.value, [parts[0], parts[1]]


def _scalar(fn, *args, ctx):
    v = C.c_double()
    _check(fn(*args, C.byref(v)), ctx)
    return v.value


def total_plaq(U):
    """total_plaq (field-matrix.h:299-330)."""
    return _scalar(_lib.lib().mag_total_plaq, U.handle, ctx=U.ctx.handle)


def total_rectangular(U):
    """total_rectangular (field-matrix.h:354-385)."""
    return _scalar(_lib.lib().mag_total_rectangular, U.handle, ctx=U.ctx.handle)


def avg_plaquette(U):
    """avg_plaquette (field-matrix.h:387-389)."""
    return _scalar(_lib.lib().mag_avg_plaquette, U.handle, ctx=U.ctx.handle)


def gauge_action(U, arg):
    a = _act(arg)
    return _scalar(_lib.lib().mag_gauge_action, U.handle, C.byref(a), ctx=U.ctx.handle)


def kinetic_energy(P, mag=0):
    return _scalar(_lib.lib().mag_kinetic_energy, P.handle, mag, ctx=P.ctx.handle)


def reunitarize(U):
    """reunitarize (field-matrix.h:83-100): returns the max deviation."""
    return _scalar(_lib.lib().mag_reunitarize, U.handle, ctx=U.ctx.handle)


def init_momentum(P, seed, traj=0, m=1.0):
    """init_momentum (field-hmc.h:646-693) from the Philox momentum stream."""
    _check(_lib.lib().mag_init_momentum(P.handle, seed, traj, m), P.ctx.handle)


def stout_smear(out, U, rho, nlev):
    _check(_lib.lib().mag_stout_smear(out.handle, U.handle, rho, nlev), U.ctx.handle)


def avg_wilson_loop(U, r, T):
    """avg_wilson_loop (field-wloop.h:58-98)."""
    return _scalar(_lib.lib().mag_avg_wilson_loop, U.handle, (C.c_int * 3)(*r), T, ctx=U.ctx.handle)


# Asynchronous reductions: d_out is a device address (an integer, e.g. a
# torch tensor's data_ptr()); the result is written in the context's stream
# order and nothing synchronises.
def total_plaq_async(U, d_out):
    _check(_lib.lib().mag_total_plaq_async(U.handle, C.c_void_p(d_out)), U.ctx.handle)


def total_rectangular_async(U, d_out):
    _check(_lib.lib().mag_total_rectangular_async(U.handle, C.c_void_p(d_out)), U.ctx.handle)


def get_hamiltonian_async(U, P, arg, d_out):
    """d_out[0:3] = (kinetic, potential, H)."""
    a = _act(arg)
    _check(_lib.lib().mag_hamiltonian_async(U.handle, P.handle, C.byref(a), C.c_void_p(d_out)), U.ctx.handle)


def get_hamiltonian_noalloc(U, P, arg, d_out):
    """get_hamiltonian_async of a thin action without any allocation (partial
    sums in the library's fixed per-GPU array): d_out[0:3] = (kinetic,
    potential, H). Calls overlapping in time on one GPU race."""
    a = _act(arg)
    _check(_lib.lib().mag_hamiltonian_noalloc(U.handle, P.handle, C.byref(a), C.c_void_p(d_out)), U.ctx.handle)


def avg_wilson_loop_async(U, r, T, d_out):
    _check(_lib.lib().mag_avg_wilson_loop_async(U.handle, (C.c_int * 3)(*r), T, C.c_void_p(d_out)), U.ctx.handle)


def _batch(ctx, fn, mats, prec):
    arrs = [np.ascontiguousarray(np.asarray(m, dtype=np.complex128)).view(np.float64).reshape(-1, 18) for m in mats]
    n = arrs[0].shape[0]
    if any(a.shape[0] != n for a in arrs):
        raise ValueError("batched matrices must have the same count")
    out = np.empty((n, 18))
    _check(fn(ctx.handle, *[_dptr(a) for a in arrs], _dptr(out), n, prec), ctx.handle)
    return out.view(np.complex128).reshape(np.asarray(mats[0]).shape)


def expiQ(ctx, Q, prec=F64):
    """expiQ (field-hmc.h:82-123) on the device for a batch (..., 3, 3)."""
    return _batch(ctx, _lib.lib().mag_expiQ_batch, [Q], prec)


def compute_Lambda(ctx, Q, SigmaP, U, prec=F64):
    """compute_Lambda (field-multi.h:180-269) on the device for a batch."""
    return _batch(ctx, _lib.lib().mag_compute_Lambda_batch, [Q, SigmaP, U], prec)


def DexpiQ(ctx, Q, DQ, prec=F64):
    """DexpiQ (field-multi.h:90-178) on the device for a batch: d exp(iQ)/dt along DQ."""
    return _batch(ctx, _lib.lib().mag_DexpiQ_batch, [Q, DQ], prec)


def chair_staple_dagger(S, U):
    """chair_staple_dagger (field-staple.h:141-151) for every link into the FP64
    18-real field S (the undaggered chair sum, as the reference returns it)."""
    _check(_lib.lib().mag_chair_staple_dagger(S.handle, U.handle), U.ctx.handle)
    return S


def twist_staple_dagger(S, U):
    """twist_staple_dagger (field-staple.h:156-165) for every link into S (undaggered)."""
    _check(_lib.lib().mag_twist_staple_dagger(S.handle, U.handle), U.ctx.handle)
    return S


def demon_microcanonical(U, nsweeps, E_max, seed, E=None):
    """demon_microcanonical (field-demon.h:25-126): returns (E[4], (accept, reject))."""
    e = np.zeros(4) if E is None else np.array(E, dtype=np.float64)
    ar = (C.c_long * 2)()
    _check(_lib.lib().mag_demon_microcanonical(U.handle, nsweeps, E_max, seed, _dptr(e), ar), U.ctx.handle)
    return e, (ar[0], ar[1])


def demon_checkerboard(U, nsweeps, E_max, seed, demons=None):
    """Checkerboard demon (demon_microcanonical with a demon per link, 128
    commuting classes per direction). demons: four REAL1 fields (plaquette,
    rectangle, chair, twist energies per link), created at zero when None.
    Returns (demons, (accepted, rejected))."""
    if demons is None:
        vol = int(np.prod(U.dims))
        demons = [Field(U.ctx, U.dims, REAL1, F64).upload(np.zeros((vol, 4))) for _ in range(4)]
    ar = (C.c_long * 2)()
    _check(_lib.lib().mag_demon_checkerboard(U.handle, *[d.handle for d in demons], nsweeps, E_max, seed, ar),
           U.ctx.handle)
    return demons, (ar[0], ar[1])


# ----------------------------------------------------------------- local Metropolis
ORDER_REFERENCE, ORDER_CHECKERBOARD = 0, 1


def local_metropolis(U, beta, mag, num_iter, nhits, small, seed, order=ORDER_CHECKERBOARD):
    """algGCtrnHeatBath::run (alg_gchb.h:290-387) on FP64 links: returns
    (acc_rej[4] = accept, reject, acceptCons, rejectCons; average plaquette
    per iteration; constrained plaquette per iteration)."""
    acc = (C.c_long * 4)()
    p, cp = np.zeros(max(num_iter, 1)), np.zeros(max(num_iter, 1))
    _check(_lib.lib().mag_local_metropolis(U.handle, beta, mag, num_iter, nhits, small, seed, order, acc,
                                           _dptr(p), _dptr(cp)), U.ctx.handle)
    return list(acc), p[:num_iter], cp[:num_iter]


@dataclass
class GchbArg:
    """gchbArg (alg_gchb.h:98-105)."""
    mag: int = 0
    numIter: int = 1
    nHits: int = 1
    small: float = 0.5


class AlgGCtrnHeatBath:
    """algGCtrnHeatBath (alg_gchb.h:107-387) over a device link field; the
    Wilson beta is passed explicitly (GJP.Beta() in CPS)."""

    def __init__(self, U, beta, arg, seed=1234, order=ORDER_CHECKERBOARD):
        self.U, self.beta, self.arg, self.seed, self.order = U, beta, arg, seed, order
        self.accept = self.reject = self.acceptCons = self.rejectCons = 0

    def run(self):
        """numIter iterations; returns the rows the reference prints:
        (iteration, avgPlaq, avgConsPlaq)."""
        a = self.arg
        acc, p, cp = local_metropolis(self.U, self.beta, a.mag, a.numIter, a.nHits, a.small, self.seed, self.order)
        self.accept += acc[0]
        self.reject += acc[1]
        self.acceptCons += acc[2]
        self.rejectCons += acc[3]
        return [(i, p[i], cp[i]) for i in range(a.numIter)]

    def accept_rate(self):
        """accpetRate (alg_gchb.h:264-276): (normal, constrained, total)."""
        n, c = self.accept + self.reject, self.acceptCons + self.rejectCons
        return (self.accept / n if n else 0.0, self.acceptCons / c if c else 0.0,
                (self.accept + self.acceptCons) / (n + c) if n + c else 0.0)


# ----------------------------------------------------------------- chain
class Chain:
    """Device-resident HMC chain (run_chmc / update_field, field-hmc.h:770-986)."""

    def __init__(self, ctx, dims, arg, seed=1234, precision=F64, integrator=0, accept_mode=0, demon_emax=0.0,
                 reunit=True, recon12=None):
        """recon12: FP64 links stored as 12 reals (FP32 links always are)."""
        self.ctx = ctx
        self.dims = tuple(dims)
        self.arg = arg
        self.act = _act(arg)
        r12 = precision == F32 or bool(recon12)
        self.params = HmcParams(arg.trajectory_length, arg.dt, integrator, arg.num_forced_accept_step, accept_mode,
                                demon_emax, seed, 1 if reunit else 0, precision, 1 if r12 else 0)
        self._h = C.c_void_p()
        _check(_lib.lib().mag_chain_create(ctx.handle, (C.c_int * 4)(*self.dims), C.byref(self.act),
                                           C.byref(self.params), C.byref(self._h)), ctx.handle)
        ctx.adopt(self)

    def _release(self):
        if self._h:
            _lib.lib().mag_chain_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def set_links(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        _check(_lib.lib().mag_chain_set_links_aos(self._h, _dptr(a)), self.ctx.handle)

    def get_links(self):
        a = np.empty((int(np.prod(self.dims)), 4, 18))
        _check(_lib.lib().mag_chain_get_links_aos(self._h, _dptr(a)), self.ctx.handle)
        return a

    def haar(self, seed):
        _check(_lib.lib().mag_chain_haar_links(self._h, seed), self.ctx.handle)

    def links(self):
        h = C.c_void_p()
        _check(_lib.lib().mag_chain_links(self._h, C.byref(h)), self.ctx.handle)
        kind = LINK12 if self.params.recon12 else LINK18
        return Field(self.ctx, self.dims, kind, self.params.precision, _handle=h)

    def reset_demon(self):
        """Re-centre the demon window on the links the chain holds now (after
        writing them through links())."""
        _check(_lib.lib().mag_chain_reset_demon(self._h), self.ctx.handle)

    def run(self, ntraj):
        res = (TrajResult * max(ntraj, 1))()
        _check(_lib.lib().mag_chain_run(self._h, ntraj, res), self.ctx.handle)
        return [res[i] for i in range(ntraj)]

    def reversibility(self):
        d, dh = C.c_double(), C.c_double()
        _check(_lib.lib().mag_chain_reversibility(self._h, C.byref(d), C.byref(dh)), self.ctx.handle)
        return d.value, dh.value

    def demon_state(self):
        s = np.zeros(4)
        _check(_lib.lib().mag_chain_demon_state(self._h, _dptr(s)), self.ctx.handle)
        return s


# --------------------------------------------------- mag-blocked lattices
def _mag(arg_or_mag):
    return arg_or_mag if isinstance(arg_or_mag, int) else arg_or_mag.mag


def coarsen_links(fine, mag, coarse):
    """Coarse links = ordered products of mag fine links (double_multigrid's
    coarse initialisation, field-hmc.h:1068-1078)."""
    _check(_lib.lib().mag_coarsen_links(fine.handle, mag, coarse.handle), fine.ctx.handle)
    return coarse


def update_constrain(gField, gField_coarse, arg):
    """update_constrain (field-hmc.h:988-1009); arg is an Arg_chmc (its mag) or mag."""
    _check(_lib.lib().mag_update_constrain(gField.handle, gField_coarse.handle, _mag(arg)), gField.ctx.handle)


def check_constrained_plaquette(gField, mag):
    """check_constrained_plaquette (field-matrix.h:407-443)."""
    return _scalar(_lib.lib().mag_constrained_plaquette, gField.handle, mag, ctx=gField.ctx.handle)


def lattice_expansion(mag, original, expanded, original_config=None, expanded_config=None):
    """lattice_expansion (expansion/main.C:19-62): expanded = unit links with
    the original links at the fine sites mag*x. With file names, the original
    is imported from and the expansion exported to NERSC files (two-row
    format, label "lattice expansion"), as the reference does. Returns
    (average plaquette, constrained plaquette) of the expansion."""
    if original_config is not None:
        import_config_nersc(original, original_config)
    _check(_lib.lib().mag_lattice_expansion(original.handle, mag, expanded.handle), original.ctx.handle)
    res = avg_plaquette(expanded), check_constrained_plaquette(expanded, mag)
    if expanded_config is not None:
        export_config_nersc(expanded, expanded_config, Arg_export(0.0, 0, "lattice expansion"), does_skip_third=True)
    return res


# ------------------------------------------------------------ NERSC I/O
@dataclass
class Arg_export:
    """Arg_export (field-matrix.h:445-450)."""
    beta: float = 0.0
    sequence_num: int = 0
    ensemble_label: str = ""


def _header_dict(h):
    return {"dims": tuple(h.dims), "skip_third": bool(h.skip_third), "link_trace": h.link_trace,
            "plaquette": h.plaquette, "beta": h.beta, "sequence_number": h.sequence_number,
            "ensemble_label": h.ensemble_label.decode(), "checksum": h.checksum, "crc32": h.crc32,
            "checksum_swapped": bool(h.checksum_swapped)}


def export_config_nersc(U, path, arg=None, does_skip_third=False):
    """export_config_nersc (field-matrix.h:452-568)."""
    arg = arg or Arg_export()
    _check(_lib.lib().mag_export_config_nersc(U.handle, str(path).encode(), arg.beta, arg.sequence_num,
                                              arg.ensemble_label.encode(), 1 if does_skip_third else 0),
           U.ctx.handle)


def import_config_nersc(U, path):
    """import_config_nersc (field-matrix.h:592-713); returns the header as a dict."""
    h = _lib.NerscHeader()
    _check(_lib.lib().mag_import_config_nersc(U.handle, str(path).encode(), C.byref(h)), U.ctx.handle)
    return _header_dict(h)


def nersc_write(path, links, dims, skip_third=False, plaquette=0.0, link_trace=0.0, beta=0.0, sequence_number=0,
                ensemble_label=""):
    """Host-only NERSC writer for (V, 4, 18) links; returns (checksum, crc32)."""
    a = np.ascontiguousarray(links, dtype=np.float64)
    h = _lib.NerscHeader((C.c_int * 4)(*dims), 1 if skip_third else 0, link_trace, plaquette, beta,
                         sequence_number, ensemble_label.encode())
    _check(_lib.lib().mag_nersc_write(str(path).encode(), _dptr(a), C.byref(h)))
    return h.checksum, h.crc32


def nersc_read(path, header_only=False):
    """Host-only NERSC reader: (links (V, 4, 18) or None, header dict)."""
    h = _lib.NerscHeader()
    _check(_lib.lib().mag_nersc_read(str(path).encode(), None, C.byref(h)))
    if header_only:
        return None, _header_dict(h)
    a = np.empty((int(np.prod(h.dims)), 4, 18))
    _check(_lib.lib().mag_nersc_read(str(path).encode(), _dptr(a), C.byref(h)))
    return a, _header_dict(h)


def run_chmc(ctx, U_host, arg, seed=1234, **kw):
    """run_chmc (field-hmc.h:770-904) without file output: returns (final links, per-trajectory results)."""
    dims = kw.pop("dims")
    ch = Chain(ctx, dims, arg, seed=seed, **kw)
    ch.set_links(U_host)
    res = ch.run(arg.num_trajectory)
    return ch.get_links(), res
