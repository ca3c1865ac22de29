"""ORACLE — test infrastructure only.

ctypes binding of the FP64 CPU restatement of the reference constrained-HMC
path (oracle/oracle.cpp). Only tests/, ``__graft_entry__.smoke()`` and the
benchmark's CPU-baseline leg import this module; the shipped library never
does. Fields are numpy float64 arrays of shape (V, 4, 3, 3) complex128 or the
equivalent (V, 4, 18) reals, in the reference's qlat layout.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Action(C.Structure):
    _fields_ = [("type", C.c_int), ("beta", C.c_double), ("c1", C.c_double), ("mag", C.c_int),
                ("n_stout", C.c_int), ("rho", C.c_double)]


class Hmc(C.Structure):
    _fields_ = [("trajectory_length", C.c_int), ("dt", C.c_double), ("integrator", C.c_int),
                ("num_forced_accept", C.c_int), ("accept_mode", C.c_int), ("demon_emax", C.c_double),
                ("seed", C.c_uint64), ("reunitarize", C.c_int)]


class Traj(C.Structure):
    _fields_ = [("h_old", C.c_double), ("h_new", C.c_double), ("dh", C.c_double), ("die_roll", C.c_double),
                ("accept", C.c_int), ("avg_plaq", C.c_double), ("s_gauge", C.c_double),
                ("demon_e", C.c_double * 2), ("reunit_dev", C.c_double)]


class Multi(C.Structure):
    _fields_ = [("rho", C.c_double), ("xi0", C.c_double), ("md_alpha", C.c_double)]


class MultiTraj(C.Structure):
    _fields_ = [("h_old", C.c_double), ("h_new", C.c_double), ("dh", C.c_double), ("die_roll", C.c_double),
                ("accept", C.c_int), ("plaq_fine", C.c_double), ("plaq_coarse", C.c_double),
                ("parts_old", C.c_double * 3), ("parts_new", C.c_double * 3), ("reunit_dev", C.c_double * 2)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U32 = C.POINTER(C.c_uint32)


def _declare(l):
    sig = {
        "orc_philox4x32_10": (None, [_U32, _U32, _U32]),
        "orc_generators": (None, [_D]),
        "orc_expiQ": (None, [_D, _D, C.c_long]),
        "orc_DexpiQ": (None, [_D, _D, _D, C.c_long]),
        "orc_compute_Lambda": (None, [_D, _D, _D, _D, C.c_long]),
        "orc_haar_links": (None, [_I, C.c_uint64, _D]),
        "orc_unit_links": (None, [_I, _D]),
        "orc_init_momentum": (None, [_I, C.c_uint64, C.c_uint64, C.c_double, _D]),
        "orc_uniform": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint64]),
        "orc_path_product": (None, [_I, _D, _I, _I, C.c_int, _D]),
        "orc_staple_dagger": (None, [_I, _D, C.c_long, C.c_int, _D]),
        "orc_rectangular_dagger": (None, [_I, _D, C.c_long, C.c_int, _D]),
        "orc_chair_staple_dagger": (None, [_I, _D, C.c_long, C.c_int, _D]),
        "orc_staple_index": (C.c_int, [C.c_int, C.c_int, _I, C.c_int]),
        "orc_twist_staple_dagger": (None, [_I, _D, C.c_long, C.c_int, _D]),
        "orc_is_constrained": (C.c_int, [_I, C.c_int, C.c_int]),
        "orc_total_plaq": (C.c_double, [_I, _D]),
        "orc_total_rectangular": (C.c_double, [_I, _D]),
        "orc_avg_plaquette": (C.c_double, [_I, _D]),
        "orc_kinetic": (C.c_double, [_I, _D, C.c_int]),
        "orc_gauge_action": (C.c_double, [_I, _D, C.POINTER(Action)]),
        "orc_hamiltonian": (C.c_double, [_I, _D, _D, C.POINTER(Action), _D]),
        "orc_reunitarize": (C.c_double, [_I, _D]),
        "orc_get_force": (None, [_I, _D, _D, C.POINTER(Action)]),
        "orc_evolve_gauge_field": (None, [_I, _D, _D, C.c_double, C.c_int]),
        "orc_evolve_momentum": (None, [_I, _D, _D, C.c_double, C.c_double]),
        "orc_stout_smear": (None, [_I, _D, C.c_double, C.c_int, _D]),
        "orc_stout_Q": (None, [_I, _D, C.c_double, _D]),
        "orc_force_gradient_integrator": (C.c_double, [_I, _D, _D, C.POINTER(Action), C.c_int, C.c_double, C.c_int]),
        "orc_leap_frog_integrator": (None, [_I, _D, _D, C.POINTER(Action), C.c_int, C.c_double]),
        "orc_avg_wilson_loop": (C.c_double, [_I, _D, _I, C.c_int]),
        "orc_run_hmc": (C.c_int, [_I, _D, C.POINTER(Action), C.POINTER(Hmc), C.c_int, C.c_int, _D,
                                  C.POINTER(Traj)]),
        "orc_demon_microcanonical": (None, [_I, _D, C.c_int, C.c_double, C.c_uint64, _D,
                                            C.POINTER(C.c_long)]),
        "orc_coarsen": (None, [_I, _D, C.c_int, _D]),
        "orc_update_constrain": (None, [_I, _D, _D, C.c_int]),
        "orc_constrained_plaquette": (C.c_double, [_I, _D, C.c_int]),
        "orc_expand": (None, [_I, _D, C.c_int, _D]),
        "orc_xi_energy": (C.c_double, [_I, _D, _D, _D, C.POINTER(Multi)]),
        "orc_multi_gb": (None, [_I, _D, C.POINTER(Multi), _D]),
        "orc_multi_Q": (None, [_I, _D, C.POINTER(Multi), _D]),
        "orc_multi_fforce": (None, [_I, _D, _D, _D, C.POINTER(Action), C.POINTER(Multi), _D]),
        "orc_multi_cforce": (None, [_I, _D, _D, C.POINTER(Multi), _D]),
        "orc_multi_DQ": (None, [_I, _D, _D, C.POINTER(Multi), _D]),
        "orc_multi_heatbath": (None, [_I, _D, _D, C.POINTER(Multi), C.c_uint64, C.c_uint64]),
        "orc_multi_zeta_T": (None, [_I, _D, _D, _D, C.c_double, C.c_int, C.POINTER(Multi)]),
        "orc_multi_integrate": (None, [_I, _D, _D, _D, _D, C.POINTER(Action), C.POINTER(Multi), C.c_int, C.c_int,
                                       C.c_double, _D]),
        "orc_multi_hamiltonian": (C.c_double, [_I, _D, _D, _D, _D, C.POINTER(Action), C.POINTER(Multi), _D]),
        "orc_run_hmc_multi": (C.c_int, [_I, _D, _D, C.POINTER(Action), C.POINTER(Hmc), C.POINTER(Multi), C.c_int,
                                        C.c_int, C.POINTER(MultiTraj)]),
        "orc_local_metropolis": (None, [_I, _D, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                        C.c_int, C.POINTER(C.c_long), _D, _D]),
        "orc_metropolis_link": (None, [_D, _D, C.c_double, C.c_int, C.c_double, C.c_uint64, C.c_uint64,
                                       C.c_uint32, C.POINTER(C.c_long)]),
        "orc_demon_checkerboard": (None, [_I, _D, C.c_int, C.c_double, C.c_uint64, _D, C.POINTER(C.c_long)]),
        "orc_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args


# ----------------------------------------------------------------- helpers
def _dims(L):
    return (C.c_int * 4)(*[int(v) for v in L])


def _dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def as_real(m):
    """complex (..., 3, 3) -> float64 (..., 18) view-compatible copy."""
    m = np.ascontiguousarray(m, dtype=np.complex128)
    return m.view(np.float64).reshape(m.shape[:-2] + (18,))


def as_cplx(r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    return r.view(np.complex128).reshape(r.shape[:-1] + (3, 3))


def volume(L):
    return int(np.prod(L))


def action(type=0, beta=6.0, c1=0.0, mag=0, n_stout=0, rho=0.1):
    return Action(type, beta, c1, mag, n_stout, rho)


def hmc(trajectory_length=8, dt=0.125, integrator=0, num_forced_accept=0, accept_mode=0, demon_emax=0.0,
        seed=1234, reunitarize=1):
    return Hmc(trajectory_length, dt, integrator, num_forced_accept, accept_mode, demon_emax, seed, reunitarize)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def generators():
    out = np.zeros((8, 18))
    lib().orc_generators(_dptr(out))
    return as_cplx(out)


def _batch(fn, *mats):
    arrs = [as_real(np.asarray(m)).reshape(-1, 18) for m in mats]
    n = arrs[0].shape[0]
    out = np.zeros((n, 18))
    fn(*[_dptr(a) for a in arrs], _dptr(out), n)
    return as_cplx(out).reshape(np.asarray(mats[0]).shape)


def expiQ(Q):
    return _batch(lib().orc_expiQ, Q)


def DexpiQ(Q, DQ):
    return _batch(lib().orc_DexpiQ, Q, DQ)


def compute_Lambda(Q, S, U):
    return _batch(lib().orc_compute_Lambda, Q, S, U)


def haar_links(L, seed):
    U = np.zeros((volume(L), 4, 18))
    lib().orc_haar_links(_dims(L), seed, _dptr(U))
    return U


def unit_links(L):
    U = np.zeros((volume(L), 4, 18))
    lib().orc_unit_links(_dims(L), _dptr(U))
    return U


def init_momentum(L, seed, traj, m=1.0):
    P = np.zeros((volume(L), 4, 18))
    lib().orc_init_momentum(_dims(L), seed, traj, m, _dptr(P))
    return P


def uniform(seed, stream, index):
    return lib().orc_uniform(seed, stream, index)


def path_product(L, U, x, dirs):
    out = np.zeros(18)
    xs = (C.c_int * 4)(*x)
    ds = (C.c_int * len(dirs))(*dirs)
    lib().orc_path_product(_dims(L), _dptr(U), xs, ds, len(dirs), _dptr(out))
    return as_cplx(out)


def _site(fn, L, U, site, mu):
    out = np.zeros(18)
    fn(_dims(L), _dptr(U), site, mu, _dptr(out))
    return as_cplx(out)


def staple_dagger(L, U, site, mu):
    return _site(lib().orc_staple_dagger, L, U, site, mu)


def rectangular_dagger(L, U, site, mu):
    return _site(lib().orc_rectangular_dagger, L, U, site, mu)


def chair_staple_dagger(L, U, site, mu):
    return _site(lib().orc_chair_staple_dagger, L, U, site, mu)


def twist_staple_dagger(L, U, site, mu):
    return _site(lib().orc_twist_staple_dagger, L, U, site, mu)


def staple_index(mu, twist):
    """The oracle's chair (twist=False) or twist path table of direction mu:
    a list of 5-step direction-code paths (codes as path_product takes them)."""
    buf = (C.c_int * (5 * 256))()
    n = lib().orc_staple_index(mu, 1 if twist else 0, buf, 256)
    return [list(buf[5 * i:5 * i + 5]) for i in range(n)]


def is_constrained(x, mu, mag):
    return lib().orc_is_constrained((C.c_int * 4)(*x), mu, mag)


def total_plaq(L, U):
    return lib().orc_total_plaq(_dims(L), _dptr(U))


def total_rectangular(L, U):
    return lib().orc_total_rectangular(_dims(L), _dptr(U))


def avg_plaquette(L, U):
    return lib().orc_avg_plaquette(_dims(L), _dptr(U))


def kinetic(L, P, mag=0):
    return lib().orc_kinetic(_dims(L), _dptr(P), mag)


def gauge_action(L, U, act):
    return lib().orc_gauge_action(_dims(L), _dptr(U), C.byref(act))


def hamiltonian(L, U, P, act):
    parts = np.zeros(2)
    h = lib().orc_hamiltonian(_dims(L), _dptr(U), _dptr(P), C.byref(act), _dptr(parts))
    return h, parts


def reunitarize(L, U):
    return lib().orc_reunitarize(_dims(L), _dptr(U))


def get_force(L, U, act):
    F = np.zeros_like(U)
    lib().orc_get_force(_dims(L), _dptr(U), _dptr(F), C.byref(act))
    return F


def evolve_gauge_field(L, U, P, dt, mag=0):
    lib().orc_evolve_gauge_field(_dims(L), _dptr(U), _dptr(P), dt, mag)


def evolve_momentum(L, P, F, dt, m=1.0):
    lib().orc_evolve_momentum(_dims(L), _dptr(P), _dptr(F), dt, m)


def stout_smear(L, U, rho, nlev):
    out = np.zeros_like(U)
    lib().orc_stout_smear(_dims(L), _dptr(U), rho, nlev, _dptr(out))
    return out


def stout_Q(L, U, rho):
    out = np.zeros_like(U)
    lib().orc_stout_Q(_dims(L), _dptr(U), rho, _dptr(out))
    return out


def force_gradient_integrator(L, U, P, act, steps, dt, reunit=True):
    return lib().orc_force_gradient_integrator(_dims(L), _dptr(U), _dptr(P), C.byref(act), steps, dt,
                                               1 if reunit else 0)


def leap_frog_integrator(L, U, P, act, steps, dt):
    lib().orc_leap_frog_integrator(_dims(L), _dptr(U), _dptr(P), C.byref(act), steps, dt)


def avg_wilson_loop(L, U, r, T):
    return lib().orc_avg_wilson_loop(_dims(L), _dptr(U), (C.c_int * 3)(*r), T)


def run_hmc(L, U, act, h, first_traj, ntraj, demon_state=None):
    out = (Traj * ntraj)()
    ds = np.zeros(4) if demon_state is None else demon_state
    nacc = lib().orc_run_hmc(_dims(L), _dptr(U), C.byref(act), C.byref(h), first_traj, ntraj, _dptr(ds), out)
    return nacc, [out[i] for i in range(ntraj)], ds


def demon_microcanonical(L, U, nsweeps, E_max, seed, E=None):
    E = np.zeros(4) if E is None else E
    ar = (C.c_long * 2)()
    lib().orc_demon_microcanonical(_dims(L), _dptr(U), nsweeps, E_max, seed, _dptr(E), ar)
    return E, (ar[0], ar[1])


def coarsen(L, U, mag):
    """Coarse links of a mag-blocked lattice (double_multigrid, field-hmc.h:1068-1078)."""
    Lc = [v // mag for v in L]
    out = np.zeros((volume(Lc), 4, 18))
    lib().orc_coarsen(_dims(L), _dptr(U), mag, _dptr(out))
    return out


def update_constrain(L, U, Uc, mag):
    """update_constrain (field-hmc.h:988-1009), in place on U."""
    lib().orc_update_constrain(_dims(L), _dptr(U), _dptr(Uc), mag)


def constrained_plaquette(L, U, mag):
    """check_constrained_plaquette (field-matrix.h:407-443)."""
    return lib().orc_constrained_plaquette(_dims(L), _dptr(U), mag)


def expand(Lc, Uc, mag):
    """lattice_expansion (expansion/main.C:19-62) without file I/O."""
    Lf = [v * mag for v in Lc]
    out = np.zeros((volume(Lf), 4, 18))
    lib().orc_expand(_dims(Lc), _dptr(Uc), mag, _dptr(out))
    return out


# ------------------------------------------- md:: coarse/fine HMC (field-multi-md.h)
# Fine extents L (even); coarse extents L/2; xi is (Vc, 4) doubles.
def multi(rho=0.1, xi0=1.0, md_alpha=0.5):
    return Multi(rho, xi0, md_alpha)


def _coarse(L):
    return [v // 2 for v in L]


def xi_energy(L, U, Uc, xi, mp):
    """get_xi_energy (field-multi-md.h:30-72)."""
    return lib().orc_xi_energy(_dims(L), _dptr(U), _dptr(Uc), _dptr(np.ascontiguousarray(xi, dtype=np.float64)),
                               C.byref(mp))


def multi_gb(L, U, mp):
    """get_Gb (field-multi-md.h:771-783): e^{iQ} U(2x) U(2x+mu) per coarse link."""
    out = np.zeros((volume(_coarse(L)), 4, 18))
    lib().orc_multi_gb(_dims(L), _dptr(U), C.byref(mp), _dptr(out))
    return out


def multi_Q(L, U, mp):
    """get_Q (field-multi.h:33-39) at every coarse link."""
    out = np.zeros((volume(_coarse(L)), 4, 18))
    lib().orc_multi_Q(_dims(L), _dptr(U), C.byref(mp), _dptr(out))
    return out


def multi_fforce(L, U, Uc, xi, act, mp):
    """md::get_Fforce (field-multi-md.h:116-368)."""
    out = np.zeros((volume(L), 4, 18))
    lib().orc_multi_fforce(_dims(L), _dptr(U), _dptr(Uc), _dptr(np.ascontiguousarray(xi, dtype=np.float64)),
                           C.byref(act), C.byref(mp), _dptr(out))
    return out


def multi_cforce(L, U, P, mp):
    """md::get_Cforce (field-multi-md.h:398-425)."""
    out = np.zeros((volume(_coarse(L)), 4, 18))
    lib().orc_multi_cforce(_dims(L), _dptr(U), _dptr(P), C.byref(mp), _dptr(out))
    return out


def multi_DQ(L, U, P, mp):
    """md::get_DQ (field-multi-md.h:370-396) at every coarse link."""
    out = np.zeros((volume(_coarse(L)), 4, 18))
    lib().orc_multi_DQ(_dims(L), _dptr(U), _dptr(P), C.byref(mp), _dptr(out))
    return out


def multi_heatbath(L, Uc, U, mp, seed, traj):
    """heatbath (field-multi-md.h:757-769), in place on Uc."""
    lib().orc_multi_heatbath(_dims(L), _dptr(Uc), _dptr(U), C.byref(mp), seed, traj)


def multi_zeta_T(L, U, P, Uc, dt, mag, mp):
    """force_gradient_integrator_zeta_T (field-multi-md.h:441-475), in place."""
    lib().orc_multi_zeta_T(_dims(L), _dptr(U), _dptr(P), _dptr(Uc), dt, mag, C.byref(mp))


MULTI_ZETA_H, MULTI_FG, MULTI_NESTED = 0, 1, 2


def multi_integrate(L, U, P, Uc, xi, act, mp, which, steps, dt):
    """zeta_H (477-597), md::force_gradient_integrator (599-661) or
    nested_integrator (663-695), in place; returns the reunitarize deviations."""
    dev = np.zeros(2)
    lib().orc_multi_integrate(_dims(L), _dptr(U), _dptr(P), _dptr(Uc),
                              _dptr(np.ascontiguousarray(xi, dtype=np.float64)), C.byref(act), C.byref(mp), which,
                              steps, dt, _dptr(dev))
    return dev


def multi_hamiltonian(L, U, P, Uc, xi, act, mp):
    """md::get_hamiltonian_multi (field-multi-md.h:74-114): (H, [K, S, S_xi])."""
    parts = np.zeros(3)
    h = lib().orc_multi_hamiltonian(_dims(L), _dptr(U), _dptr(P), _dptr(Uc),
                                    _dptr(np.ascontiguousarray(xi, dtype=np.float64)), C.byref(act), C.byref(mp),
                                    _dptr(parts))
    return h, parts


def run_hmc_multi(L, U, Uc, act, h, mp, first_traj, ntraj):
    """md::run_hmc_multi (field-multi-md.h:785-988) without file output, in place."""
    out = (MultiTraj * ntraj)()
    nacc = lib().orc_run_hmc_multi(_dims(L), _dptr(U), _dptr(Uc), C.byref(act), C.byref(h), C.byref(mp),
                                   first_traj, ntraj, out)
    return nacc, [out[i] for i in range(ntraj)]


# ------------------------------------ local Metropolis (alg_gchb.h:107-387)
def local_metropolis(L, U, beta, mag, num_iter, nhits, small, seed, order=0):
    """algGCtrnHeatBath::run in place on U; returns (acc[4], plaq[iter], cons_plaq[iter])."""
    acc = (C.c_long * 4)()
    p, cp = np.zeros(num_iter), np.zeros(num_iter)
    lib().orc_local_metropolis(_dims(L), _dptr(U), beta, mag, num_iter, nhits, small, seed, order, acc,
                               _dptr(p), _dptr(cp))
    return list(acc), p, cp


def metropolis_link(U, sigma, beta, nhits, small, seed, it, link, acc=None):
    """metropolisKernel (alg_gchb.h:128-183) on one 3x3 link with a fixed staple sum, in place."""
    acc = (C.c_long * 2)() if acc is None else acc
    lib().orc_metropolis_link(_dptr(U), _dptr(sigma), beta, nhits, small, seed, it, link, acc)
    return acc


def demon_checkerboard(L, U, nsweeps, E_max, seed, Elink=None):
    """Checkerboard demon with per-link demons, in place on U and Elink
    ((V, 4, 4): site, mu, {plaquette, rectangle, chair, twist}); returns (Elink, (acc, rej))."""
    Elink = np.zeros((volume(L), 4, 4)) if Elink is None else Elink
    ar = (C.c_long * 2)()
    lib().orc_demon_checkerboard(_dims(L), _dptr(U), nsweeps, E_max, seed, _dptr(Elink), ar)
    return Elink, (ar[0], ar[1])


def num_threads():
    return lib().orc_num_threads()
