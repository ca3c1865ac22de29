"""The coupled coarse/fine HMC of field-multi-md.h (namespace md::), with the
reference's names, over the device kernels of libmag_amd.so (FP64).

Fine links carry even extents; the coarse links have half of them, coarse site
c sitting at fine site 2c. The reference leaves rho and XI0 to its includer
and fixes md_alpha = 0.5 (field-multi-md.h:28); here they are the fields of
MultiParams.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import ALG8, F64, LINK18, REAL1, HmcParams, MultiParams, MultiResult
from . import Field, _act, _check, _dptr

MD_ALPHA = 0.5
ZETA_H, FG, NESTED = 0, 1, 2


def params(rho, xi0, md_alpha=MD_ALPHA):
    return MultiParams(rho, xi0, md_alpha)


def coarse_dims(dims):
    return tuple(d // 2 for d in dims)


def xi_field(ctx, fine_dims, values):
    """The CxField (qlat Field<double>, one real per coarse link) from a
    scalar or a (Vc, 4) array."""
    x = Field(ctx, coarse_dims(fine_dims), REAL1, F64)
    vc = int(np.prod(coarse_dims(fine_dims)))
    return x.upload(np.broadcast_to(np.asarray(values, dtype=np.float64), (vc, 4)))


def init_xi(CxField, mp):
    """init_xi (field-multi-md.h:697-745): xi = XI0 on every coarse link."""
    vc = int(np.prod(CxField.dims))
    CxField.upload(np.full((vc, 4), mp.xi0))


def get_xi_energy(FgField, CgField, CxField, mp):
    """get_xi_energy (field-multi-md.h:30-72)."""
    v = C.c_double()
    _check(_lib.lib().mag_xi_energy(FgField.handle, CgField.handle, CxField.handle, C.byref(mp), C.byref(v)),
           FgField.ctx.handle)
    return v.value


def get_Q(CQField, FgField, mp):
    """get_Q (field-multi.h:32-39) for every coarse link into a coarse FP64 algebra field."""
    _check(_lib.lib().mag_multi_get_Q(CQField.handle, FgField.handle, C.byref(mp)), FgField.ctx.handle)
    return CQField


def get_Gb(CGbField, FgField, mp):
    """get_Gb (field-multi-md.h:771-783)."""
    _check(_lib.lib().mag_multi_gb(CGbField.handle, FgField.handle, C.byref(mp)), FgField.ctx.handle)
    return CGbField


def get_Fforce(FfField, FgField, CgField, Farg, CxField, mp):
    """md::get_Fforce (field-multi-md.h:116-368)."""
    act = _act(Farg)
    _check(_lib.lib().mag_multi_fforce(FfField.handle, FgField.handle, CgField.handle, CxField.handle, C.byref(act),
                                       C.byref(mp)), FgField.ctx.handle)
    return FfField


def get_Cforce(CfField, FgField, FmField, mp):
    """md::get_Cforce (field-multi-md.h:398-425)."""
    _check(_lib.lib().mag_multi_cforce(CfField.handle, FgField.handle, FmField.handle, C.byref(mp)),
           FgField.ctx.handle)
    return CfField


def heatbath(CgField, FgField, mp, seed, traj):
    """heatbath (field-multi-md.h:757-769): one Metropolis hit per coarse link."""
    _check(_lib.lib().mag_multi_heatbath(CgField.handle, FgField.handle, C.byref(mp), seed, traj),
           FgField.ctx.handle)


def force_gradient_integrator_zeta_T(FgField, FmField, CgField, dt, Farg, mp):
    """force_gradient_integrator_zeta_T (field-multi-md.h:441-475)."""
    _check(_lib.lib().mag_multi_zeta_T(FgField.handle, FmField.handle, CgField.handle, dt, Farg.mag, C.byref(mp)),
           FgField.ctx.handle)


def _integrate(which, FgField, FmField, CgField, CxField, Farg, mp, steps):
    act = _act(Farg)
    dev = np.zeros(2)
    _check(_lib.lib().mag_multi_integrate(FgField.handle, FmField.handle, CgField.handle, CxField.handle,
                                          C.byref(act), C.byref(mp), which, steps, Farg.dt, _dptr(dev)),
           FgField.ctx.handle)
    return dev


def force_gradient_integrator_zeta_H(FgField, FmField, CgField, CxField, Farg, mp):
    """force_gradient_integrator_zeta_H (field-multi-md.h:477-597), Farg.trajectory_length
    steps of Farg.dt; returns the reunitarize deviations (fine, coarse)."""
    return _integrate(ZETA_H, FgField, FmField, CgField, CxField, Farg, mp, Farg.trajectory_length)


def force_gradient_integrator(FgField, FmField, Farg, steps, CgField, CxField, mp):
    """md::force_gradient_integrator (field-multi-md.h:599-661): fine fields only,
    step Farg.dt (the reference ignores its dt argument)."""
    return _integrate(FG, FgField, FmField, CgField, CxField, Farg, mp, steps)


def nested_integrator(FgField, FmField, Farg, CgField, CxField, mp):
    """nested_integrator (field-multi-md.h:663-695)."""
    return _integrate(NESTED, FgField, FmField, CgField, CxField, Farg, mp, Farg.trajectory_length)


def get_hamiltonian_multi(FgField, FmField, Farg, CgField, CxField, mp):
    """md::get_hamiltonian_multi (field-multi-md.h:74-114): (H, [kinetic, gauge, xi])."""
    act = _act(Farg)
    parts = np.zeros(3)
    h = C.c_double()
    _check(_lib.lib().mag_multi_hamiltonian(FgField.handle, FmField.handle, CgField.handle, CxField.handle,
                                            C.byref(act), C.byref(mp), _dptr(parts), C.byref(h)),
           FgField.ctx.handle)
    return h.value, parts


class MultiChain:
    """md::run_hmc_multi (field-multi-md.h:785-988) as a device-resident chain."""

    def __init__(self, ctx, dims, Farg, mp, seed=1234):
        self.ctx = ctx
        self.dims = tuple(dims)
        self.act = _act(Farg)
        self.mp = mp
        self.params = HmcParams(Farg.trajectory_length, Farg.dt, 0, Farg.num_forced_accept_step, 0, 0.0, seed, 1,
                                F64, 0)
        self._h = C.c_void_p()
        _check(_lib.lib().mag_multi_create(ctx.handle, (C.c_int * 4)(*self.dims), C.byref(self.act),
                                           C.byref(self.params), C.byref(mp), C.byref(self._h)), ctx.handle)
        ctx.adopt(self)

    def _release(self):
        if self._h:
            _lib.lib().mag_multi_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def fields(self):
        """Borrowed (fine, coarse) link fields of the chain."""
        f, c = C.c_void_p(), C.c_void_p()
        _check(_lib.lib().mag_multi_fields(self._h, C.byref(f), C.byref(c)), self.ctx.handle)
        return (Field(self.ctx, self.dims, LINK18, F64, _handle=f),
                Field(self.ctx, coarse_dims(self.dims), LINK18, F64, _handle=c))

    def run(self, ntraj):
        res = (MultiResult * max(ntraj, 1))()
        _check(_lib.lib().mag_multi_run(self._h, ntraj, res), self.ctx.handle)
        return [res[i] for i in range(ntraj)]


def run_hmc_multi(ctx, dims, FgField_host, Farg, CgField_host, mp, seed=1234):
    """run_hmc_multi without file output, from host links of fine extents dims:
    returns (fine links, coarse links, per-trajectory results) after
    Farg.num_trajectory trajectories."""
    ch = MultiChain(ctx, dims, Farg, mp, seed)
    UF, UC = ch.fields()
    UF.upload(FgField_host)
    UC.upload(CgField_host)
    res = ch.run(Farg.num_trajectory)
    return UF.download(), UC.download(), res


__all__ = ["MD_ALPHA", "params", "coarse_dims", "xi_field", "init_xi", "get_xi_energy", "get_Gb", "get_Fforce",
           "get_Cforce", "heatbath", "force_gradient_integrator_zeta_T", "force_gradient_integrator_zeta_H",
           "force_gradient_integrator", "nested_integrator", "get_hamiltonian_multi", "MultiChain",
           "run_hmc_multi"]
