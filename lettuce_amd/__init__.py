"""lettuce_amd -- MI355X-native gauge-field hot path of Lettuce.

A thin Python mirror of the reference's call surface over the C ABI of
liblettuce_hip.so (include/lettuce_hip.h).  Names follow the reference:

  reference (LettuceGauge/...)                       here
  GaugeField4D (lattice.hpp:87-394)                  GaugeField
  Iterator::Checkerboard / Checkerboard4             Iterator.Checkerboard / .Checkerboard4
  HeatbathKernel (updates/heatbath.hpp:30-130)       HeatbathKernel
  OverrelaxationSubgroupKernel (overrelaxation.hpp)  OverrelaxationSubgroupKernel
  PlaquetteSum (observables/plaquette.hpp:26-43)     PlaquetteSum
  TopChargeClover (topological_charge.hpp:97-231)    TopChargeClover, TopChargeCloverTimeslice
  EnergyDensity (field_strength_tensor.hpp:136-196)  EnergyDensity.Plaquette / .Clover / .CloverTimeslice
  WilsonAction::Action (wilson_action.hpp:23-41)     WilsonAction.Action
  StoutSmearing4D / StoutSmearingN (stout_smearing)  StoutSmearing4D, StoutSmearingN
  GradientFlowKernel (smearing/gradient_flow.hpp)    GradientFlowKernel
  SaveConfigBMW[Full] / LoadConfigBMW[Full]          same names (io.py; packed on the device)
  CheckpointManager (checkpoint_manager.hpp)         CheckpointManager (io.py)

Every call runs HIP kernels on the GPU; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import LettuceError, call
from .io import (CheckpointManager, LoadConfigBMW, LoadConfigBMWFull, PhiloxState, SaveConfigBMW,  # noqa: F401
                 SaveConfigBMWFull)

__all__ = [
    "Context", "GaugeField", "Iterator", "HeatbathKernel", "OverrelaxationSubgroupKernel", "PlaquetteSum",
    "TopChargeClover", "TopChargeCloverTimeslice", "TopChargeCloverDensity", "EnergyDensity", "WilsonAction",
    "StoutSmearing4D", "StoutSmearingN", "GradientFlowKernel", "Observables", "LettuceError",
    "CB2", "CB4", "SaveConfigBMW", "SaveConfigBMWFull", "LoadConfigBMW", "LoadConfigBMWFull", "CheckpointManager",
    "PhiloxState", "RectangularAction", "TopChargePlaquette", "WilsonLoop", "PolyakovLoop",
    "ParityUpdate", "MetadynamicsTemperingKernel", "LuescherWeiszAction", "IwasakiAction", "DBW2Action",
    "StapleAction", "IntegratorSchedule", "HMCKernel", "HMCMetaDKernel", "MetaBiasPotential",
    "CollectiveVariable", "MatrixField", "ReplicaBatch", "CloverDerivative", "MetaDRun", "TemperingRun",
    "UpdateSweeps",
]

CB2 = 2
CB4 = 4
FLOW_EULER = 1
FLOW_RK3 = 3

_DP = ctypes.POINTER(ctypes.c_double)


def _dp(a):
    return a.ctypes.data_as(_DP)


class Context:
    """One device and one HIP stream (lt_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        call("lt_ctx_create", device, ctypes.byref(h))
        self._h = h

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        call("lt_ctx_synchronize", self._h)

    def stream(self):
        s = ctypes.c_void_p()
        call("lt_ctx_stream", self._h, ctypes.byref(s))
        return s.value

    def profile(self, enable=True):
        call("lt_ctx_profile", self._h, 1 if enable else 0)

    def mem_info(self):
        """(free, total) device memory in bytes."""
        f = ctypes.c_size_t()
        t = ctypes.c_size_t()
        call("lt_ctx_mem_info", self._h, ctypes.byref(f), ctypes.byref(t))
        return f.value, t.value

    def read_rate(self, footprint_bytes=1 << 21):
        """Bytes/s that 16-byte lane loads read from a buffer of footprint_bytes (a power of two;
        below one XCD's L2: the L2 read ceiling)."""
        r = ctypes.c_double()
        call("lt_ctx_read_rate", self._h, footprint_bytes, ctypes.byref(r))
        return r.value

    def kernel_time(self, name):
        n = ctypes.c_int()
        ms = ctypes.c_double()
        call("lt_ctx_profile_query", self._h, name.encode(), ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("lt_ctx_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GaugeField:
    """Device gauge field with runtime extents (Nt, Nx, Ny, Nz), all even."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(x) for x in shape)
        h = ctypes.c_void_p()
        call("lt_field_create", ctx.handle, *self.shape, ctypes.byref(h))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def volume(self):
        return int(np.prod(self.shape))

    def upload(self, aos):
        a = np.ascontiguousarray(aos, dtype=np.float64)
        if a.size != self.volume * 72:
            raise ValueError("upload: array size does not match the lattice")
        call("lt_field_upload", self._h, _dp(a))

    def download(self):
        a = np.empty((self.volume, 4, 18), dtype=np.float64)
        call("lt_field_download", self._h, _dp(a))
        return a

    def SetToIdentity(self):
        call("lt_field_set_identity", self._h)

    def SetToHaar(self, seed):
        call("lt_field_set_haar", self._h, seed)

    def reunitarize(self):
        call("lt_field_reunitarize", self._h)

    def assign(self, other):
        """Copy assignment U = other (lattice.hpp:110-127)."""
        call("lt_field_copy", self._h, other.handle)

    def check(self):
        dev = ctypes.c_double()
        bad = ctypes.c_longlong()
        call("lt_field_check", self._h, ctypes.byref(dev), ctypes.byref(bad))
        return dev.value, bad.value

    def set_rng(self, seed, counter=0):
        call("lt_field_set_rng", self._h, seed, counter)

    def set_tile(self, tt, tx, ty, tz):
        """Thread-block tile (t, x, y, z/2) of the link kernels; does not change results."""
        call("lt_field_set_tile", self._h, tt, tx, ty, tz)

    def get_tile(self):
        t = (ctypes.c_int * 4)()
        call("lt_field_get_tile", self._h, t)
        return tuple(t)

    def autotune(self, family="sweep", cache=None):
        """Time the candidate tiles of the sweep or flow kernels on scratch copies of this
        field, apply the fastest; cache: optional JSON file of tuned tiles.
        Returns (tile, ms per launch)."""
        t = (ctypes.c_int * 4)()
        ms = ctypes.c_double()
        code = {"sweep": 1, "flow": 2}[family]
        call("lt_autotune_field", self._h, code, cache.encode() if cache else None, t, ctypes.byref(ms))
        return tuple(t), ms.value

    def get_rng(self):
        s = ctypes.c_uint64()
        c = ctypes.c_uint64()
        call("lt_field_get_rng", self._h, ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value

    def save_bmw(self, path, full=False, overwrite=False):
        """Write the field as a BMW file ("#BMW", rows 0-1; full: "#18BMW", all rows),
        packed on the device; returns the header's checksum (hex)."""
        buf = ctypes.create_string_buffer(32)
        call("lt_field_save_bmw", self._h, str(path).encode(), 1 if full else 0, 1 if overwrite else 0, buf, 32)
        return buf.value.decode()

    def load_bmw(self, path, full=False):
        """Read a BMW file into the field; returns (checksum_ok, max deviation from SU(3)).
        Raises LettuceError for a missing / malformed file, other extents, or links that
        are not SU(3) within 1e-6."""
        ok = ctypes.c_int()
        dev = ctypes.c_double()
        call("lt_field_load_bmw", self._h, str(path).encode(), 1 if full else 0, ctypes.byref(ok), ctypes.byref(dev))
        return bool(ok.value), dev.value

    def bmw_checksum(self, full=False):
        """Checksum_Adler64 of the field as save_bmw would write it (an integer)."""
        v = ctypes.c_ulonglong()
        call("lt_field_bmw_checksum", self._h, 1 if full else 0, ctypes.byref(v))
        return v.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("lt_field_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def Swap(a, b):
    """Pointer swap of two fields (lattice.hpp:238-246)."""
    call("lt_field_swap", a.handle, b.handle)


def Staple(U, out):
    """Staple of every link into out (rectangular_action.hpp:103-184)."""
    call("lt_staple_field", U.handle, out.handle)


class HeatbathKernel:
    """HeatbathKernel(U, Action, prng) (heatbath.hpp:106-108); beta plays the Action/global role,
    the field's Philox state plays the PRNG role."""

    def __init__(self, U, beta, action=None):
        self.U = U
        self.beta = float(beta)
        self.action = action

    def sweep(self, n, schedule):
        if self.action is None:
            call("lt_heatbath_sweep", self.U.handle, self.beta, n, schedule)
        else:
            call("lt_heatbath_sweep_action", self.U.handle, self.beta, self.action.c_plaq, self.action.c_rect, n,
                 schedule)


class OverrelaxationSubgroupKernel:
    """OverrelaxationSubgroupKernel(U, Action) (overrelaxation.hpp:68-117); action: a
    RectangularAction (default: the Wilson action)."""

    def __init__(self, U, beta=None, action=None):
        self.U = U
        self.action = action

    def sweep(self, n, schedule):
        if self.action is None:
            call("lt_or_sweep", self.U.handle, n, schedule)
        else:
            call("lt_or_sweep_action", self.U.handle, self.action.c_plaq, self.action.c_rect, n, schedule)


class Iterator:
    """Iterator::Checkerboard / Checkerboard4 (iterators/iterators.hpp:65-169): whole sweeps."""

    @staticmethod
    def Checkerboard(kernel, n_sweep=1):
        kernel.sweep(n_sweep, CB2)

    @staticmethod
    def Checkerboard4(kernel, n_sweep=1):
        kernel.sweep(n_sweep, CB4)


def UpdateSweeps(U, beta, n_heatbath=1, n_orelax=4, n_update=1, schedule=CB4, measure=True):
    """n_update iterations of the local-update loop body of SU3_main.cpp:358-385
    (Checkerboard[4] heatbath then overrelaxation sweeps), replayed as one captured graph
    per update; returns PlaquetteSum after each update (None when measure is False)."""
    out = np.zeros(n_update) if measure else None
    call("lt_update_sweeps", U.handle, beta, n_heatbath, n_orelax, schedule, n_update,
         _dp(out) if measure else None)
    return out


def Observables(U):
    """Fused plaquette / clover-Q / energy pass; returns a dict."""
    nt = U.shape[0]
    out = np.zeros(4 + 2 * nt)
    call("lt_observables", U.handle, _dp(out))
    return {"plaquette_sum": out[0], "Q": out[1], "E_plaq": out[2], "E_clov": out[3],
            "Q_t": out[4:4 + nt].copy(), "E_t": out[4 + nt:].copy()}


def PlaquetteSum(U):
    v = ctypes.c_double()
    call("lt_plaquette_sum", U.handle, ctypes.byref(v))
    return v.value


def TopChargeClover(U):
    v = ctypes.c_double()
    call("lt_topcharge_clover", U.handle, ctypes.byref(v), None)
    return v.value


def TopChargeCloverDensity(U):
    """Per-site summand of TopChargeClover (topological_charge.hpp:184-199), shape (Nt, Nx, Ny, Nz)."""
    q = np.zeros(U.shape)
    call("lt_topcharge_clover_density", U.handle, _dp(q))
    return q


def TopChargeCloverTimeslice(U):
    q = ctypes.c_double()
    qt = np.zeros(U.shape[0])
    call("lt_topcharge_clover", U.handle, ctypes.byref(q), _dp(qt))
    return qt


class EnergyDensity:
    @staticmethod
    def Plaquette(U):
        e = ctypes.c_double()
        call("lt_energy", U.handle, ctypes.byref(e), None, None)
        return e.value

    @staticmethod
    def Clover(U):
        e = ctypes.c_double()
        call("lt_energy", U.handle, None, ctypes.byref(e), None)
        return e.value

    @staticmethod
    def CloverTimeslice(U):
        et = np.zeros(U.shape[0])
        call("lt_energy", U.handle, None, None, _dp(et))
        return et


class RectangularAction:
    """GaugeAction::Rectangular<2>(beta, c_plaq, c_rect) (rectangular_action.hpp:29-100);
    ActionNormalized does not depend on beta.  Symanzik: c_plaq = 1 + 8/12, c_rect = -1/12."""

    def __init__(self, beta, c_plaq, c_rect):
        self.beta, self.c_plaq, self.c_rect = beta, c_plaq, c_rect

    def ActionNormalized(self, U):
        v = ctypes.c_double()
        call("lt_action_rectangular", U.handle, self.c_plaq, self.c_rect, ctypes.byref(v))
        return v.value

    def Action(self, U):
        return self.ActionNormalized(U) * 6.0 * self.beta * U.volume


def ParityUpdate(U, U_copy):
    """ParityUpdate (updates/parity_update.hpp:22-47): spatial reflection, Q -> -Q."""
    call("lt_parity_update", U.handle, U_copy.handle)


def LuescherWeiszAction(beta):
    """Rectangular<2>(beta, 1 + 8/12, -1/12) (rectangular_action.hpp:294)."""
    return RectangularAction(beta, 1.0 + 8.0 / 12.0, -1.0 / 12.0)


def IwasakiAction(beta):
    """Rectangular<2>(beta, 1 + 8 * 0.331, -0.331) (rectangular_action.hpp:295)."""
    return RectangularAction(beta, 1.0 + 8.0 * 0.331, -0.331)


def DBW2Action(beta):
    """Rectangular<2>(beta, 1 + 8 * 1.4088, -1.4088) (rectangular_action.hpp:296)."""
    return RectangularAction(beta, 1.0 + 8.0 * 1.4088, -1.4088)


def StapleAction(U, out, action):
    """Rectangular<R>::Staple for every link into the MatrixField out."""
    call("lt_staple_field_action", U.handle, action.c_plaq, action.c_rect, out.handle)


def TopChargePlaquette(U):
    """TopChargePlaquette (topological_charge.hpp:366-416)."""
    v = ctypes.c_double()
    call("lt_topcharge_plaquette", U.handle, ctypes.byref(v))
    return v.value


def WilsonLoop(n_start, n_end, reset, U, U_chain):
    """WilsonLoop<n_start, n_end, reset>(U, U_chain) (wilson_loop.hpp:65-126); U_chain is a
    MatrixField that keeps the link chains between calls."""
    v = ctypes.c_double()
    call("lt_wilson_loop", U.handle, U_chain.handle, n_start, n_end, 1 if reset else 0, ctypes.byref(v))
    return v.value


def PolyakovLoop(U):
    """PolyakovLoop (polyakov_loop.hpp:31-49), complex."""
    re = ctypes.c_double()
    im = ctypes.c_double()
    call("lt_polyakov_loop", U.handle, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


class WilsonAction:
    @staticmethod
    def Action(U, beta):
        v = ctypes.c_double()
        call("lt_wilson_action", U.handle, beta, ctypes.byref(v))
        return v.value


def StoutSmearing4D(U_in, U_out, rho=0.12):
    call("lt_stout", U_in.handle, U_out.handle, rho, 1)


def StoutSmearingN(U_in, U_out, n, rho=0.12):
    call("lt_stout", U_in.handle, U_out.handle, rho, n)


class GradientFlowKernel:
    """GradientFlowKernel(U_unflowed, U_flowed, Integrator, Action, epsilon) (gradient_flow.hpp:252-367)."""

    def __init__(self, U_unflowed, U_flowed, integrator="rk3", epsilon=0.01):
        code = {"rk3": FLOW_RK3, "euler": FLOW_EULER}[integrator.lower()]
        h = ctypes.c_void_p()
        call("lt_flow_create", U_unflowed.handle, U_flowed.handle, code, epsilon, ctypes.byref(h))
        self._h = h
        self.U_unflowed = U_unflowed
        self.U_flowed = U_flowed

    def __call__(self, n_step):
        call("lt_flow_run", self._h, n_step)

    def Resume(self, n_step):
        call("lt_flow_resume", self._h, n_step)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("lt_flow_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# HMC, collective variable + adjoint, metadynamics (updates/hmc_gauge.hpp,
# updates/hmc_metadynamics.hpp, actions/bias_potential/metadynamics.hpp)

LEAPFROG = 1
OMF2 = 2
OMF4 = 4
# multiple-timescale MetaD integrators (hmc_metadynamics.hpp:28-238), MetaD-HMC only
LEAPFROG_OMF4 = 5
OMF2_OMF4_SLOW = 6
OMF2_OMF4 = 7
# without merged momentum updates (hmc_gauge.hpp:21-46,121-153,195-237)
LEAPFROG_SLOW = 8
OMF2_SLOW = 9
OMF4_SLOW = 10
_INTEGRATORS = {"leapfrog": LEAPFROG, "omf2": OMF2, "omf4": OMF4, "leapfrog_omf4": LEAPFROG_OMF4,
                "omf2_omf4_slow": OMF2_OMF4_SLOW, "omf2_omf4": OMF2_OMF4, "leapfrog_slow": LEAPFROG_SLOW,
                "omf2_slow": OMF2_SLOW, "omf4_slow": OMF4_SLOW}


def IntegratorSchedule(integrator, tau, n_step):
    """The kick / drift events of an integrator as [(kind, eps, sub)]: kind 'P' (both
    forces), 'G' (gauge force), 'M' (bias force) or 'U' (field update)."""
    code = _INTEGRATORS[integrator]
    n = ctypes.c_int()
    call("lt_integrator_schedule", code, tau, n_step, None, None, None, 0, ctypes.byref(n))
    kinds = ctypes.create_string_buffer(max(n.value, 1))
    eps = np.zeros(max(n.value, 1))
    sub = (ctypes.c_int * max(n.value, 1))()
    call("lt_integrator_schedule", code, tau, n_step, kinds, _dp(eps), sub, n.value, ctypes.byref(n))
    return [(kinds.raw[i:i + 1].decode(), float(eps[i]), int(sub[i])) for i in range(n.value)]


class MatrixField(GaugeField):
    """Field of general 3x3 matrices (momenta, forces)."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(x) for x in shape)
        h = ctypes.c_void_p()
        call("lt_matrix_field_create", ctx.handle, *self.shape, ctypes.byref(h))
        self._h = h


def CloverDerivative(U, F, scale=1.0):
    """CalculateClover<1> + CloverDerivative for every link (observables/clover.hpp:59-89,
    238-321) into the matrix field F, times scale (dV/dQ in the MetaD force)."""
    call("lt_clover_derivative", U.handle, F.handle, scale)


def RandomMomentum(P, seed, counter=0):
    call("lt_hmc_random_momentum", P.handle, seed, counter)


def UpdateMomenta(U, P, eps, beta):
    call("lt_hmc_update_momenta", U.handle, P.handle, eps, beta)


def UpdateFields(U, P, eps):
    call("lt_hmc_update_fields", U.handle, P.handle, eps)


def Hamiltonian(U, P, beta):
    h = ctypes.c_double()
    call("lt_hmc_hamiltonian", U.handle, P.handle, beta, ctypes.byref(h))
    return h.value


def Integrate(U, P, beta, integrator="omf2", tau=1.0, n_step=10):
    call("lt_hmc_integrate", U.handle, P.handle, beta, _INTEGRATORS[integrator], tau, n_step)


class HMCKernel:
    """GaugeUpdates::HMCKernel(U, U_copy, Momentum, Integrator, Action, prng, ...) (hmc_gauge.hpp:381-600);
    momenta come from U's Philox state, one counter value per trajectory."""

    def __init__(self, U, U_copy, P, integrator="omf2", beta=6.0, trajectory_length=1.0):
        self.U, self.U_copy, self.P = U, U_copy, P
        self.integrator = _INTEGRATORS[integrator]
        self.beta = beta
        self.trajectory_length = trajectory_length
        self.delta_H = 0.0

    def __call__(self, n_step, metropolis_step=True):
        d = ctypes.c_double()
        a = ctypes.c_int()
        call("lt_hmc_trajectory", self.U.handle, self.P.handle, self.U_copy.handle, self.beta, self.integrator,
             self.trajectory_length, n_step, 1 if metropolis_step else 0, ctypes.byref(d), ctypes.byref(a))
        self.delta_H = d.value
        return bool(a.value)

    def ReversibilityTest(self, n_step):
        """HMCKernel::ReversibilityTest (hmc_gauge.hpp:585-599): integrate, reverse the
        momenta, integrate back; returns (H_old, H_new).  U is left at the end point."""
        h0, h1 = ctypes.c_double(), ctypes.c_double()
        call("lt_hmc_reversibility_test", self.U.handle, self.P.handle, self.beta, self.integrator,
             self.trajectory_length, n_step, ctypes.byref(h0), ctypes.byref(h1))
        return h0.value, h1.value


class CollectiveVariable:
    """Workspace of the smeared clover charge and its adjoint (HMCMetaDData, hmc_metadynamics.hpp:251-260)."""

    def __init__(self, ctx, shape, n_levels=4, rho=0.12, checkpoint=0):
        """checkpoint = k > 0 keeps only every k-th stout level and recomputes the rest
        in the backward pass (identical results, less memory)."""
        h = ctypes.c_void_p()
        call("lt_cv_create_checkpointed", ctx.handle, *shape, n_levels, rho, checkpoint, ctypes.byref(h))
        self._h = h
        self.ctx = ctx      # the context must outlive the workspace
        self.shape = tuple(shape)

    @property
    def handle(self):
        return self._h

    def forward(self, U):
        q = ctypes.c_double()
        call("lt_cv_forward", self._h, U.handle, ctypes.byref(q))
        return q.value

    def backward(self, dvdq, F):
        call("lt_cv_backward", self._h, dvdq, F.handle)

    def bytes(self):
        n = ctypes.c_size_t()
        call("lt_cv_bytes", self._h, ctypes.byref(n))
        return n.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("lt_cv_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MetaBiasPotential:
    """MetaBiasPotential(CV_min, CV_max, bin_number, weight, width, well_tempered_parameter, threshold_weight)
    (metadynamics.hpp:58-91); well_tempered selects the update rule (a compile-time switch in the reference)."""

    def __init__(self, cv_min, cv_max, bins, weight, width, well_tempered_parameter=1.0, threshold_weight=1000.0,
                 well_tempered=False):
        h = ctypes.c_void_p()
        call("lt_bias_create", cv_min, cv_max, bins, weight, width, 1 if well_tempered else 0,
             well_tempered_parameter, threshold_weight, ctypes.byref(h))
        self._h = h
        self.bins = bins

    @property
    def handle(self):
        return self._h

    def UpdatePotential(self, cv):
        call("lt_bias_deposit", self._h, cv, 0)

    def UpdatePotentialSymmetric(self, cv):
        call("lt_bias_deposit", self._h, cv, 1)

    def ReturnPotential(self, cv):
        v = ctypes.c_double()
        call("lt_bias_potential", self._h, cv, ctypes.byref(v))
        return v.value

    def ReturnDerivative(self, cv):
        v = ctypes.c_double()
        call("lt_bias_derivative", self._h, cv, ctypes.byref(v))
        return v.value

    def grid(self):
        a = np.zeros(self.bins + 1)
        call("lt_bias_grid", self._h, _dp(a), self.bins + 1, 0)
        return a

    def set_grid(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        call("lt_bias_grid", self._h, _dp(a), self.bins + 1, 1)

    def SaveParameters(self, filename, overwrite=False):
        """MetaBiasPotential::SaveParameters (metadynamics.hpp:304-324); appends unless overwrite."""
        call("lt_bias_save_parameters", self._h, str(filename).encode(), 1 if overwrite else 0)

    def SavePotential(self, filename, overwrite=False):
        """MetaBiasPotential::SavePotential (metadynamics.hpp:326-346); appends unless overwrite."""
        call("lt_bias_save_potential", self._h, str(filename).encode(), 1 if overwrite else 0)

    def LoadPotential(self, filename):
        """MetaBiasPotential::LoadPotential (metadynamics.hpp:351-505): parameters and the
        last potential snapshot of the file.  Returns False (and leaves the potential
        unchanged) where the reference returns false."""
        try:
            call("lt_bias_load_potential", self._h, str(filename).encode())
        except LettuceError:
            return False
        self.bins = self.info()["bin_number"]
        return True

    def info(self):
        p = np.zeros(6)
        n = ctypes.c_int()
        e = ctypes.c_ulonglong()
        call("lt_bias_info", self._h, _dp(p), ctypes.byref(n), ctypes.byref(e))
        return {"CV_min": p[0], "CV_max": p[1], "weight": p[2], "width": p[3], "well_tempered_parameter": p[4],
                "threshold_weight": p[5], "bin_number": n.value, "exceeded_count": e.value}

    def ReturnCV_current(self):
        v = ctypes.c_double()
        call("lt_bias_cv", self._h, ctypes.byref(v), 0)
        return v.value

    def SetCV_current(self, cv):
        v = ctypes.c_double(cv)
        call("lt_bias_cv", self._h, ctypes.byref(v), 1)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("lt_bias_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HMCMetaDKernel:
    """GaugeUpdates::HMCMetaDKernel (hmc_metadynamics.hpp:264-675) with the adjoint force.

    integrator: "leapfrog", "omf2", "omf4" (both forces in every kick) or the
    multiple-timescale "leapfrog_omf4", "omf2_omf4_slow", "omf2_omf4" (bias force on
    the outer scale, gauge force on an inner OMF4).  With hmc_log_path set, every
    trajectory appends its CV path to that file as LogCVData does (the visited values;
    with the Metropolis step a second line with the values deposited)."""

    def __init__(self, U, U_copy, P, F, bias, cv, integrator="omf2", beta=6.0, trajectory_length=1.0,
                 update_bias=True, hmc_log_path=None):
        self.U, self.U_copy, self.P, self.F, self.bias, self.cv = U, U_copy, P, F, bias, cv
        self.integrator = _INTEGRATORS[integrator]
        self.beta = beta
        self.trajectory_length = trajectory_length
        self.update_bias = update_bias
        self.hmc_log_path = hmc_log_path
        self.delta_H = 0.0
        self.cv_value = 0.0

    def __call__(self, n_step, metropolis_step=True):
        d = ctypes.c_double()
        a = ctypes.c_int()
        q = ctypes.c_double()
        call("lt_metad_hmc_trajectory", self.U.handle, self.P.handle, self.U_copy.handle, self.F.handle,
             self.cv.handle, self.bias.handle, self.beta, self.integrator, self.trajectory_length, n_step,
             1 if metropolis_step else 0, 1 if self.update_bias else 0, ctypes.byref(d), ctypes.byref(a),
             ctypes.byref(q))
        self.delta_H = d.value
        self.cv_value = q.value
        if self.hmc_log_path is not None and n_step > 0:
            self.LogCVData(False)
            if metropolis_step:
                self.LogCVData(True)
        return bool(a.value)

    def ReversibilityTest(self, n_step):
        """HMCMetaDKernel::ReversibilityTest (hmc_metadynamics.hpp:657-674): integrate with the
        bias, reverse the momenta, integrate back; returns (H_old, H_new) with V(Q) included.
        U is left at the end point; the bias is not updated."""
        h0, h1 = ctypes.c_double(), ctypes.c_double()
        call("lt_metad_reversibility_test", self.U.handle, self.P.handle, self.F.handle, self.cv.handle,
             self.bias.handle, self.beta, self.integrator, self.trajectory_length, n_step, ctypes.byref(h0),
             ctypes.byref(h1))
        return h0.value, h1.value

    def path_samples(self, deposited=False):
        """cv_path_samples of the last trajectory: the visited CVs, or the deposited values."""
        n = ctypes.c_int()
        call("lt_cv_path_samples", self.cv.handle, 1 if deposited else 0, None, 0, ctypes.byref(n))
        out = np.zeros(n.value)
        call("lt_cv_path_samples", self.cv.handle, 1 if deposited else 0, _dp(out), n.value, ctypes.byref(n))
        return out

    def LogCVData(self, deposited=False):
        """HMCMetaDKernel::LogCVData (hmc_metadynamics.hpp:537-545): one line appended to hmc_log_path."""
        call("lt_cv_log_path", self.cv.handle, 1 if deposited else 0, str(self.hmc_log_path).encode())


class MetadynamicsTemperingKernel:
    """MetadynamicsTemperingKernel (updates/tempering.hpp:24-73): swap proposals between
    the unbiased stream U and the MetaD stream U_temper; each call draws from Philox
    stream 4 at (seed, counter) and advances the counter."""

    def __init__(self, U, U_temper, bias, cv, seed=0, counter=0):
        self.U, self.U_temper, self.bias, self.cv = U, U_temper, bias, cv
        self.seed, self.counter = seed, counter
        self.delta_V = 0.0
        self.acceptances = 0

    def __call__(self):
        acc = ctypes.c_int()
        dv = ctypes.c_double()
        call("lt_metad_tempering_swap", self.U.handle, self.U_temper.handle, self.cv.handle, self.bias.handle,
             self.seed, self.counter, ctypes.byref(acc), ctypes.byref(dv))
        self.counter += 1
        self.delta_V = dv.value
        self.acceptances += acc.value
        return bool(acc.value)

    def get_rng(self):
        """(seed, counter) of the swap proposals' Philox stream (saved by PhiloxState)."""
        return self.seed, self.counter

    def set_rng(self, seed, counter=0):
        self.seed, self.counter = int(seed), int(counter)


class _ReplicaField(GaugeField):
    """Gauge field of replica r of a ReplicaBatch; the batch owns its memory and
    the view keeps the batch alive.  After the batch is closed the view raises."""

    def __init__(self, batch, r):
        self.ctx = batch.ctx
        self.shape = batch.shapes[r]
        self._batch = batch
        self._r = r

    @property
    def _h(self):
        bh = self._batch.handle
        if bh is None or not bh.value:
            raise LettuceError("lt_batch_field", -4, "the batch of this replica was closed")
        f = ctypes.c_void_p()
        call("lt_batch_field", bh, self._r, ctypes.byref(f))
        return f

    def close(self):
        pass


SWAP_MODES = {"none": 0, "even": 1, "odd": 2, "alternate": 3}


class ReplicaBatch:
    """n lattices (equal or mixed extents) advanced by the same launches (lt_batch).

    Generalises the reference's two-stream tempering (MetadynamicsTemperingKernel,
    updates/tempering.hpp:24-73, SU3_main.cpp:549-621) to a ladder of bias
    strengths: replica r runs MetaD-HMC (HMCMetaDKernel) with bias
    strengths[label[r]] * V(Q); adjacent labels exchange with probability
    min(1, exp(-(s_k - s_k+1) (V(Q_j) - V(Q_i)))); the replica with the top label
    deposits its path samples into the shared bias.
    """

    def __init__(self, ctx, shapes, n_levels=4, rho=0.12):
        self.ctx = ctx
        self.shapes = [tuple(int(x) for x in s) for s in shapes]
        flat = (ctypes.c_int * (4 * len(self.shapes)))(*[x for s in self.shapes for x in s])
        h = ctypes.c_void_p()
        call("lt_batch_create", ctx.handle, len(self.shapes), flat, n_levels, rho, ctypes.byref(h))
        self._h = h
        self.n = len(self.shapes)
        self._n_sub = 0

    @property
    def handle(self):
        return self._h

    def field(self, r):
        if not 0 <= r < self.n:
            raise IndexError("replica index out of range")
        return _ReplicaField(self, r)

    def set_rng(self, r, seed, counter=0):
        call("lt_batch_set_rng", self._h, r, seed, counter)

    def get_rng(self, r):
        s = ctypes.c_uint64()
        c = ctypes.c_uint64()
        call("lt_batch_get_rng", self._h, r, ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value

    def set_ladder(self, strengths, labels=None, pt_seed=0, pt_counter=0):
        s = np.ascontiguousarray(strengths, dtype=np.float64)
        if s.size != self.n:
            raise ValueError("one strength per replica")
        lab = None
        if labels is not None:
            lab = (ctypes.c_int * self.n)(*[int(x) for x in labels])
        call("lt_batch_set_ladder", self._h, _dp(s), lab, pt_seed, pt_counter)

    def set_bias(self, bias):
        call("lt_batch_set_bias", self._h, bias.handle)

    def get_bias(self, bias):
        call("lt_batch_get_bias", self._h, bias.handle)

    def use_graph(self, enable=True):
        call("lt_batch_use_graph", self._h, 1 if enable else 0)

    def autotune(self, cache=None):
        """Tune the thread-block tile of the batched MetaD kernels per replica extent."""
        call("lt_batch_autotune", self._h, cache.encode() if cache else None)

    def metad_hmc(self, n_step, beta=6.0, integrator="omf2", tau=1.0, metropolis=True, update_bias=False,
                  swap="none", n_traj=1):
        call("lt_batch_metad_hmc", self._h, beta, _INTEGRATORS[integrator], tau, n_step, 1 if metropolis else 0,
             1 if update_bias else 0, SWAP_MODES[swap], n_traj)
        self._n_sub = n_step

    def results(self):
        dH = np.zeros(self.n)
        cv = np.zeros(self.n)
        acc = (ctypes.c_int * self.n)()
        lab = (ctypes.c_int * self.n)()
        call("lt_batch_results", self._h, _dp(dH), acc, _dp(cv), lab)
        return {"dH": dH, "accepted": np.array(acc[:], dtype=bool), "cv": cv, "labels": np.array(lab[:])}

    def samples(self):
        a = np.zeros((self.n, self._n_sub))
        call("lt_batch_samples", self._h, _dp(a), a.size)
        return a

    def swap_stats(self):
        a = (ctypes.c_longlong * (2 * self.n))()
        call("lt_batch_swap_stats", self._h, a, 2 * self.n)
        return np.array(a[:], dtype=np.int64).reshape(self.n, 2)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("lt_batch_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .driver import MetaDRun, TemperingRun  # noqa: E402,F401
