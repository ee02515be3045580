"""The update loops of the reference's main program over the library's kernels.

  reference (SU3_main.cpp)                                  here
  MetaD-HMC loop, random ParityUpdate (:523-531)            MetaDRun.step
  two-stream tempering loop (:600-620)                      TemperingRun.step

Each step does what one iteration of the reference's loop does to the gauge fields
and the bias; measurements, bias snapshots and checkpoints stay with the caller, who
calls Observables / SavePotential / CheckpointManager at its own periods as the
reference's loop body does.
"""
import numpy as np

from . import Iterator, ParityUpdate


class MetaDRun:
    """HMC_MetaD(n_hmc, accept_reject) followed by ParityUpdate(U, U_copy) with
    probability 1/2 (SU3_main.cpp:523-531).  The coin comes from parity_rng (a numpy
    Generator; the reference draws it from a host std::uniform_int_distribution).  As
    in the reference, the bias's current CV is not flipped with the field: the next
    trajectory measures CV_old afresh."""

    def __init__(self, hmc_metad, U_copy, parity_rng=None):
        self.hmc_metad = hmc_metad
        self.U_copy = U_copy
        self.rng = parity_rng if parity_rng is not None else np.random.default_rng(0)

    def step(self, n_hmc, accept_reject=True):
        """Returns (accepted, parity_flipped)."""
        accepted = self.hmc_metad(n_hmc, accept_reject)
        flipped = bool(self.rng.integers(0, 2))
        if flipped:
            ParityUpdate(self.hmc_metad.U, self.U_copy)
        return accepted, flipped


class TemperingRun:
    """One iteration of the two-stream tempering loop (SU3_main.cpp:600-616):
    tempering_nonmetadynamics_sweeps rounds of Checkerboard4 heatbath (n_heatbath
    sweeps) and overrelaxation (n_orelax sweeps) on the unbiased stream, one MetaD-HMC
    trajectory on the biased stream, and a swap proposal every tempering_swap_period
    iterations (MetadynamicsTemperingKernel)."""

    def __init__(self, heatbath, overrelaxation, hmc_metad, tempering, n_heatbath=1, n_orelax=4,
                 tempering_nonmetadynamics_sweeps=1, tempering_swap_period=1):
        if tempering.U is not heatbath.U or tempering.U_temper is not hmc_metad.U:
            raise ValueError("TemperingRun: the tempering kernel must swap the two streams the updates run on")
        if tempering_swap_period < 1:
            raise ValueError("TemperingRun: tempering_swap_period must be >= 1")
        self.heatbath, self.overrelaxation = heatbath, overrelaxation
        self.hmc_metad, self.tempering = hmc_metad, tempering
        self.n_heatbath, self.n_orelax = int(n_heatbath), int(n_orelax)
        self.nonmetadynamics_sweeps = int(tempering_nonmetadynamics_sweeps)
        self.swap_period = int(tempering_swap_period)

    def step(self, n_count, n_hmc, accept_reject=True):
        """Returns (MetaD trajectory accepted, swap accepted or None when none was proposed)."""
        for _ in range(self.nonmetadynamics_sweeps):
            Iterator.Checkerboard4(self.heatbath, self.n_heatbath)
            Iterator.Checkerboard4(self.overrelaxation, self.n_orelax)
        accepted = self.hmc_metad(n_hmc, accept_reject)
        swapped = self.tempering() if n_count % self.swap_period == 0 else None
        return accepted, swapped
