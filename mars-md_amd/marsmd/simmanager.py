"""The SimManager step loop above one Patch: force calculation, integration, energy
output with momenta settled first, and the native restart format with firstStep
continuation.

Follows src/SimManager.cpp:
  * run()                        :235-303 (steps first_step+1 .. first_step+num_steps,
                                  then settle_momenta_for_output(last) and the final
                                  restart)
  * execute_force_calculation()  :337-425 (maybe_reorder, nonbonded and bonded forces,
                                  energy on steps where step % energyOutputPeriod == 0)
  * handle_output()              :541-573
  * settle_momenta_for_output()  :520-539 (a pending BAOAB closing half-kick is applied
                                  with forces recomputed at the current positions)
  * write_energy_output()        :1000-1075 ("Kinetic Energy: <kT> (kT) " and "Potential
                                  Energy: <kcal/mol> (kcal/mol) " lines, std::ostream's
                                  default 6 significant digits)
  * write_restart_files()        :1083-1140 ("type x y z" per particle in global-id
                                  order, std::to_chars shortest round-trip floats;
                                  <name>.restart, and <name>.0.momentum.restart for
                                  Langevin)
Steps between outputs run through the Patch's fused step (Patch.run); the steps with
output go through the same calls, so a run's trajectory does not depend on its output
periods. Trajectory (DCD) writing is not part of this module.
"""
from __future__ import annotations

from dataclasses import replace
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

from . import _capi as api
from .patch import Patch
from .system import SystemSpec

SQRT_CAL_TO_JOULE = 2.046167337  # src/Constants.h:34


# --------------------------------------------------------------------- number formats
def float_to_chars(v) -> str:
    """std::to_chars(float) without a format: the shortest string that parses back to
    the same float, in fixed or scientific notation, whichever is shorter (fixed on a
    tie); scientific as printf's %e writes it (sign and at least two exponent digits)."""
    f = np.float32(v)
    if np.isnan(f):
        return "-nan" if np.signbit(f) else "nan"
    if np.isinf(f):
        return "-inf" if f < 0 else "inf"
    sign = "-" if np.signbit(f) else ""
    a = np.abs(f)
    if a == 0:
        return sign + "0"
    sci = np.format_float_scientific(a, unique=True, trim="-")  # d[.ddd]e[+-]x
    mant, exp = sci.split("e")
    e = int(exp)
    digits = mant.replace(".", "")
    # scientific: d[.ddd]e+XX
    s_sci = digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + ("e-" if e < 0 else "e+") + f"{abs(e):02d}"
    # fixed: the same significant digits placed by the exponent
    if e >= len(digits) - 1:
        s_fix = digits + "0" * (e - len(digits) + 1)
    elif e >= 0:
        s_fix = digits[:e + 1] + "." + digits[e + 1:]
    else:
        s_fix = "0." + "0" * (-e - 1) + digits
    return sign + (s_fix if len(s_fix) <= len(s_sci) else s_sci)


def ostream_double(v: float) -> str:
    """std::ostream << double with the default flags (%g, 6 significant digits)."""
    return "%g" % float(v)


def restart_line(type_id: int, v) -> str:
    """append_restart_line (SimManager.cpp:14-28): "type x y z\\n"."""
    return f"{int(type_id)} {float_to_chars(v[0])} {float_to_chars(v[1])} {float_to_chars(v[2])}\n"


def read_restart(path) -> Tuple[np.ndarray, np.ndarray]:
    """A native .restart file ("type x y z" per line, file order = global id) as
    (type ids (n,) int32, vectors (n, 3) float32), each value parsed as a float."""
    types, vecs = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if len(tok) != 4:
                raise ValueError(f"{path}: expected 'type x y z', got {line!r}")
            types.append(int(tok[0]))
            vecs.append([np.float32(t) for t in tok[1:]])
    return np.asarray(types, np.int32), np.asarray(vecs, np.float32).reshape(-1, 3)


def restart_spec(spec: SystemSpec, restart_path, momentum_path=None) -> SystemSpec:
    """The spec with positions (and momenta) from restart files, types checked against
    it, for a run continued with firstStep."""
    types, pos = read_restart(restart_path)
    if len(types) != spec.n or not np.array_equal(types, np.asarray(spec.type_ids, np.int32)):
        raise ValueError(f"{restart_path}: particle count or types differ from the system")
    mom = spec.momenta
    if momentum_path is not None:
        mtypes, mom = read_restart(momentum_path)
        if not np.array_equal(mtypes, types):
            raise ValueError(f"{momentum_path}: types differ from {restart_path}")
    return replace(spec, positions=pos, momenta=mom)


# ----------------------------------------------------------------------- the manager
class SimManager:
    """SimManager over one Patch. `output_period` / `energy_output_period` in steps
    (0: never); energy output writes <output_name>.energy.dat and the restart files."""

    def __init__(self, spec: SystemSpec, output_name, num_steps: int, first_step: int = 0, output_period: int = 0,
                 energy_output_period: int = 0, patch: Optional[Patch] = None):
        self.spec = spec
        self.output_name = str(output_name)
        self.num_steps = int(num_steps)
        self.first_step = int(first_step)
        self.output_period = int(output_period)
        self.energy_output_period = int(energy_output_period)
        self.params = replace(spec.params, energy_period=self.energy_output_period)
        self.patch = patch if patch is not None else Patch(spec)
        self.has_momentum_output = spec.params.integrator == api.INTEGRATOR_LANGEVIN
        self._energy_file = None
        self.energies = []  # (step, kinetic kT, potential kcal/mol) per energy output

    def close(self):
        if self._energy_file is not None:
            self._energy_file.close()
            self._energy_file = None

    # ---------------------------------------------------------------- step pieces
    def _compute_energy(self, step: int) -> bool:
        return self.energy_output_period > 0 and step % self.energy_output_period == 0

    def execute_force_calculation(self, step: int):
        P = self.params
        self.patch.maybe_reorder(step, P)
        self.patch.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, P.electric_field,
                                              P.interpolation, self._compute_energy(step))
        self.patch._prepared = True
        self.patch.calculate_bonded_forces(self._compute_energy(step))

    def execute_integration(self, step: int):
        P = self.params
        self.patch.integrate_motion(P.timestep, P.kT, P.integrator, step)

    def settle_momenta_for_output(self, step: int):
        if not self.patch.has_deferred_kick():
            return
        self.execute_force_calculation(step)
        self.patch.finish_deferred_kick(self.params.timestep)

    def handle_output(self, step: int):
        trajectory_due = self.output_period > 0 and step % self.output_period == 0
        energy_due = self._compute_energy(step)
        if (trajectory_due and self.has_momentum_output) or energy_due:
            self.settle_momenta_for_output(step)
        if energy_due:
            self.write_energy_output(step)

    # ---------------------------------------------------------------- output
    def gather(self, need_energy: bool = False) -> dict:
        """Particles in global-id order (gather_particle_data_from_patches)."""
        h = self.patch.copy_particles_to_host()
        order = np.argsort(h["id"], kind="stable")
        out = {k: v[order] for k, v in h.items()}
        if not need_energy:
            out.pop("force")
        return out

    def energies_of(self, particles: dict) -> Tuple[float, float]:
        """(kinetic energy in kT, potential energy in kcal/mol) as write_energy_output
        forms them: FP32 energies summed in FP64; 0.5 |p|^2 / m with |p|^2 in FP32,
        summed in FP64 and divided by SQRT_CAL_TO_JOULE^2 and kT."""
        # in particle order, one FP64 add at a time (SimManager.cpp:1007-1010)
        e = particles["force"][:, 3].astype(np.float64)
        pot = float(np.cumsum(e)[-1]) if len(e) else 0.0
        mom = particles["mom"][:, :3].astype(np.float32)
        p2 = (mom[:, 0] * mom[:, 0] + mom[:, 1] * mom[:, 1]) + mom[:, 2] * mom[:, 2]
        mass = np.array([t.mass for t in self.spec.types], np.float32)[particles["type"]].astype(np.float64)
        terms = 0.5 * p2.astype(np.float64) / mass
        ke = (float(np.cumsum(terms)[-1]) if len(terms) else 0.0) / (SQRT_CAL_TO_JOULE * SQRT_CAL_TO_JOULE)
        kT = float(np.float32(self.params.kT))
        return (ke / kT if kT > 0.0 else 0.0), pot

    def write_energy_output(self, step: int):
        particles = self.gather(need_energy=True)
        ke, pot = self.energies_of(particles)
        if self._energy_file is None:
            self._energy_file = open(self.output_name + ".energy.dat", "w")
        self._energy_file.write(f"Kinetic Energy: {ostream_double(ke)} (kT) \n")
        self._energy_file.write(f"Potential Energy: {ostream_double(pot)} (kcal/mol) \n")
        self._energy_file.flush()
        self.energies.append((step, ke, pot))
        self.write_restart_files(particles)

    def write_restart_files(self, particles: Optional[dict] = None):
        p = particles if particles is not None else self.gather()
        with open(self.output_name + ".restart", "w") as f:
            f.write("".join(restart_line(t, v) for t, v in zip(p["type"], p["pos"][:, :3])))
        if self.has_momentum_output:
            with open(self.output_name + ".0.momentum.restart", "w") as f:
                f.write("".join(restart_line(t, v) for t, v in zip(p["type"], p["mom"][:, :3])))

    def write_final_restart(self):
        self.write_restart_files()

    # ---------------------------------------------------------------- the loop
    def _next_output(self, step: int, last: int) -> int:
        nxt = last
        for period in (self.output_period, self.energy_output_period):
            if period > 0:
                nxt = min(nxt, (step // period + 1) * period)
        return nxt

    def run(self):
        """Steps first_step+1 .. first_step+num_steps. The steps up to each output step
        run fused on the device; the output step's output follows it."""
        step = self.first_step
        last = self.first_step + self.num_steps
        while step < last:
            nxt = self._next_output(step, last)
            self.patch.run(nxt - step, first_step=step, params=self.params)
            step = nxt
            self.handle_output(step)
        self.settle_momenta_for_output(last)
        self.write_final_restart()
        self.close()
        return self


def continued_run(spec: SystemSpec, output_name, restart_prefix, first_step: int, num_steps: int,
                  **periods) -> SimManager:
    """A run that resumes from <restart_prefix>.restart (and its momentum file for
    Langevin) with firstStep = first_step, so every step draws the noise of its
    absolute step number."""
    mom = None
    if spec.params.integrator == api.INTEGRATOR_LANGEVIN:
        mom = Path(str(restart_prefix) + ".0.momentum.restart")
    resumed = restart_spec(spec, str(restart_prefix) + ".restart", mom)
    return SimManager(resumed, output_name, num_steps, first_step=first_step, **periods)


__all__ = ["SimManager", "continued_run", "float_to_chars", "read_restart", "restart_line", "restart_spec",
           "ostream_double", "SQRT_CAL_TO_JOULE"]
