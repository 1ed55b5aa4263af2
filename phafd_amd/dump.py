"""The dump command over the C ABI (Dump, src/dump.cpp:28-313): one file per
dumped timestep, its parallel header naming the one piece, and the .pvd
collection listing the timesteps.

    dump <name> grid <base> <every> attributes n phi chempot ...
    dump <name> atom/owned <base> <every> attributes n x ux ix F ID molID type
         [decimalplaces p] [fmt lf|e]
    dump <name> lammpsatom <base> <every> attributes n ...   (one text file)
    dump <name> ftgrid <base> <every> attributes n c_<id> f_<id> ...   (compute complex, fix grid/ave)

File names follow src/dump.cpp:315-334: <base>_p0_<step>.vti|vtp for the
data, <base>_<step>.pvti|pvtp for the header, <base>.pvd for the collection.
"""
import os

import numpy as np

from . import api as A


class Dump:
    def __init__(self, system, name, dump_type, base_name, every, attributes, decimalplaces=None, fmt="lf",
                 async_=True, computes=None):
        if dump_type == "lammpsatom":
            self.kind, self.fext = "lammps", "lmmpstrj"
        elif dump_type.startswith("atom"):
            which = dump_type[5:]
            if which != "owned":
                raise ValueError("only atom/owned dumps are written here (one subdomain holds every atom)")
            self.kind, self.fext = "atom", "vtp"
        elif dump_type == "grid":
            self.kind, self.fext = "grid", "vti"
        elif dump_type == "ftgrid":
            self.kind, self.fext = "ftgrid", "vti"
        else:
            raise ValueError("Invalid dump type (must be atom or grid or ftgrid).")
        if not attributes:
            raise ValueError("Need attributes to write dump files.")
        if len(set(attributes)) != len(attributes):
            raise ValueError("Duplicate attribute values in dump file.")
        if decimalplaces is not None and decimalplaces <= 0:
            raise ValueError("Dump precision must be greater than or equal to zero.")
        # c_<id> attributes: compute pair arrays, {id: pair index}
        self.computes = dict(computes or {})
        for w in attributes:
            if self.kind == "ftgrid" and w.startswith("f_") and w[2:] in self.computes:
                continue   # a fix grid/ave over a compute complex
            if w.startswith("c_") and w[2:] not in self.computes:
                raise ValueError("Cannot write dump, compute ID " + w[2:] + "doesn't exist.")
            if w.startswith("c_") and self.kind == "grid" and not hasattr(self.computes[w[2:]], "threshold"):
                raise ValueError("Cannot write dump, compute ID " + w[2:] + "is not a per_grid quantity.")
            if self.kind == "ftgrid" and not (w.startswith("c_") or w.startswith("f_")):
                raise ValueError("Dump error: Attribute does not exist.")
        self._computed = None
        self.system, self.name, self.every = system, name, int(every)
        self.attributes = list(attributes)
        self.precision = -1 if decimalplaces is None else int(decimalplaces)
        self.fmt = A._fmt(fmt)
        self.async_ = async_
        self.base_name = base_name
        self.nopath_base_name = os.path.basename(base_name)
        self.collection_name = base_name + ".pvd"

    def write_collection_header(self):
        """The collection header; a lammpsatom dump instead creates its one
        file (Dump::setup, src/dump.cpp:204-207, 334-346)."""
        if self.kind == "lammps":
            open(f"{self.base_name}_p0.{self.fext}", "w").close()
            return
        A.vtk_collection(self.collection_name, A.PVD_HEADER)

    def write_collection_footer(self):
        self.system.dump_wait()
        if self.kind != "lammps":
            A.vtk_collection(self.collection_name, A.PVD_FOOTER)

    def due(self, timestep):
        return timestep % self.every == 0

    def needs_capture(self, timestep):
        """An ftgrid dump of a compute complex needs the in_fourier spectra of
        the step it is written at."""
        from .computes import ComputeComplex
        return self.kind == "ftgrid" and self.due(timestep) and any(
            isinstance(c, ComputeComplex) for c in self.computes.values())

    def write_collection_middle(self, timestep, dt):
        """The data-set line, then this timestep's file and its header."""
        self.dt = dt
        s = self.system
        if self.kind == "ftgrid":
            self._write_ftgrid(timestep)
            return
        if self.kind == "lammps":   # no collection line, one file for every step
            vals = self._atom_values()
            top = [lo + p for lo, p in zip(s.boxlo, s.period)]
            A.lammps_write_step(f"{self.base_name}_p0.{self.fext}", timestep, [vals[w] for w in self.attributes],
                                self.attributes, s.boxlo, top, self.precision, self.fmt)
            return
        instance = f"{self.base_name}_p0_{timestep}.{self.fext}"
        pinstance = f"{self.base_name}_{timestep}.p{self.fext}"
        source = f"{self.nopath_base_name}_p0_{timestep}.{self.fext}"
        A.vtk_collection(self.collection_name, A.PVD_DATASET, timestep * dt,
                         f"{self.nopath_base_name}_{timestep}.p{self.fext}")
        if self.kind == "atom":
            if self._computed is not None or any(w.startswith("c_") for w in self.attributes):
                self._atoms_with_computes(instance)
            else:
                s.dump_atoms(instance, self.attributes, self.precision, self.fmt, self.async_)
            A.vtk_write_patoms(pinstance, self.attributes, source)
        else:
            spacing = [p / n for p, n in zip(s.period, s.shape[::-1])]
            if any(w.startswith("c_") for w in self.attributes):
                # with per-grid computes the file is written from host arrays
                if self._computed is None:
                    self.end_of_step()
                fields = {"phi": A.PHI, "chempot": A.MU, "gradphi_x": A.GRADX, "gradphi_y": A.GRADY,
                          "gradphi_z": A.GRADZ}
                arrays = [self._computed[w] if w.startswith("c_") else s.get_field(fields[w])
                          for w in self.attributes]
                self._computed = None
                A.vtk_write_grid(instance, arrays, self.attributes, s.boxlo, spacing)
            else:
                s.dump_grid(instance, self.attributes, self.async_)
            A.vtk_write_pgrid(pinstance, s.shape, self.attributes, s.boxlo, spacing, source)


    def _write_ftgrid(self, timestep):
        """Dump::write_grid_timestep for ftgrid (src/dump.cpp:703-723): the
        per-mode compute arrays in the ft order [ky][kz][kx], origin 0 and the
        spacings dqx, dqy = 2pi/Lz, dqz = 2pi/Ly of the reference's Domain
        (src/domain.cpp:206-208, pi as 3.141592).  One rank: each array holds
        exactly its Ny*Nz*(Nx/2+1) values (the reference's one-rank writer
        announces one plane more than the compute holds)."""
        from .computes import ComputeComplex
        s = self.system
        instance = f"{self.base_name}_p0_{timestep}.{self.fext}"
        pinstance = f"{self.base_name}_{timestep}.p{self.fext}"
        source = f"{self.nopath_base_name}_p0_{timestep}.{self.fext}"
        A.vtk_collection(self.collection_name, A.PVD_DATASET, timestep * self.dt,
                         f"{self.nopath_base_name}_{timestep}.p{self.fext}")
        arrays = []
        for w in self.attributes:
            c = self.computes[w[2:]]
            arrays.append(c.array() if isinstance(c, ComputeComplex) else c.array)
        lx, ly, lz = s.period
        dq = [2 * 3.141592 / lx, 2 * 3.141592 / lz, 2 * 3.141592 / ly]
        A.vtk_write_grid(instance, arrays, self.attributes, [0.0, 0.0, 0.0], dq)
        A.vtk_write_pgrid(pinstance, arrays[0].shape, self.attributes, [0.0, 0.0, 0.0], dq, source)

    def end_of_step(self):
        """The computes this dump holds (ComputePair::end_of_step), evaluated
        before any dump of the step is written, as the reference's step runs
        its computes' end_of_step before its dumps (src/integrate.cpp:172-190):
        a grid-atom compute's second chempot deposit is then in a chempot
        dump of the same step."""
        if self.kind == "ftgrid":   # its computes are read when the file is written
            self._computed = None
            return
        if self.kind == "grid":     # per-grid computes (grid clusters) of this step
            self._computed = {w: self.computes[w[2:]].array() for w in self.attributes if w.startswith("c_")} or None
            return
        self._computed = {w: self.system.compute_pair(self.computes[w[2:]])
                          for w in self.attributes if w.startswith("c_")} or None

    def _atom_values(self):
        """Every atom attribute of this step from the device, computes included."""
        s = self.system
        if self._computed is None:
            self.end_of_step()
        x, ux, img = s.get_atoms()
        vals = {"x": x, "ux": ux, "ix": img, "F": s.get_forces(), "label": np.zeros(len(x), np.int32)}
        vals.update(s.atom_ints)
        vals.update(self._computed or {})
        self._computed = None
        return vals

    def _atoms_with_computes(self, instance):
        """An atom dump holding compute pair arrays, written from the host."""
        vals = self._atom_values()
        A.vtk_write_atoms(instance, [vals[w] for w in self.attributes], self.attributes, self.precision, self.fmt)


def run_with_dumps(system, nsteps, dumps, dt, first_step=0, aves=()):
    """Integrate::run with dumps and fix grid/ave: the runs stop at every step
    a dump or an average needs (src/integrate.cpp:105-203).  At such a step
    the computes run first (ComputePair::end_of_step), then the averages take
    their samples (the run's last step kept its in_fourier spectra for them),
    then the dumps are written (src/integrate.cpp:172-190).  The runs between
    stops are single calls, so the steps are not synchronised for them."""
    step = first_step
    end = first_step + nsteps
    # FixGridAve::setup refuses a dump of the fix whose every differs from
    # the fix's freq (src/fixgrid_ave.cpp:101-107)
    for d in dumps:
        for w in d.attributes:
            fix = d.computes.get(w[2:]) if w.startswith("f_") else None
            if hasattr(fix, "freq") and d.every != fix.freq:
                raise ValueError("Using fixgrid/ave with frequency that doesn't match dump frequency - "
                                 "waste of time.")
    for a in aves:   # FixGridAve::setup at the start of every run (src/fixgrid_ave.cpp:99-125)
        a.setup(first_step)
    while step < end:
        nxt = min([end] + [((step // d.every) + 1) * d.every for d in dumps] + [a.next_sample(step) for a in aves])
        sampling = [a for a in aves if a.samples_at(nxt)]
        spectra = bool(sampling) or any(d.needs_capture(nxt) for d in dumps)
        if aves or any(d.kind == "ftgrid" for d in dumps):
            system.fourier_capture(spectra)
        system.run(nxt - step)
        step = nxt
        due = [d for d in dumps if d.due(step)]
        for d in due:
            d.end_of_step()
        for a in sampling:
            a.end_of_step(step)
        for d in due:
            d.write_collection_middle(step, dt)
    if aves or any(d.kind == "ftgrid" for d in dumps):
        system.fourier_capture(False)
    return step
