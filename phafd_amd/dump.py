"""The dump command over the C ABI (Dump, src/dump.cpp:28-300): one file per
dumped timestep, its parallel header naming the one piece, and the .pvd
collection listing the timesteps.

    dump <name> grid <base> <every> attributes n phi chempot ...
    dump <name> atom/owned <base> <every> attributes n x ux ix F ID molID type
         [decimalplaces p] [fmt lf|e]

File names follow src/dump.cpp:302-318: <base>_p0_<step>.vti|vtp for the
data, <base>_<step>.pvti|pvtp for the header, <base>.pvd for the collection.
"""
import os

from . import api as A


class Dump:
    def __init__(self, system, name, dump_type, base_name, every, attributes, decimalplaces=None, fmt="lf",
                 async_=True):
        if dump_type.startswith("atom"):
            which = dump_type[5:]
            if which != "owned":
                raise ValueError("only atom/owned dumps are written here (one subdomain holds every atom)")
            self.kind, self.fext = "atom", "vtp"
        elif dump_type == "grid":
            self.kind, self.fext = "grid", "vti"
        else:
            raise ValueError("Invalid dump type (must be atom or grid or ftgrid).")
        if not attributes:
            raise ValueError("Need attributes to write dump files.")
        if len(set(attributes)) != len(attributes):
            raise ValueError("Duplicate attribute values in dump file.")
        if decimalplaces is not None and decimalplaces <= 0:
            raise ValueError("Dump precision must be greater than or equal to zero.")
        self.system, self.name, self.every = system, name, int(every)
        self.attributes = list(attributes)
        self.precision = -1 if decimalplaces is None else int(decimalplaces)
        self.fmt = A._fmt(fmt)
        self.async_ = async_
        self.base_name = base_name
        self.nopath_base_name = os.path.basename(base_name)
        self.collection_name = base_name + ".pvd"

    def write_collection_header(self):
        A.vtk_collection(self.collection_name, A.PVD_HEADER)

    def write_collection_footer(self):
        self.system.dump_wait()
        A.vtk_collection(self.collection_name, A.PVD_FOOTER)

    def due(self, timestep):
        return timestep % self.every == 0

    def write_collection_middle(self, timestep, dt):
        """The data-set line, then this timestep's file and its header."""
        s = self.system
        instance = f"{self.base_name}_p0_{timestep}.{self.fext}"
        pinstance = f"{self.base_name}_{timestep}.p{self.fext}"
        source = f"{self.nopath_base_name}_p0_{timestep}.{self.fext}"
        A.vtk_collection(self.collection_name, A.PVD_DATASET, timestep * dt,
                         f"{self.nopath_base_name}_{timestep}.p{self.fext}")
        if self.kind == "atom":
            s.dump_atoms(instance, self.attributes, self.precision, self.fmt, self.async_)
            A.vtk_write_patoms(pinstance, self.attributes, source)
        else:
            s.dump_grid(instance, self.attributes, self.async_)
            spacing = [p / n for p, n in zip(s.period, s.shape[::-1])]
            A.vtk_write_pgrid(pinstance, s.shape, self.attributes, s.boxlo, spacing, source)


def run_with_dumps(system, nsteps, dumps, dt, first_step=0):
    """Integrate::run with dumps: step, and after every step write the dumps
    that are due (src/integrate.cpp:180-190); the runs between dumps are
    single calls, so the steps are not synchronised for the dumps."""
    step = first_step
    end = first_step + nsteps
    while step < end:
        nxt = min([end] + [((step // d.every) + 1) * d.every for d in dumps])
        system.run(nxt - step)
        step = nxt
        for d in dumps:
            if d.due(step):
                d.write_collection_middle(step, dt)
    return step
