"""compute complex and fix grid/ave over the C ABI.

    compute <id> complex ft_phi|ft_chempot real|imag|modulus|norm
        ComputeComplex (src/compute_complex.cpp): the spectrum as it is at
        the in_fourier hook, per mode, in the reference's ft array order
    fix <id> grid/ave <every> <repeat> <freq> c_<id>
        FixGridAve (src/fixgrid_ave.cpp:92-215): the average of repeat
        samples taken every `every` steps, the last at a multiple of freq;
        accumulated on the device
"""
from . import api as A

ARRAYS = {"ft_phi": A.PHIHAT, "ft_chempot": A.MUHAT}
QUANTS = {"real": A.CPX_REAL, "imag": A.CPX_IMAG, "modulus": A.CPX_MODULUS, "norm": A.CPX_NORM}


class ComputeComplex:
    def __init__(self, system, name, arrname, quant):
        if arrname not in ARRAYS:
            raise ValueError("Invalid array to compute in compute modulus.")
        if quant not in QUANTS:
            raise ValueError("Invalid argument in compute complex.")
        self.system, self.name = system, name
        self.which, self.quant = ARRAYS[arrname], QUANTS[quant]

    def array(self):
        """The compute at the last step of the last run made with
        system.fourier_capture(True)."""
        return self.system.compute_complex(self.which, self.quant)


class ComputeGridClusters:
    """compute <id> grid/clusters phi <threshold> lt|gt
    (src/compute_grid_clusters.cpp:22-60): a per-grid compute, evaluated on
    the current phi when read."""

    def __init__(self, system, name, arrname, threshold, condition):
        if arrname != "phi":
            raise ValueError("Must use phi as array for clustering.")
        if condition not in ("lt", "gt"):
            raise ValueError("Must specify lt or gt for compute grid cluster.")
        self.system, self.name = system, name
        self.threshold, self.condition = float(threshold), condition

    def array(self):
        return self.system.compute_grid_clusters(self.threshold, self.condition)


def save_steps(every, repeat, freq, nwindows):
    """The step counts (steps since the fix's setup) at which grid/ave samples,
    over nwindows windows (src/fixgrid_ave.cpp:113-140, 170-185)."""
    out = []
    for w in range(1, nwindows + 1):
        out += [w * freq - every * i for i in range(repeat - 1, -1, -1)]
    return out


class FixGridAve:
    def __init__(self, system, name, every, repeat, freq, compute, slot=0):
        if every <= 0 or repeat <= 0 or freq <= 0:
            raise ValueError("Every, repeat, and freq must all be >= 0.")
        if every * repeat > freq:
            raise ValueError("Every and repeat exceed freq in fixgridave.")
        if not isinstance(compute, ComputeComplex):
            raise ValueError("Need per_grid or per_ftgrid compute for fixgridave.")
        self.system, self.name, self.compute, self.slot = system, name, compute, slot
        self.every, self.repeat, self.freq = int(every), int(repeat), int(freq)
        self.setup()

    def setup(self, first_step=0):
        self.first_step = first_step
        self.next_freq = self.freq
        self.savesteps = [self.next_freq - self.every * i for i in range(self.repeat - 1, -1, -1)]
        self.array = None

    def next_sample(self, step):
        """The first sample step (absolute) after step."""
        c = step - self.first_step
        nxt = [s for s in self.savesteps if s > c]
        if nxt:
            return self.first_step + nxt[0]
        return self.first_step + self.next_freq + self.freq - self.every * (self.repeat - 1)

    def samples_at(self, step):
        return (step - self.first_step) in self.savesteps

    def end_of_step(self, step):
        """Called after a step that ran with the Fourier capture on."""
        c = step - self.first_step
        if c not in self.savesteps:
            return
        first, last = c == self.savesteps[0], c == self.savesteps[-1]
        out = self.system.grid_ave(self.slot, self.compute.which, self.compute.quant, first, last, self.repeat,
                                   want=last)
        if last:
            self.array = out
            self.next_freq += self.freq
            self.savesteps = [self.next_freq - self.every * i for i in range(self.repeat - 1, -1, -1)]


def run_with_aves(system, nsteps, aves, first_step=0):
    """Steps with grid/ave sampling: runs stop at sample steps, whose last step
    keeps its spectra (the capture is off for the other runs)."""
    step, end = first_step, first_step + nsteps
    for a in aves:   # FixGridAve::setup at the start of every run (src/fixgrid_ave.cpp:99-125)
        a.setup(first_step)
    while step < end:
        nxt = min([end] + [a.next_sample(step) for a in aves])
        due = [a for a in aves if a.samples_at(nxt)]
        system.fourier_capture(bool(due))
        system.run(nxt - step)
        step = nxt
        for a in due:
            a.end_of_step(step)
    system.fourier_capture(False)
    return step
