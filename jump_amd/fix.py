"""EnergiesFix -- mirror of Fix.jl:104-117, 248-270.

The energies are reduced on the GPU (one deterministic reduction over all
particles); the fix records them and appends `time KE SE` lines to its file in
Julia's shortest-float text form."""
import os


def julia_float(x):
    """Shortest round-trip digits with Julia's exponent spelling (1.0e-5
    rather than 1e-05); the digits are exact, the switch point between plain
    and exponent notation follows Python's repr."""
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "Inf" if x > 0 else "-Inf"
    r = repr(x)
    if "e" in r:
        mant, exp = r.split("e")
        if "." not in mant:
            mant += ".0"
        return f"{mant}e{int(exp)}"
    return r


class FixBase:
    pass


class EnergiesFix(FixBase):
    def __init__(self, solids, filename):
        self.solids = solids
        self.filename = filename
        if filename:
            d = os.path.dirname(filename)
            if d:
                os.makedirs(d, exist_ok=True)
            if os.path.isfile(filename):
                os.remove(filename)
            self.file = open(filename, "a")
        else:
            self.file = None
        self.kinEnergy = []
        self.strEnergy = []
        self.recordTime = []
        self._source = None  # set by the solver: callable returning (KE, SE)

    def record(self, time, KE, SE):
        self.kinEnergy.append(KE)
        self.strEnergy.append(SE)
        self.recordTime.append(time)
        if self.file:
            self.file.write(f"{julia_float(time)} {julia_float(KE)} {julia_float(SE)}\n")


def compute(fix, time):
    """compute(fix, t): Fix.jl:248-270 (energies from the device state)."""
    if isinstance(fix, (list, tuple)):
        for f in fix:
            compute(f, time)
        return
    if isinstance(fix, EnergiesFix):
        KE, SE = fix._source()
        fix.record(time, KE, SE)


def closeFile(fix):
    if isinstance(fix, (list, tuple)):
        for f in fix:
            closeFile(f)
        return
    if getattr(fix, "file", None):
        fix.file.close()
        fix.file = None
