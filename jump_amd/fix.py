"""EnergiesFix -- mirror of Fix.jl:104-117, 248-270.

The energies are reduced on the GPU (one deterministic reduction over all
particles); the fix records them and appends `time KE SE` lines to its file in
Julia's shortest-float text form."""
import math
import os


def julia_float(x):
    """print(io, x::Float64) as Julia writes it: the shortest digits that
    read back as x (the same digits as Python's repr); plain notation when
    -4 < pt <= 6 (pt = decimal exponent + 1), else d.ddde<exp> with an
    unpadded exponent; integers keep ".0" (1.0e6, 100000.0, 0.0001, 1.0e-5).
    The library's dump writer formats with the same rule
    (jump_amd/csrc/mpm_output.hpp: julia_float)."""
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "Inf" if x > 0 else "-Inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    x = abs(x)
    if x == 0.0:
        return sign + "0.0"
    # shortest digits and decimal exponent from repr
    r = repr(x)
    if "e" in r:
        m, e = r.split("e")
        e = int(e)
    else:
        m, e = r, 0
    if "." in m:
        ip, fp = m.split(".")
    else:
        ip, fp = m, ""
    digits = (ip + fp).lstrip("0")
    # position of the decimal point relative to the first significant digit
    lead = len(ip + fp) - len((ip + fp).lstrip("0"))
    e10 = len(ip) - 1 - lead + e
    digits = digits.rstrip("0") or "0"
    pt = e10 + 1
    if -4 < pt <= 6:
        if pt <= 0:
            return sign + "0." + "0" * (-pt) + digits
        if pt >= len(digits):
            return sign + digits + "0" * (pt - len(digits)) + ".0"
        return sign + digits[:pt] + "." + digits[pt:]
    return sign + digits[0] + "." + (digits[1:] or "0") + "e" + str(e10)


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
