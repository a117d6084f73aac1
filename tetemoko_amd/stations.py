"""Fault-station output in the reference's file format.

AMRLevelLinElast opens one file per fault station and level, named
"<prefix>Fx<x>z<z>.L<level>.dat" with x and z printed as %5.4f (AMRLevelLinElast.cpp:283-287),
writes a three-line text header (number of values, size of a real, the column names;
:298-308) and then appends, after every level step, the raw doubles
time, value[0..numBdryVars-1], level, procID (:1781-1810).  The values come from
capi.Stations, which records them on the device.
"""
import math

import numpy as np

BDRY_NAMES = {
    # RSIBC::bdryNames (RSIBC.cpp:380-396), SWIBC::bdryNames (SWIBC.H:132-140)
    "rs": ["Vx", "Vz", "V", "sxy", "syz", "syy", "slip x", "slip z", "slip total", "rupture time", "psi"],
    "sw": ["Vx", "Vz", "sxy", "syz", "V", "slip_x", "slip_z", "slip_total", "rupture"],
}


def placement(x, dx):
    """Face index and weight of a station at x metres from the domain's lo x edge
    (AMRLevelLinElast.cpp:262-270): i = floor(x/dx - 0.5), d = x - (i + 0.5) dx."""
    i = int(math.floor(x / dx - 0.5))
    return i, -(i + 0.5) * dx + x


def station_file_name(prefix, x, z, level):
    return "%sFx%5.4fz%5.4f.L%d.dat" % (prefix, x, z, level)


class StationFiles:
    """The reference's station files for one level: headers at creation, records appended."""

    def __init__(self, prefix, x, fault, level=0, z=None, proc_id=0):
        self.names = BDRY_NAMES[fault]
        z = [0.0] * len(x) if z is None else z
        self.paths = [station_file_name(prefix, xx, zz, level) for xx, zz in zip(x, z)]
        self.level, self.proc_id = float(level), float(proc_id)
        for p in self.paths:
            with open(p, "w") as f:
                f.write("#number of values = %d\n" % len(self.names))
                f.write("#size of real = %d\n" % 8)
                f.write("#time" + "".join("\t%s" % n for n in self.names) + "\tlevel\tprocID\n")

    def append(self, records):
        """records: [nrecords, nstations, 1 + numBdryVars] from capi.Stations.read()."""
        records = np.asarray(records, dtype=np.float64)
        for k, p in enumerate(self.paths):
            rows = np.empty((records.shape[0], records.shape[2] + 2))
            rows[:, :records.shape[2]] = records[:, k, :]
            rows[:, -2] = self.level
            rows[:, -1] = self.proc_id
            with open(p, "ab") as f:
                f.write(rows.astype("<f8").tobytes())


def read_station_file(path):
    """(header lines, records [n, 1 + numBdryVars + 2]) of a station file."""
    with open(path, "rb") as f:
        data = f.read()
    lines = []
    pos = 0
    for _ in range(3):
        end = data.index(b"\n", pos)
        lines.append(data[pos:end].decode())
        pos = end + 1
    nv = int(lines[0].split("=")[1])
    rec = np.frombuffer(data[pos:], dtype="<f8").reshape(-1, nv + 3)
    return lines, rec
