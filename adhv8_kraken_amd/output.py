"""Solution output and hot-start files in the reference's SMS DATASET format.

The time-step writer follows ssw_2d_print_ts (src/structs/ssw_2d.c:262-300)
for the files sio_init names (src/structs/sio.c:105-115): <stem>_dep.dat,
_old_dep.dat, _vel.dat and _old_vel.dat, each a DATASET with the header of
pdata / tc_timeunits (src/initio/print_header.c:9-23,
src/tools/tc_timeunits.c:15), a "TS 0 <time>" card per output time, node
values as "%16.8e" (vectors "%16.8e %16.8e %16.8e", the third component 0:
src/structs/svect2d.c:51) and ENDDS (src/initio/print_trailer.c:9).

The hot-start writer produces what the reference's hot-start reader takes
(src/structs/ssw.c:164-545): IOH (depth), IV (velocity), IPH and IPV
(previous depth and velocity).  With precise=True values are written with
17 significant digits, so a run resumed from the file continues bit for bit
(checkpoint / resume); the reference's 9-digit format is the default.
"""
from __future__ import annotations

import numpy as np

SCALAR_FMT = "%16.8e"
VECTOR_FMT = "%16.8e %16.8e %16.8e"
PRECISE_SCALAR_FMT = "%25.17e"
PRECISE_VECTOR_FMT = "%25.17e %25.17e %25.17e"


def _header(f, title, begtype, nnodes, nelems):
    f.write("DATASET\n")
    f.write('OBJTYPE "mesh2d"\n')
    f.write(f"{begtype}\n")
    f.write(f"ND {nnodes}\n")
    f.write(f"NC {nelems}\n")
    f.write(f'NAME "{title}"\n')
    f.write("TIMEUNITS SECONDS\n")


def _values(f, values, vector, precise):
    if vector:
        v = np.asarray(values, dtype=np.float64).reshape(-1, 2)
        fmt = PRECISE_VECTOR_FMT if precise else VECTOR_FMT
        f.write("".join(fmt % (a, b, 0.0) + "\n" for a, b in v))
    else:
        fmt = PRECISE_SCALAR_FMT if precise else SCALAR_FMT
        f.write("".join(fmt % a + "\n" for a in np.asarray(values, dtype=np.float64)))


class SolutionWriter:
    """Time series of depth and depth-averaged velocity (current and
    previous time level), one DATASET file each, as the reference writes
    them at its output times."""

    FILES = (("_dep.dat", "Depth", False), ("_old_dep.dat", "Depth", False),
             ("_vel.dat", "Depth-Averaged Velocity", True), ("_old_vel.dat", "Depth-Averaged Velocity", True))

    def __init__(self, stem, nnodes, nelems, precise=False):
        self.precise = precise
        self.files = []
        for suffix, title, vec in self.FILES:
            f = open(stem + suffix, "w")
            _header(f, title, "BEGVEC" if vec else "BEGSCL", nnodes, nelems)
            self.files.append((f, vec))

    def write(self, time, head, vel, old_head, old_vel):
        for (f, vec), data in zip(self.files, (head, old_head, vel, old_vel)):
            f.write("TS 0 %15.8e\n" % time)
            _values(f, data, vec, self.precise)

    def close(self):
        for f, _ in self.files:
            f.write("ENDDS\n")
            f.close()
        self.files = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_hotstart(path, nnodes, nelems, time, head, vel, old_head, old_vel, precise=False):
    """A hot-start file the reference (and adhv8_kraken_amd.deck.read_hot)
    reads: IOH, IV, IPH, IPV at `time`."""
    with open(path, "w") as f:
        for name, data, vec in (("IOH", head, False), ("IV", vel, True), ("IPH", old_head, False),
                                ("IPV", old_vel, True)):
            _header(f, name, "BEGVEC" if vec else "BEGSCL", nnodes, nelems)
            f.write("TS 0 %15.8e\n" % time)
            _values(f, data, vec, precise)
            f.write("ENDDS\n")


def read_dataset(path):
    """Reads a DATASET file back: (times, [values per time]) with vectors as
    [N, 2] (the third component dropped)."""
    times, sets = [], []
    cur, vec, nd = None, False, None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "BEGVEC":
                vec = True
            elif t[0] == "BEGSCL":
                vec = False
            elif t[0] == "ND":
                nd = int(t[1])
            elif t[0] == "TS":
                if cur is not None:
                    sets.append(np.array(cur))
                times.append(float(t[2]))
                cur = []
            elif t[0] == "ENDDS":
                if cur is not None:
                    sets.append(np.array(cur))
                cur = None
            elif cur is not None:
                vals = [float(v) for v in t]
                cur.append(vals[:2] if vec else vals[0])
    for s in sets:
        if len(s) != nd:
            raise ValueError(f"{path}: a time step holds {len(s)} values for {nd} nodes")
    return np.array(times), sets
