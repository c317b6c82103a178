"""The reference example's command line (examples/advection.cpp:43-519) on
this backend:

    python -m cajitafluids_amd.cli [-n cells] [-s size] [-t time] [-i deltat]
        [-d density] [-g gravity] [-p driver] [-m matrix-solver]
        [-c preconditioner] [-x input-x] [-y input-y] [-w input-width]
        [-h input-height] [-q input-quantity] [-u input-velocity-x]
        [-v input-velocity-y] [--help]

Flags, long names, defaults, validation and the printed banner follow
parseInput / main as written, including what they do rather than what the
help text says: -w sets the inflow width and -h the inflow height (the
write frequency stays 20; --help prints the help), -i's check tests the
final time.  -p accepts the reference's drivers (serial, cuda, openmp,
pthreads) and hip; every one runs the HIP kernels.  The run is advect()
(advection.cpp:436-461): SOLID walls, the inflow box, body force
(0, -gravity), zero initial state, createSolver with the solver and
preconditioner names, then Solver::solve(t_final, write_freq) with output
under ./data.  --write-freq and --output-dir (long only) are additions.
"""
import getopt
import re
import sys

DRIVERS = ("serial", "cuda", "openmp", "pthreads", "hip")

SHORT = "n:s:t:i:d:g:p:m:c:x:y:w:h:q:u:v:"
LONG = ["cells=", "size=", "time=", "deltat=", "density=", "gravity=", "driver=",
        "matrix-solver=", "preconditioner=", "input-x=", "input-y=", "input-width=",
        "input-height=", "input-quantity=", "input-velocity-x=", "input-velocity-y=",
        "help", "write-freq=", "output-dir="]
LONG_TO_SHORT = {"cells": "n", "size": "s", "time": "t", "deltat": "i", "density": "d",
                 "gravity": "g", "driver": "p", "matrix-solver": "m", "preconditioner": "c",
                 "input-x": "x", "input-y": "y", "input-width": "w", "input-height": "h",
                 "input-quantity": "q", "input-velocity-x": "u", "input-velocity-y": "v",
                 "help": "j"}


class Exit(Exception):
    """parseInput's exit(code)."""

    def __init__(self, code):
        super().__init__(code)
        self.code = code


def defaults():
    """parseInput's defaults (advection.cpp:172-189)."""
    return dict(device="serial", t_final=4.0, delta_t=0.005, write_freq=20,
                num_cells=[128, 128], bbox=[0.0, 0.0, 1.0, 1.0], in_location=[0.2, 0.45],
                in_size=[0.02, 0.1], in_velocity=[1.0, 0.0], in_quantity=3.0, density=0.1,
                gravity=0.0, solver="PCG", precon="PFMG", output_dir="data")


def help_text(progname):
    """help() (advection.cpp:104-148), entries as printed there."""
    rows = [("-p", "On-node Parallelism Model (default serial)"),
            ("-s", "Size of domain (default 1.0)"),
            ("-n", "Number of Cells (default 128)"),
            ("-t", "Amount of time to simulate (default 4.0)"),
            ("-t", "Timestep increment (default 0.005)"),
            ("-w", "Write Frequency (default 20)"),
            ("-d", "Fluid Density (default 0.1)"),
            ("-x", "Inflow X Location (default 0.2)"),
            ("-y", "Inflow Y Location (default 0.45)"),
            ("-w", "Inflow Width (default 0.001)"),
            ("-h", "Inflow Height (default 0.1)"),
            ("-u", "Inflow X Velocity (default 0)"),
            ("-v", "Inflow Y Velocity (default 1.0)"),
            ("-q", "Inflow Quantity (default 3.0)"),
            ("-g", "Gravity (default 0.0)"),
            ("-h", "Print Help Message")]
    return "Usage: %s\n" % progname + "".join("%-10s%-40s\n" % r for r in rows)


def _atof(s):
    """C atof: the longest leading decimal prefix, 0.0 if there is none."""
    m = re.match(r"\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?|inf|nan)", s, re.I)
    return float(m.group(0)) if m else 0.0


def _atoi(s):
    m = re.match(r"\s*[+-]?\d+", s)
    return int(m.group(0)) if m else 0


def parse_input(argv, progname="advection", err=sys.stderr, out=sys.stdout):
    """parseInput (advection.cpp:158-346).  Raises Exit(code) where the
    reference calls exit(code)."""
    cl = defaults()

    def fail(msg):
        err.write(msg + "\n")
        out.write(help_text(progname))
        raise Exit(-1)

    try:
        opts, _ = getopt.getopt(argv, SHORT, LONG)
    except getopt.GetoptError:
        err.write("Invalid argument.\n")
        out.write(help_text(progname))
        raise Exit(-1)
    for flag, arg in opts:
        key = flag[2:] if flag.startswith("--") else flag[1:]
        if key == "write-freq":
            cl["write_freq"] = _atoi(arg)
            continue
        if key == "output-dir":
            cl["output_dir"] = arg
            continue
        ch = LONG_TO_SHORT.get(key, key)
        if ch == "n":
            cl["num_cells"][0] = _atoi(arg)
            if cl["num_cells"][0] <= 0:
                fail("Invalid cell number argument.")
            cl["num_cells"][1] = cl["num_cells"][0]
        elif ch == "s":
            cl["bbox"][2] = _atof(arg)
            if cl["bbox"][3] <= 0.0:
                fail("Invalid bounding box size argument.")
            cl["bbox"][0] = 0.0
            cl["bbox"][1] = 0.0
            cl["bbox"][3] = cl["bbox"][2]
        elif ch == "t":
            cl["t_final"] = _atof(arg)
            if cl["t_final"] <= 0.0:
                fail("Invalid final time argument.")
        elif ch == "i":
            cl["delta_t"] = _atof(arg)
            if cl["t_final"] <= 0.0:
                fail("Invalid timestep increment argument.")
        elif ch == "d":
            cl["density"] = _atof(arg)
            if cl["density"] <= 0.0:
                fail("Invalid fluid density argument.")
        elif ch == "p":
            if arg not in DRIVERS:
                fail("Invalid  parallel device argument.")
            cl["device"] = arg
        elif ch == "g":
            cl["gravity"] = _atof(arg)
        elif ch == "j":
            out.write(help_text(progname))
            raise Exit(0)
        elif ch == "m":
            cl["solver"] = arg
        elif ch == "c":
            cl["precon"] = arg
        elif ch == "x":
            cl["in_location"][0] = _atof(arg)
            if cl["in_location"][0] < 0.0:
                fail("Invalid inflow x argument.")
        elif ch == "y":
            cl["in_location"][1] = _atof(arg)
            if cl["in_location"][1] < 0.0:
                fail("Invalid inflow y argument.")
        elif ch == "w":
            cl["in_size"][0] = _atof(arg)
            if cl["in_size"][0] <= 0.0:
                fail("Invalid inflow width argument.")
        elif ch == "h":
            cl["in_size"][1] = _atof(arg)
            if cl["in_size"][1] <= 0.0:
                fail("Invalid inflow height argument.")
        elif ch == "q":
            cl["in_quantity"] = _atof(arg)
            if cl["in_quantity"] < 0.0:
                fail("Invalid inflow quantity argument.")
        elif ch == "u":
            cl["in_velocity"][0] = _atof(arg)
        elif ch == "v":
            cl["in_velocity"][1] = _atof(arg)
    return cl


def banner(cl):
    """main's echo of the options (advection.cpp:480-512)."""
    def w(x, n=8):
        return ("%-" + str(n) + "s") % (("%g" % x) if isinstance(x, float) else x)
    lines = ["CajitaFluids", "=======Command line arguments======="]
    lines.append("%-20s: %s" % ("Thread Setting", w(cl["device"])))
    lines.append("%-20s: %s%s" % ("Cells", w(cl["num_cells"][0]), w(cl["num_cells"][1])))
    lines.append("%-20s: %s%s" % ("Domain", w(cl["bbox"][2]), w(cl["bbox"][3])))
    lines.append("%-20s: %s at Location (%s%s ) Size (%s%s ) Velocity (%s%s )" % (
        "Input Flow", w(cl["in_quantity"]), w(cl["in_location"][0]), w(cl["in_location"][1]),
        w(cl["in_size"][0]), w(cl["in_size"][1]), w(cl["in_velocity"][0]),
        w(cl["in_velocity"][1])))
    lines.append("%-20s: %s" % ("Total Simulation Time", w(cl["t_final"])))
    lines.append("%-20s: %s" % ("Timestep Size", w(cl["delta_t"])))
    lines.append("%-20s: %s" % ("Write Frequency", w(cl["write_freq"])))
    lines.append("====================================")
    return "\n".join(lines) + "\n"


def config_of(cl):
    """advect()'s createSolver arguments (advection.cpp:436-458): SOLID
    walls, the inflow box, body force (0, -gravity), as a configuration."""
    from .configs import SOLID
    from .solver import BodyForce, BoundaryCondition, InflowSource, solver_config
    bc = BoundaryCondition([SOLID] * 4)
    source = InflowSource(list(cl["in_location"]), list(cl["in_size"]),
                          list(cl["in_velocity"]), cl["in_quantity"])
    body = BodyForce([0.0, -cl["gravity"]])
    return solver_config("hip", list(cl["bbox"]), list(cl["num_cells"]), cl["density"], bc,
                         source, body, cl["delta_t"], cl["solver"], cl["precon"])


def main(argv=None, progname="advection"):
    argv = sys.argv[1:] if argv is None else argv
    try:
        cl = parse_input(argv, progname)
    except Exit as e:
        return e.code & 0xFF
    sys.stdout.write(banner(cl))
    sys.stdout.flush()
    try:
        cfg = config_of(cl)
    except RuntimeError as e:
        sys.stderr.write("%s\n" % e)
        return 1
    from .solver import Solver
    with Solver(cfg) as solver:
        solver.solve(cl["t_final"], cl["write_freq"], output_dir=cl["output_dir"],
                     log=lambda s: print(s, flush=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
