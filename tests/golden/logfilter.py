"""The lines of an omega_<name>.txt log that a run determines.

Kept: the settings block, the atom-type table and totals, the boundary fit
(axes, RMSD, enlargement), the cluster summary, the velocity window, every
progress / completed / Omega line, and the final statistics.  Dropped: dates,
wall times, absolute paths, the per-integral thread line, and the setup
diagnostics the engine does not print (inertia tensor, multipoles, fit
uncertainties).  Used both to cut the reference's example logs into
fixtures and to read the command-line program's own logs.
"""
import re

PREFIXES = (
    "\t Project name:", "\t Number of CCS integrals:", "\t Number of trajectories per CCS integral:",
    "\t Trajectory time step", "\t Temperature of the molecule", "\t Random number seed",
    "\t Total charge of molecule (e):",
    "\t Total number of atoms:", "\t Total mass of molecule:", "\t Reduced mass:",
    " Axis Length", "\t RMSD:", " Ellipsoid axes further expanded",
    "\t by center of charge:", "\t Number of clusters formed:", "\t Average cluster size:",
    "\t Average cluster radius:", "\t Multipole cut-off radius:", "\t Dispersion cut-off radius:",
    "\t Maximum velocity:", "\t Minimum velocity:", "\t Mean velocity:", "\t Mode velocity:",
    "\t distribution used:",
    " progress:", " Number of completed trajectories:", " Omega:",
    "   Mean Cross-section:", "   Standard Dev:", "   StdDev Percent:", "   Weighted CCS Mean:",
    "   Weighted Error:", "   Weighted Pct. Error:", "   Max Energy Deviation:",
)
# force-model sentence of the settings block, and rows of the atom-type table
PATTERNS = (
    re.compile(r"^\t .* will be used\.$"),
    re.compile(r"^\t \S+\s+\S{1,2}\s+[\d.]+\s+[\d.]+\s+[\d.]+\s+\d+\s+\d+$"),
)


def log_lines(text: str):
    out = []
    for line in text.splitlines():
        if line.startswith(PREFIXES) or any(p.match(line) for p in PATTERNS):
            out.append(line.rstrip())
    return out
