"""Writes evrard_lattice_n30.json: the particle count and the first and last
100 positions (as printed, 6 significant digits) of the Evrard initial
lattice in the reference's sample output sample/evrard/results_new/00000.dat
(N = 30, sample/evrard/evrard.json).  Run from this directory with the
reference tree at /root/reference."""
import json
from pathlib import Path

SRC = Path("/root/reference/sample/evrard/results_new/00000.dat")
rows = [ln.split()[:3] for ln in SRC.read_text().splitlines() if ln and not ln.startswith("#")]
out = {"source": "sample/evrard/results_new/00000.dat", "N": 30, "count": len(rows),
       "first": rows[:100], "last": rows[-100:]}
Path(__file__).with_name("evrard_lattice_n30.json").write_text(json.dumps(out) + "\n")
print(len(rows))
