"""Writes sic8_refout.json: the energies the reference's own test
SiC_orthogonal_quick_relax (standard accuracy) prints for its first SCF
(tests/SiC_orthogonal_quick_relax/standard/SiC_orthogonal_quick_relax.refout)
and the atomic forces at that geometry (RELAXSTEP 1 of the .refgeopt),
together with the input it ran (cell, grid, atoms, settings) and the
effective input parameters it echoes; and copies the test's .inpt / .ion
input files (data) to sic8/.

Run in a container that holds the reference tree:
    python tests/golden/make_sic8_fixture.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
case = root / "tests" / "SiC_orthogonal_quick_relax" / "standard"
out = (case / "SiC_orthogonal_quick_relax.refout").read_text()
first = out.split("Self Consistent Field (SCF#2)")[0]


def grab(label):
    m = re.search(re.escape(label) + r"\s*:\s*([-+0-9.Ee]+)", first)
    return float(m.group(1))


hist = [(int(a), float(b), float(c)) for a, b, c in
        re.findall(r"^(\d+)\s+([-+0-9.E]+)\s+([-+0-9.E]+)\s+[0-9.]+\s*$", first, re.M)]
ion = (case / "SiC_orthogonal_quick_relax.ion").read_text()
blocks = re.findall(r"ATOM_TYPE:\s*(\w+).*?PSEUDO_POT:\s*(\S+).*?COORD_FRAC:[^\n]*\n((?:\s*[-0-9.]+\s+[-0-9.]+\s+[-0-9.]+\s*\n?)+)", ion, re.S)
atoms = {t: {"psp": Path(p).name, "frac": [[float(v) for v in ln.split()] for ln in c.strip().splitlines()]}
         for t, p, c in blocks}
fix = {
    "source": "tests/SiC_orthogonal_quick_relax/standard/SiC_orthogonal_quick_relax.refout (SCF#1)",
    "cell": float(re.search(r"LATVEC_SCALE:\s*([0-9.]+)", first).group(1)),
    "fd_grid": [int(v) for v in re.search(r"FD_GRID:\s*(\d+)\s+(\d+)\s+(\d+)", first).groups()],
    "nstates": int(re.search(r"NSTATES:\s*(\d+)", first).group(1)),
    "cheb_degree": int(re.search(r"CHEB_DEGREE:\s*(\d+)", first).group(1)),
    "smearing": float(re.search(r"SMEARING:\s*([0-9.Ee+-]+)", first).group(1)),
    "xc": re.search(r"EXCHANGE_CORRELATION:\s*(\S+)", first).group(1),
    "tol_scf": float(re.search(r"TOL_SCF:\s*([0-9.Ee+-]+)", first).group(1)),
    "tol_pseudocharge": float(re.search(r"TOL_PSEUDOCHARGE:\s*([0-9.Ee+-]+)", first).group(1)),
    "pseudocharge_radii": [float(v) for v in re.findall(r"Pseudocharge radii of atom type \d+\s*:\s*([0-9.]+)", first)],
    "atoms": atoms,
    "free_energy_per_atom": grab("Free energy per atom"),
    "total_free_energy": grab("Total free energy"),
    "band_structure_energy": grab("Band structure energy"),
    "exchange_correlation_energy": grab("Exchange correlation energy"),
    "self_and_correction_energy": grab("Self and correction energy"),
    "entropy_term": grab("-Entropy*kb*T"),
    "fermi_level": grab("Fermi level"),
    "scf_history": hist,
    "rms_force": grab("RMS force"),
    "max_force": grab("Maximum force"),
}
geo = (case / "SiC_orthogonal_quick_relax.refgeopt").read_text().split(":RELAXSTEP: 2")[0]
fblock = geo.split(":F(Ha/Bohr):")[1].split(":CELL:")[0]
fix["forces_step1"] = [[float(v) for v in ln.split()] for ln in fblock.strip().splitlines()]
rblock = geo.split(":R(Bohr):")[1].split(":F(Ha/Bohr):")[0]
fix["positions_step1"] = [[float(v) for v in ln.split()] for ln in rblock.strip().splitlines()]
# effective parameters echoed under "Input parameters" (the defaults it applied)
head = first.split("Input parameters")[1].split("Cell")[0]
fix["echo"] = {k: v.strip() for k, v in re.findall(r"^([A-Z_0-9]+):[ \t]*(.*)$", head, re.M)
               if k not in ("LATVEC",)}
here = Path(__file__).resolve().parent
(here / "sic8").mkdir(exist_ok=True)
for ext in (".inpt", ".ion"):
    src = case / ("SiC_orthogonal_quick_relax" + ext)
    (here / "sic8" / src.name).write_text(src.read_text())
dst = here / "sic8_refout.json"
dst.write_text(json.dumps(fix, indent=1) + "\n")
print(f"wrote {dst}")
