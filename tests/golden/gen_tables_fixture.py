"""Writes tests/golden/ff_tables.json from the reference's force-field source
(src/amber.rs, read as text): the AMBER_NB table (src/amber.rs:46-63) and,
per residue in RESIDUE_TOPOLOGY order (src/amber.rs:357-378), each heavy
atom's (AtomType index, charge, hydrophobic flag) (src/amber.rs:126-351).

Run in the build container (where the reference is mounted):
    python tests/golden/gen_tables_fixture.py /root/reference
The JSON is data (numbers only) and is what the CPU tests compare against.
"""
import json
import re
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
src = (ref / "src" / "amber.rs").read_text()

enum = re.search(r"pub enum AtomType \{(.*?)\}", src, re.S).group(1)
types = {m.group(1): int(m.group(2)) for m in re.finditer(r"(\w+)\s*=\s*(\d+)", enum)}

nb_block = re.search(r"pub static AMBER_NB.*?=\s*\[(.*?)\];", src, re.S).group(1)
amber_nb = [[float(a), float(b)] for a, b in
            re.findall(r"r_min_half:\s*([-\d.]+),\s*epsilon:\s*([-\d.]+)", nb_block)]

residues = {}
for m in re.finditer(r"static (\w+)_ATOMS: &\[ResAtom\] = &\[(.*?)\];", src, re.S):
    atoms = []
    for a in re.finditer(r'atom!\("[^"]*",\s*(\w+),\s*([-\d.]+)(,\s*H)?\)', m.group(2)):
        atoms.append([types[a.group(1)], float(a.group(2)), 1 if a.group(3) else 0])
    residues[m.group(1)] = atoms

topo = re.search(r"pub static RESIDUE_TOPOLOGY.*?=\s*\[(.*?)\];", src, re.S).group(1)
order = re.findall(r"atoms:\s*(\w+)_ATOMS", topo)
out = {"amber_nb": amber_nb, "residues": [residues[name] for name in order],
       "residue_order": order}
dst = Path(__file__).resolve().parent / "ff_tables.json"
dst.write_text(json.dumps(out, indent=1) + "\n")
print(f"wrote {dst}: {len(amber_nb)} atom types, {len(order)} residues, "
      f"{sum(len(r) for r in out['residues'])} atoms")
