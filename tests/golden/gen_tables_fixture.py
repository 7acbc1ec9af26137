"""Writes tests/golden/ff_tables.json from the reference's force-field source
(src/amber.rs, read as text): the AMBER_NB table (src/amber.rs:46-63), the
H-bond donor / acceptor classes (src/amber.rs:72-84), the EEF1 table
(src/amber.rs:414-431) and, per residue in RESIDUE_TOPOLOGY order
(src/amber.rs:357-378), each heavy atom's (AtomType index, charge,
hydrophobic flag) and name (src/amber.rs:126-351).

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

def bools(name):
    block = re.search(r"pub static " + name + r".*?=\s*\[(.*?)\];", src, re.S).group(1)
    block = re.sub(r"//[^\n]*", "", block)
    return [1 if v == "true" else 0 for v in re.findall(r"\b(true|false)\b", block)]

eef1_block = re.search(r"pub static EEF1.*?=\s*\[(.*?)\];", src, re.S).group(1)
eef1 = [[float(a), float(b), float(c)] for a, b, c in re.findall(
    r"dg_ref:\s*([-\d.]+),\s*vol:\s*([-\d.]+),\s*lambda:\s*([-\d.]+)", eef1_block)]

residues, names = {}, {}
for m in re.finditer(r"static (\w+)_ATOMS: &\[ResAtom\] = &\[(.*?)\];", src, re.S):
    atoms, nm = [], []
    for a in re.finditer(r'atom!\("([^"]*)",\s*(\w+),\s*([-\d.]+)(,\s*H)?\)', m.group(2)):
        atoms.append([types[a.group(2)], float(a.group(3)), 1 if a.group(4) else 0])
        nm.append(a.group(1).strip())
    residues[m.group(1)] = atoms
    names[m.group(1)] = nm

topo = re.search(r"pub static RESIDUE_TOPOLOGY.*?=\s*\[(.*?)\];", src, re.S).group(1)
order = re.findall(r"atoms:\s*(\w+)_ATOMS", topo)
out = {"amber_nb": amber_nb, "residues": [residues[name] for name in order],
       "residue_order": order, "atom_names": [names[name] for name in order],
       "hbond_donor": bools("HBOND_DONOR"), "hbond_acceptor": bools("HBOND_ACCEPTOR"),
       "eef1": eef1}
dst = Path(__file__).resolve().parent / "ff_tables.json"
dst.write_text(json.dumps(out, indent=1) + "\n")
print(f"wrote {dst}: {len(amber_nb)} atom types, {len(order)} residues, "
      f"{sum(len(r) for r in out['residues'])} atoms")
