"""Writes tests/golden/ff_tables.json from the reference's force-field source
(src/amber.rs, src/rotamer.rs, read as text): the AMBER_NB table
(src/amber.rs:46-63), the H-bond donor / acceptor classes (src/amber.rs:72-84),
the EEF1 table (src/amber.rs:414-431), per residue in RESIDUE_TOPOLOGY order
(src/amber.rs:357-378) each heavy atom's (AtomType index, charge, hydrophobic
flag) and name (src/amber.rs:126-351), and per residue in AminoAcid order the
rotamer chi angles (src/rotamer.rs:33-184) and the side-chain NERF placements
(ref_a, ref_b, ref_c, bond length, angle, fixed dihedral, chi index or 255;
src/rotamer.rs:280-440).

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

rsrc = (ref / "src" / "rotamer.rs").read_text()
num = r"\s*([-\d.]+)\s*"
rot = {}
for m in re.finditer(r"static ROT_(\w+): &\[Rotamer\] = &\[(.*?)\];", rsrc, re.S):
    rot[m.group(1)] = [[float(v) for v in r[:4]] for r in re.findall(
        r"Rotamer::new\(" + ",".join([num] * 5) + r"\)", m.group(2))]
lib = re.search(r"pub static ROTAMER_LIB.*?=\s*\[(.*?)\];", rsrc, re.S).group(1)
rot_order = re.findall(r"ROT_(\w+)", lib)
bd = {}
for m in re.finditer(r"static BD_(\w+): &\[BondDef\] = &\[(.*?)\];", rsrc, re.S):
    body = re.sub(r"//[^\n]*", "", m.group(2))
    defs = []
    for kind, args in re.findall(r"BondDef::(chi|fixed)\(([^)]*)\)", body):
        a = [float(v) for v in args.split(",")]
        if kind == "chi":
            defs.append([a[0], a[1], a[2], a[3], a[4], 0.0, a[5]])
        else:
            defs.append([a[0], a[1], a[2], a[3], a[4], a[5], 255.0])
    bd[m.group(1)] = defs
sc = re.search(r"pub static SIDECHAIN_DEFS.*?=\s*\[(.*?)\];", rsrc, re.S).group(1)
bd_order = re.findall(r"BD_(\w+)", sc)

topo = re.search(r"pub static RESIDUE_TOPOLOGY.*?=\s*\[(.*?)\];", src, re.S).group(1)
order = re.findall(r"atoms:\s*(\w+)_ATOMS", topo)
out = {"amber_nb": amber_nb, "residues": [residues[name] for name in order],
       "residue_order": order, "atom_names": [names[name] for name in order],
       "hbond_donor": bools("HBOND_DONOR"), "hbond_acceptor": bools("HBOND_ACCEPTOR"),
       "eef1": eef1, "rotamers": [rot[n] for n in rot_order],
       "sidechain": [bd[n] for n in bd_order]}
dst = Path(__file__).resolve().parent / "ff_tables.json"
dst.write_text(json.dumps(out, indent=1) + "\n")
print(f"wrote {dst}: {len(amber_nb)} atom types, {len(order)} residues, "
      f"{sum(len(r) for r in out['residues'])} atoms")
