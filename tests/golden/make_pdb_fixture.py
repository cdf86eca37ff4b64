"""Writes mini.pdb and mini_expected.json: a small hand-built PDB file that exercises the
reader's rules (ATOM records only, alternate locations, element column, radii, the END
record), and the atoms a reader must return from it.

The atoms are synthetic (seeded), laid out as a few residues of a peptide-like chain A, a
second chain B, HETATM waters, and a second model after ENDMDL that must not be read.
Run: python tests/golden/make_pdb_fixture.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RADIUS = {"O": 1.52, "C": 1.70, "N": 1.55, "H": 1.20, "S": 1.80, "P": 1.80}


def atom_line(rec, serial, name, alt, res, chain, resseq, x, y, z, element, occ=1.0, tf=20.0):
    # PDB v3.3 fixed columns; element right-justified in columns 77-78
    return "%-6s%5d %-4s%1s%3s %1s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          %2s  " % (
        rec, serial, name, alt, res, chain, resseq, x, y, z, occ, tf, element)


def main():
    rng = np.random.default_rng(20240611)
    lines = ["HEADER    SYNTHETIC TEST MOLECULE", "REMARK   1 BUILT BY make_pdb_fixture.py"]
    kept = []  # (x, y, z, radius) in file order
    serial = 1
    names = [("N", "N"), ("CA", "C"), ("C", "C"), ("O", "O"), ("CB", "C"), ("SG", "S"), ("H", "H")]
    for chain, nres in (("A", 4), ("B", 2)):
        for resseq in range(1, nres + 1):
            base = rng.uniform(-8, 8, size=3)
            for k, (name, el) in enumerate(names):
                p = base + rng.uniform(-1.5, 1.5, size=3)
                x, y, z = (round(float(v), 3) for v in p)
                alt = " "
                if resseq == 2 and name == "CB":
                    # two conformations: A kept, B dropped
                    lines.append(atom_line("ATOM", serial, name, "A", "CYS", chain, resseq, x, y, z, el, 0.5))
                    kept.append((x, y, z, RADIUS[el]))
                    serial += 1
                    lines.append(atom_line("ATOM", serial, name, "B", "CYS", chain, resseq, x + 0.4, y, z, el, 0.5))
                    serial += 1
                    continue
                if resseq == 3 and name == "O":
                    # conformation C only: dropped
                    lines.append(atom_line("ATOM", serial, name, "C", "CYS", chain, resseq, x, y, z, el))
                    serial += 1
                    continue
                if resseq == 1 and name == "SG" and chain == "A":
                    # no element column at all (short line): radius 1.4
                    lines.append(atom_line("ATOM", serial, name, alt, "CYS", chain, resseq, x, y, z, el)[:66])
                    kept.append((x, y, z, 1.40))
                    serial += 1
                    continue
                if resseq == 4 and name == "SG":
                    # the radius is looked up by the first letter only: selenium "SE" gets the S radius
                    lines.append(atom_line("ATOM", serial, "SE", alt, "MSE", chain, resseq, x, y, z, "SE"))
                    kept.append((x, y, z, 1.80))
                    serial += 1
                    continue
                if resseq == 2 and name == "SG" and chain == "B":
                    # a first letter outside the table (iron): radius 1.4
                    lines.append(atom_line("ATOM", serial, "FE", alt, "HEM", chain, resseq, x, y, z, "FE"))
                    kept.append((x, y, z, 1.40))
                    serial += 1
                    continue
                lines.append(atom_line("ATOM", serial, name, alt, "CYS", chain, resseq, x, y, z, el))
                kept.append((x, y, z, RADIUS[el]))
                serial += 1
        lines.append("TER   %5d      CYS %1s%4d" % (serial, chain, nres))
        serial += 1
    for w in range(3):
        p = rng.uniform(-12, 12, size=3)
        lines.append(atom_line("HETATM", serial, "O", " ", "HOH", "W", w + 1, *(round(float(v), 3) for v in p), "O"))
        serial += 1
    # a phosphorus atom after the waters (still read: TER and HETATM do not stop the reader)
    p = [round(float(v), 3) for v in rng.uniform(-8, 8, size=3)]
    lines.append(atom_line("ATOM", serial, "P", " ", "DA", "C", 1, *p, "P"))
    kept.append((p[0], p[1], p[2], 1.80))
    serial += 1
    lines.append("ENDMDL")
    lines.append("MODEL        2")
    lines.append(atom_line("ATOM", serial, "CA", " ", "GLY", "A", 1, 1.0, 2.0, 3.0, "C"))
    lines.append("END")
    with open(os.path.join(HERE, "mini.pdb"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(HERE, "mini_expected.json"), "w") as f:
        json.dump({"atoms": kept}, f, indent=0)


if __name__ == "__main__":
    main()
