"""Writes air5.json: the 5-species air model (6 reactions) that ucs.x builds
from chemModels/5speciesAir.rxn and the thermo database its
chemdata/chemDB.py derives from BURCAT_FIXED.THR.  The species are numbered
as ChemModel::BuildGlobalSpeciesList (ucs/chem.tcc:163-204) numbers them:
in order of first appearance over the reactions' local species lists, which
gives [O2, O, N, N2, NO] (the file's speciesInModel line is not used for
this).  Every per-species array of the FR state follows that numbering.

The Burcat records are located the way the database builder does
(chemdata/chemDB.py:514-613): skip the 88-line header, take the text
between a CAS-number line and the next blank line, keep the last four
80-column lines whose column 80 holds the line number, and let the first
record of a compound win (chemDB.py:720-737 never overwrites); coefficients
are read in 15-character fields with the builder's repair of a blank
exponent sign (chemDB.py:697-717).  Low-range coefficients (T <= 1000 K)
come from lines 3-4, high-range from lines 2-3; MW is columns 70-78.
Species.GetDBInfo (ucs/species.tcc:140-320) reads "NASA7_burcat1/2" first,
so these are the coefficients the solver uses.

Run once with the reference tree at hand:
    python proteuscfd_amd/chem/make_air5.py /root/reference
"""
import json
import os
import re
import sys



def read_coeff(line, pos, nchar=15):
    c = line[pos:pos + nchar]
    if nchar == 15:
        if c[11:13] == "E ":
            c = c[0:12] + "+" + c[13:15]
        if c[10:12] == " E":
            c = c[0:10] + "0" + c[11:15]
    if "N/A" in c:
        return 0.0
    return float(c)


def burcat(path):
    with open(path) as f:
        text = "".join(f.readlines()[88:])
    pat = re.compile(r"[0-9]+\-[0-9]*\-[0-9].+?\n[ \t\n]*?\n", re.DOTALL)
    out = {}
    for r in pat.findall(text):
        rec = r.strip().split("\n")[-4:]
        if len(rec) != 4:
            continue
        name = rec[0].split()[0]
        rec = [x.expandtabs().rstrip() for x in rec]
        if any(len(x) != 80 or x[79] != str(i + 1) for i, x in enumerate(rec)):
            continue
        if name in out:
            continue
        hi = [read_coeff(rec[1], 0), read_coeff(rec[1], 15), read_coeff(rec[1], 30), read_coeff(rec[1], 45),
              read_coeff(rec[1], 60), read_coeff(rec[2], 0), read_coeff(rec[2], 15)]
        lo = [read_coeff(rec[2], 30), read_coeff(rec[2], 45), read_coeff(rec[2], 60), read_coeff(rec[3], 0),
              read_coeff(rec[3], 15), read_coeff(rec[3], 30), read_coeff(rec[3], 45)]
        out[name] = {"MW": float(rec[0][69:78]), "lo": lo, "hi": hi, "coeff15": read_coeff(rec[3], 60),
                     "T1": float(rec[0][47:55]), "T3": float(rec[0][56:65])}
    return out


# 5speciesAir.rxn, restated per reaction in the order the reaction parser
# builds its local species list (ucs/reaction.tcc:313-570): equation species
# left to right, merged by first occurrence, then third bodies named only in
# the M line appended with nu' = nu'' = 0.  Entries: (symbol, nu', nu'',
# third-body efficiency; 1 where the M line is silent).
# (type 2 = GuptaModArrhenius k = A T^n exp(-EA/T); third bodies 0/1)
REACTIONS = [
    (2, 1, 3.618e12, 5.94e4, -1.0,
     [("O2", 1, 0, 9.0), ("O", 0, 2, 25.0), ("N", 0, 0, 1.0), ("N2", 0, 0, 2.0), ("NO", 0, 0, 1.0)]),
    (2, 1, 1.92e11, 1.131e5, -0.5,
     [("N2", 1, 0, 2.5), ("N", 0, 2, 1.0), ("O2", 0, 0, 1.0), ("NO", 0, 0, 1.0), ("O", 0, 0, 1.0)]),
    (2, 0, 4.15e16, 1.131e5, -1.5,
     [("N2", 1, 0, 1.0), ("N", 1, 3, 1.0)]),
    (2, 1, 3.97e14, 7.56e4, -1.5,
     [("NO", 1, 0, 20.0), ("N", 0, 1, 20.0), ("O", 0, 1, 20.0), ("N2", 0, 0, 1.0), ("O2", 0, 0, 1.0)]),
    (2, 0, 3.18e3, 1.97e4, 1.0,
     [("NO", 1, 0, 1.0), ("O", 1, 0, 1.0), ("O2", 0, 1, 1.0), ("N", 0, 1, 1.0)]),
    (2, 0, 6.75e7, 3.75e4, 0.0,
     [("N2", 1, 0, 1.0), ("O", 1, 0, 1.0), ("NO", 0, 1, 1.0), ("N", 0, 1, 1.0)]),
]


def global_species(reactions):
    """ChemModel::BuildGlobalSpeciesList: first appearance over the local lists."""
    out = []
    for *_, loc in reactions:
        for x in loc:
            if x[0] not in out:
                out.append(x[0])
    return out


SPECIES = global_species(REACTIONS)


def main(ref):
    db = burcat(os.path.join(ref, "chemdata", "BURCAT_FIXED.THR"))
    sp = [{"symbol": s, **db[s]} for s in SPECIES]
    rx = []
    for typ, tb, A, EA, n, loc in REACTIONS:
        rx.append({"type": typ, "third_body": tb, "A": A, "EA": EA, "n": n,
                   "species": [SPECIES.index(x[0]) for x in loc],
                   "nup": [float(x[1]) for x in loc], "nupp": [float(x[2]) for x in loc],
                   "tb_eff": [float(x[3]) for x in loc]})
    out = {"species": sp, "reactions": rx, "source": "chemModels/5speciesAir.rxn + chemdata/BURCAT_FIXED.THR"}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "air5.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
