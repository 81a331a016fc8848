"""Transcribe the scattering-factor tables of the reference into CSV data files.

Provenance script: it reads the reference's MATLAB table files as text and
writes plain numeric CSV files into the package's data directory
(totem_amd/data/).  It runs only where the reference checkout is present; the
CSV files it produced are committed and are what the package loads.

Sources (reference files, read as text):
  Lobato_para.m:4-311   -- Lobato & Van Dyck (2014) electron scattering factors,
                           Z = 1..103, two rows per element (a1..a5 then b1..b5);
                           assembled into [Z][a1..a5 b1..b5] at Lobato_para.m:313-315.
  ABSFpara.m:5-220      -- Peng et al. (1996) neutral-atom parameters, Z = 1..98,
                           one row per element [a1..a5 b1..b5].
  ABSFpara.m:224-436    -- Peng ion parameters [Z, valence, a1..a5, b1..b5].
"""
import os
import re
import sys

NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def _numeric_rows(path, start_tag, stop_tag="];"):
    rows = []
    inside = False
    with open(path, "rb") as fh:
        for raw in fh:
            line = raw.decode("latin-1")
            if not inside:
                if start_tag in line:
                    inside = True
                    line = line.split(start_tag, 1)[1]
                else:
                    continue
            code = line.split("%", 1)[0]
            done = stop_tag in code
            code = code.split(stop_tag, 1)[0]
            vals = [float(v) for v in NUM.findall(code.replace("...", " "))]
            if vals:
                rows.append(vals)
            if done:
                break
    return rows


DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "totem_amd", "data")


def main(ref="/root/reference", out=DATA):
    lob = _numeric_rows(os.path.join(ref, "Lobato_para.m"), "para=")
    assert len(lob) == 206 and all(len(r) == 5 for r in lob), len(lob)
    lobato = [lob[2 * z] + lob[2 * z + 1] for z in range(103)]
    peng = _numeric_rows(os.path.join(ref, "ABSFpara.m"), "PengPara=[")
    assert len(peng) == 98 and all(len(r) == 10 for r in peng), len(peng)
    ion = _numeric_rows(os.path.join(ref, "ABSFpara.m"), "PengIon=[")
    assert all(len(r) == 12 for r in ion), [len(r) for r in ion]
    for name, table, hdr in (
        ("lobato.csv", lobato, "a1,a2,a3,a4,a5,b1,b2,b3,b4,b5"),
        ("peng_neutral.csv", peng, "a1,a2,a3,a4,a5,b1,b2,b3,b4,b5"),
        ("peng_ion.csv", ion, "Z,valence,a1,a2,a3,a4,a5,b1,b2,b3,b4,b5"),
    ):
        with open(os.path.join(out, name), "w") as fh:
            fh.write("# " + hdr + "\n")
            for r in table:
                fh.write(",".join(repr(v) for v in r) + "\n")
        print(name, len(table), "rows")


if __name__ == "__main__":
    main(*sys.argv[1:])
