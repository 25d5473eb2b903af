"""Trim the EPDL2017 ENDF-6 element files to the sections the photon-transport
path reads, and gzip them next to this script.

Sections kept (ENDF-6 MF/MT):
  MF23 MT502  coherent (Rayleigh) scattering cross section   [barn vs eV]
  MF23 MT504  incoherent (Compton) scattering cross section   [barn vs eV]
  MF23 MT522  total photoionization cross section             [barn vs eV]
  MF27 MT502  coherent form factor FF(x)                      [x in 1/Angstrom]
  MF27 MT504  incoherent scattering function S(x)             [x in 1/Angstrom]
plus the first HEAD record of the file (ZA, AWR).

These are the tables the reference loads into its SQLite store
(utility_scripts/sql_database_generator.ipynb cells 9, 21-29) and reads back
in src/MIDSX/Core/material_data.cpp:29-58.  The lines are copied verbatim, so
the ENDF-6 reader in the library parses exactly the numbers of the library
files.

Usage: python extract_epdl.py /path/to/EPDL [Z ...]
"""
import gzip
import os
import sys

KEEP = {(23, 502), (23, 504), (23, 522), (27, 502), (27, 504)}
# Elements of the materials in midsx_amd/data/materials.csv plus a few
# common filter / shielding metals.
DEFAULT_Z = [1, 6, 7, 8, 11, 12, 13, 15, 16, 17, 18, 19, 20, 26, 29, 42, 53, 74, 82]


def trim(src_path, dst_path):
    out = []
    with open(src_path) as f:
        lines = f.readlines()
    head = None
    for ln in lines:
        if len(ln) < 75:
            continue
        try:
            mf = int(ln[70:72])
            mt = int(ln[72:75])
        except ValueError:
            continue
        if head is None and mf == 1 and mt == 451:
            head = ln
        if (mf, mt) in KEEP:
            out.append(ln)
    with gzip.open(dst_path, "wt") as g:
        g.write(head)
        g.writelines(out)


def main():
    src = sys.argv[1]
    zs = [int(a) for a in sys.argv[2:]] or DEFAULT_Z
    here = os.path.dirname(os.path.abspath(__file__))
    for z in zs:
        trim(os.path.join(src, "%03d.txt" % z), os.path.join(here, "%03d.endf.gz" % z))


if __name__ == "__main__":
    main()
