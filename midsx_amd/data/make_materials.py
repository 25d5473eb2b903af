"""Write materials.csv, the material registry the table builder reads.

The reference looks materials up by name or ID in its SQLite store
(include/MIDSX/Core/material_helpers.h:118-121, material_properties.cpp:24-64).
That store is not shipped, so the registry is rebuilt here from what the
reference tree does hold:

* IDs 1-48 follow the NIST XrayMassCoef table-2 order the store was filled in
  (utility_scripts/sql_material_generator.ipynb cell 4, material_id = index+1):
  3 = "Air, Dry (near sea level)", 48 = "Water, Liquid".  Their compositions
  and densities are the NIST rows quoted in
  data/reference_set/01-Material Definitions.xlsx.
* IDs 48+Z are the pure elements, named by symbol
  (sql_database_generator.ipynb cell 13).  Densities: Al, Mo, W from the same
  workbook; the others are handbook values.
* IDs 150+v are the TG-195 case-5 materials, named "<Material> - C5"
  (case_5_materials_to_sql.ipynb cells 4-6,
  case_5_phantom_processing.ipynb cell 11 `data += 150`), from
  data/reference_set/02-Case 5 Material Definitions.xlsx.

Usage: python make_materials.py /path/to/reference/data/reference_set
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _xlsx  # noqa: E402

SYMBOLS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
           "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
           "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In",
           "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb",
           "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl",
           "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk",
           "Cf", "Es", "Fm"]
# Handbook densities (g/cm3) for the element rows with no workbook entry.
ELEMENT_DENSITY = {6: 2.0, 13: 2.699, 26: 7.874, 29: 8.96, 42: 10.22, 53: 4.93, 74: 19.3, 82: 11.35}


def rows_of(sheet, first_row, name_col, dens_col, z_row):
    zcols = {c: int(v) for c, v in sheet[z_row].items() if v.strip().isdigit()}
    out = []
    r = first_row
    while r in sheet and name_col in sheet[r] and not sheet[r][name_col].startswith("("):
        row = sheet[r]
        comp = []
        for c, z in sorted(zcols.items(), key=lambda kv: kv[1]):
            w = float(row.get(c, "0"))
            if w != 0.0:
                comp.append((z, w / 100.0))
        out.append((row[name_col], float(row[dens_col]), comp, row))
        r += 1
    return out


def main():
    ref = sys.argv[1]
    s1 = _xlsx.read(os.path.join(ref, "01-Material Definitions.xlsx"))["Sheet1"]
    s2 = _xlsx.read(os.path.join(ref, "02-Case 5 Material Definitions.xlsx"))["Sheet1"]
    t1 = {name: (d, comp) for name, d, comp, _ in rows_of(s1, 5, "A", "B", 3)}
    lines = ["# id,name,density_g_cm3,Z:mass_fraction;...  (see make_materials.py for provenance)"]

    def emit(mid, name, dens, comp):
        cs = ";".join("%d:%r" % (z, w) for z, w in comp)
        lines.append('%d,"%s",%r,%s' % (mid, name, dens, cs))

    # The NIST rows were stored as NIST's own 6-decimal mass fractions; the
    # case-5 rows as the workbook percentage / 100 (case_5_materials_to_sql
    # cell 6).
    def nist(name):
        d, comp = t1[name]
        return d, [(z, round(w, 6)) for z, w in comp]

    emit(3, "Air, Dry (near sea level)", *nist("Air"))
    emit(48, "Water, Liquid", *nist("Water"))
    for z, dens in sorted(ELEMENT_DENSITY.items()):
        emit(48 + z, SYMBOLS[z - 1], dens, [(z, 1.0)])
    for name, dens, comp, row in rows_of(s2, 5, "B", "C", 3):
        emit(150 + int(row["A"]), name + " - C5", dens, comp)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "materials.csv"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
