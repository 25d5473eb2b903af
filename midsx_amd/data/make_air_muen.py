"""Writes air_muen_tg195.csv (energy in keV, mu_en/rho of dry air in cm^2/g)
from the AAPM TG-195 table the reference ships
(data/reference_set/04-Case 1 Air Mass Energy Absorption Coefficients.xlsx,
sheet 1, rows 7 onward: 0.25 keV to 150 keV in 0.5 keV steps).  Reads the
workbook with the standard library (an .xlsx file is a zip of XML)."""
import csv
import os
import re
import sys
import xml.etree.ElementTree as ET
import zipfile

NS = {"m": "http://schemas.openxmlformats.org/spreadsheetml/2006/main"}


def cells(path):
    z = zipfile.ZipFile(path)
    strings = []
    if "xl/sharedStrings.xml" in z.namelist():
        for si in ET.fromstring(z.read("xl/sharedStrings.xml")).findall("m:si", NS):
            strings.append("".join(t.text or "" for t in si.iter("{%s}t" % NS["m"])))
    rows = {}
    for c in ET.fromstring(z.read("xl/worksheets/sheet1.xml")).iter("{%s}c" % NS["m"]):
        v = c.find("m:v", NS)
        if v is None:
            continue
        val = strings[int(v.text)] if c.get("t") == "s" else v.text
        col = re.match(r"[A-Z]+", c.get("r")).group()
        rows.setdefault(int(c.get("r")[len(col):]), {})[col] = val
    return rows


def main(src, dst):
    rows = cells(src)
    with open(dst, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["energy_keV", "mu_en_over_rho_cm2_per_g"])
        for r in sorted(rows):
            if r < 7 or "A" not in rows[r] or "B" not in rows[r]:
                continue
            w.writerow([rows[r]["A"], rows[r]["B"]])


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    main(os.path.join(ref, "data", "reference_set", "04-Case 1 Air Mass Energy Absorption Coefficients.xlsx"),
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "air_muen_tg195.csv"))
