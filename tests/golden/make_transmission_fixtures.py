"""Copies the data the transmission tests use from the reference tree into
tests/golden/ (data files only: a measured glass transmission curve and the
Materials.csv header with the H-ZLaF68C row that names it).  Run in the
build container, where the reference tree is present:

    python tests/golden/make_transmission_fixtures.py
"""
import os
import shutil

REF = "/root/reference/toolbox/static/expData"
HERE = os.path.dirname(os.path.abspath(__file__))

shutil.copyfile(os.path.join(REF, "materialTransmission", "H-ZLaF68C_Transmission_5mm.txt"),
                os.path.join(HERE, "H-ZLaF68C_Transmission_5mm.txt"))
with open(os.path.join(REF, "Materials.csv")) as f:
    lines = f.read().splitlines()
keep = [lines[0]] + [l for l in lines[1:] if l.split(",")[0] == "H-ZLaF68C"]
with open(os.path.join(HERE, "materials_transmission.csv"), "w") as f:
    f.write("\n".join(keep) + "\n")
