"""Builds tests/golden/peri_lake_model from the reference's own case
directory SW2D-GPU_linux/src/db/peri_lake_model (input files only: the
committed outputs of that case are missing from the repository).  The two
large rasters are stored gzip-compressed; every other input is copied as is.

Case data of LabHidro/SW2D-GPU (Carlotto et al., Environmental Modelling &
Software 2021, https://www.sciencedirect.com/science/article/abs/pii/S1364815221002474);
used here as test data, with the authors cited as the reference's README asks.

Run once in a container that has the reference at /root/reference:
    python tests/golden/make_peri_lake.py
"""
import gzip
import os
import shutil

SRC = "/root/reference/SW2D-GPU_linux/src/db/peri_lake_model/input"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "peri_lake_model", "input")
BIG = ("dem.txt", "initial_condition.txt")

os.makedirs(DST, exist_ok=True)
for name in sorted(os.listdir(SRC)):
    src = os.path.join(SRC, name)
    if name in BIG:
        with open(src, "rb") as f, gzip.GzipFile(os.path.join(DST, name + ".gz"), "wb", 9, mtime=0) as g:
            shutil.copyfileobj(f, g)
    else:
        shutil.copyfile(src, os.path.join(DST, name))
print("wrote", DST)
