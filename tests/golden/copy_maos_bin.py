"""Copy three small MAOS-written .bin data files from the reference tree into
tests/golden/maos_bin/ as fixtures for the .bin reader and writer (run once,
in a container that holds the reference; the files are data, written by
MAOS's writebin: header text, M_SKIP, M_DBL, dimensions, doubles; two of them
gzip-compressed).

    python tests/golden/copy_maos_bin.py
"""
import os
import shutil

SRC = "/root/reference/config/maos/bin"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "maos_bin")
FILES = ["hyst10p500nm.bin", "PSD_TMT_WindShake75mas_rad2.bin", "NapDelta.bin"]

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
        print(f, os.path.getsize(os.path.join(DST, f)), "bytes")
