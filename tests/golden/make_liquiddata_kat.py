"""Extract the known-answer data of the reference's recorded run into a fixture.

Source: LiquidData.csv of the reference (26 frames x 5000 particle positions at
6 significant digits, written by ParticleSystem::WriteToFile, ParticleSystem.cpp:66-84,
from the headless loop of main.cpp:36-44 on Config/PureLiquid3D.json).
Writes tests/golden/liquiddata_kat.npz with frames 0-2 (float32), the
per-frame centroid y of all 26 frames (float64), and the file's text of the
first count line + frame-0 line and of the frame-1 line, byte for byte (uint8),
which pin the text format of WriteToFile / LoadFromLine.  Data only; run once here:
    python tests/golden/make_liquiddata_kat.py /root/reference/LiquidData.csv
"""
import os
import sys

import numpy as np


def main(src):
    raw = open(src, "rb").read()
    lines = raw.decode().split("\n")
    frames = []
    k = 0
    while 2 * k + 1 < len(lines) and lines[2 * k].strip():
        n = int(lines[2 * k])
        vals = [v for v in lines[2 * k + 1].split(",") if v.strip()]
        frames.append(np.array(vals, np.float64).reshape(n, 3))
        k += 1
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liquiddata_kat.npz")
    np.savez_compressed(out, frame0=frames[0].astype(np.float32), frame1=frames[1].astype(np.float32),
                        frame2=frames[2].astype(np.float32),
                        centroid_y=np.array([f[:, 1].mean() for f in frames]),
                        max_step=np.array([np.abs(frames[i + 1] - frames[i]).max() for i in range(len(frames) - 1)]),
                        text_frame0=np.frombuffer(("\n".join(lines[:2]) + "\n").encode(), np.uint8),
                        text_frame1_line=np.frombuffer(lines[3].encode(), np.uint8))
    print("wrote", out, len(frames), "frames")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/LiquidData.csv")
