"""Writes tests/golden/controller_confetti_2k.npz from the reference's own
output of its controller (models/confetti_2k.frames: the confetti_2k scene,
2000 meshes, 100 Adams steps of dt = 0.01 s with 1e6 Monte Carlo radiance
rays per step, recorded on a GTX 780 Ti): per frame the mean, standard
deviation, minimum and maximum of the mesh temperatures, the full first and
last frames, and the recorded wall time.

Run from the repository root:  python tests/golden/make_controller_golden.py
"""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
SRC = "/root/reference/models/confetti_2k.frames"
OUT = os.path.join(ROOT, "tests", "golden", "controller_confetti_2k.npz")


def main():
    frames, steps, seconds = [], [], None
    for line in open(SRC):
        if line.startswith("newfrm"):
            frames.append([])
        elif line.startswith("step"):
            steps.append(float(line.split()[1]))
        elif line.startswith("tmprt"):
            frames[-1].append(float(line.split()[1]))
        elif line.startswith("#Done"):
            seconds = float(re.findall(r"[\d.]+", line)[0])
    T = np.array(frames, np.float32)
    np.savez_compressed(OUT, mean=T.mean(1), std=T.std(1), min=T.min(1), max=T.max(1), first=T[0], last=T[-1],
                        step=np.float32(steps[0]), seconds=np.float64(seconds), n_frames=np.int32(len(T)))
    print(f"wrote {OUT}: {T.shape}, {seconds} s")


if __name__ == "__main__":
    main()
