"""A small dataset folder in the reference's RFID spectrum format, written
for the tests (rx_pos.csv, gateway_info.yml, spectrum/NNNNN.npy)."""
import os

import numpy as np

QUAT = (0.1, -0.3, 0.2, 0.9)          # gateway orientation, x y z w (normalised by the readers)


def make_dataset(root, n=24, H=12, W=40, seed=0, quat=QUAT):
    rng = np.random.default_rng(seed)
    os.makedirs(os.path.join(root, "spectrum"), exist_ok=True)
    rx = rng.uniform([-6, -3, 0.5], [6, 5, 3.5], size=(n, 3))
    with open(os.path.join(root, "rx_pos.csv"), "w") as f:
        f.write("x,y,z\n")
        for r in rx:
            f.write(f"{r[0]:.6f},{r[1]:.6f},{r[2]:.6f}\n")
    with open(os.path.join(root, "gateway_info.yml"), "w") as f:
        f.write("gateway1:\n")
        f.write("  position: [0.5, 1.0, 2.0]\n")
        f.write(f"  orientation: [{quat[0]}, {quat[1]}, {quat[2]}, {quat[3]}]\n")
    spectra = rng.uniform(0, 0.5, size=(n, H, W)).astype(np.float32)
    for i in range(n):
        np.save(os.path.join(root, "spectrum", f"{i + 1:05d}.npy"), spectra[i])
    return rx, spectra
