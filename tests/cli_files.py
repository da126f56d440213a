"""Writes a reference-format acquisition (Program/Main.cpp:79-96,
ReconData.cpp:83-144): one float32 file per trajectory component, one
float32 DCF file, one complex64 file per coil, and a RECON.ini with the
keys ProgramOptions.cpp:30-44 reads."""
import os

import numpy as np


def write_acquisition(d, traj, dcf, data, readout, projections, rc, og=2.0, out="image.raw"):
    os.makedirs(d, exist_ok=True)
    traj = np.asarray(traj, np.float32)
    for k in range(traj.shape[1]):
        traj[:, k].tofile(os.path.join(d, f"traj_{'xyz'[k]}.float"))
    np.asarray(dcf, np.float32).tofile(os.path.join(d, "weights.float"))
    for c in range(data.shape[0]):
        np.asarray(data[c], np.complex64).tofile(os.path.join(d, f"coil_{c:03d}.c64"))
    ini = os.path.join(d, "RECON.ini")
    with open(ini, "w") as f:
        f.write("[General]\n"
                "traj_file=traj_*.float\n"
                "weight_file=weights.float\n"
                "data_file=coil_*.c64\n"
                f"out_file={out}\n"
                f"xres={readout}\n"
                f"projections={projections}\n"
                f"rcxres={rc[0]}\nrcyres={rc[1]}\nrczres={rc[2]}\n"
                f"overgridding_factor={og}\n")
    return ini
