"""Writes tests/golden/ecore_circuit_runs.npz: the reference's recorded
field-circuit transients of the E-core (data/V-1500-R-8-N-3000-step-1e3.mat
and data/V-1500-R-8-N-3000-step-5e4.mat: tutorials/run_ecore_sdc_compare.m's
TransientBDF2Solver run, 1500 V / 50 Hz source, R = 8 ohm, 3000 turns,
dt = 1 ms and 0.5 ms over 0.1 s, probe |B| at the origin).

Run in a checkout that has the reference at /root/reference:

    python tests/golden/make_circuit_fixtures.py

The .mat files are MATLAB v5 data read with scipy.io.loadmat, a parser that
executes nothing from the file.  Stored per run: times, CurrentHistory,
ProbeB_History (float64) and the sha256 of the source file.
"""
import hashlib
from pathlib import Path

import numpy as np
import scipy.io as sio

REF = Path("/root/reference/data")
OUT = Path(__file__).resolve().parent / "ecore_circuit_runs.npz"
RUNS = {"dt1e-3": "V-1500-R-8-N-3000-step-1e3.mat", "dt5e-4": "V-1500-R-8-N-3000-step-5e4.mat"}


def main():
    out = {}
    for key, name in RUNS.items():
        path = REF / name
        d = sio.loadmat(str(path), squeeze_me=True, struct_as_record=False)
        s = d["info_bdf"]
        out[f"{key}_times"] = np.asarray(s.times, dtype=np.float64).ravel()
        out[f"{key}_current"] = np.asarray(s.CurrentHistory, dtype=np.float64).ravel()
        out[f"{key}_probeB"] = np.asarray(s.ProbeB_History, dtype=np.float64).ravel()
        out[f"{key}_sha256"] = np.array(hashlib.sha256(path.read_bytes()).hexdigest())
        assert out[f"{key}_times"].size == out[f"{key}_current"].size
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
