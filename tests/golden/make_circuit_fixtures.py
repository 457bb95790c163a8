"""Writes tests/golden/ecore_circuit_runs.npz: the reference's recorded
field-circuit transients of the E-core (TransientBDF2Solver on
data/Ecore.mphtxt, primary coil of 3000 turns on a 1500 V / 50 Hz source,
probe |B| at the origin):
- dt1e-3, dt5e-4: data/V-1500-R-8-N-3000-step-{1e3,5e4}.mat,
  tutorials/run_ecore_sdc_compare.m's run (R = 8 ohm, 24-point DW465-50
  curve, tolerances 1e-4; dt = 1 ms and 0.5 ms over 0.1 s);
- r10: data/R-10-step-1e3-tol-1e3-2.mat (R = 10 ohm, dt = 1 ms, 40 steps,
  tolerances 1e-3, the 23-point curve of tutorials/run_ecore_transient.m);
- decay_r5: data/shuaijian-R-5-step-1e3-tol-1e3-5.mat (R = 5 ohm, dt = 1 ms,
  100 steps, tolerances 1e-3, 23-point curve, source switched off at
  t = 10 ms as run_ecore_sdc_compare.m's timeDisc does);
- decay_r02_1e2, decay_r02_1e3: data/shuaijian-R-0.2-step1-5e4-step2-1e-{2,3}-
  tol-1e3-5.mat (R = 0.2 ohm, the same switch-off, 20 steps of 0.5 ms then
  steps of 10 ms or 1 ms to 0.1 s: BDF2 across a step-size change);
- r100, r100_bdf1: data/R-100-step-1e3-tol-1e3-3.mat and
  data/R-100-step-2e4-tol-1e3-3-bdf1.mat (the winding of
  tutorials/run_ecore_hbfem.m: 300 turns, 100 ohm, 76 V; 23-point curve,
  tolerances 1e-3; BDF2 at 1 ms for 60 steps, and TransientBDF1Solver at
  0.2 ms for 300 steps);
- r10_dt1e-4: data/R-10-step-1e4-tol-1e3-2.mat (r10 at dt = 0.1 ms, 400 steps);
- r100_dt2e-4, r100_dt5e-4: data/R-100-step-{1e4,5e4}-tol-1e3-3.mat (r100 with
  BDF2 at 0.2 ms for 300 steps and 0.5 ms for 120 steps; the first file's name
  says 1e4, its time vector 0.2 ms);
- decay_r02, decay_r02_5e-4, decay_r5_5e-4: data/shuaijian-R-0.2-step-{1e3,5e4}-
  tol-1e3-5.mat and data/shuaijian-R-5-step-5e4-tol-1e3-5.mat (the switch-off
  runs at a constant dt of 1 ms or 0.5 ms);
- r02_long, r02_long_tol1e-2: data/R-0.2-step_1e3_tol_1e{3,2}_100.mat
  (R = 0.2 ohm, 23-point curve, no switch-off, 2000 steps of 1 ms, Newton
  tolerances 1e-3 and 1e-2); r02_long_dt1e-4: data/R-0.2-step_1e4_tol_1e3_100.mat
  (the same at 0.1 ms, 20000 steps; only the first 1000 are stored).
The file names give R, dt and the tolerance; the curve, the voltage and the
switch-off time of the last two were identified by matching the recordings
(each candidate among the tutorials' choices either reproduces a run to
~1e-10 or departs from it within a few steps).

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
RUNS = {"dt1e-3": "V-1500-R-8-N-3000-step-1e3.mat", "dt5e-4": "V-1500-R-8-N-3000-step-5e4.mat",
        "r10": "R-10-step-1e3-tol-1e3-2.mat", "decay_r5": "shuaijian-R-5-step-1e3-tol-1e3-5.mat",
        "decay_r02_1e2": "shuaijian-R-0.2-step1-5e4-step2-1e-2-tol-1e3-5.mat",
        "decay_r02_1e3": "shuaijian-R-0.2-step1-5e4-step2-1e-3-tol-1e3-5.mat",
        "r100": "R-100-step-1e3-tol-1e3-3.mat", "r100_bdf1": "R-100-step-2e4-tol-1e3-3-bdf1.mat",
        "r10_dt1e-4": "R-10-step-1e4-tol-1e3-2.mat",
        "r100_dt2e-4": "R-100-step-1e4-tol-1e3-3.mat",
        "r100_dt5e-4": "R-100-step-5e4-tol-1e3-3.mat",
        "decay_r02": "shuaijian-R-0.2-step-1e3-tol-1e3-5.mat",
        "decay_r02_5e-4": "shuaijian-R-0.2-step-5e4-tol-1e3-5.mat",
        "decay_r5_5e-4": "shuaijian-R-5-step-5e4-tol-1e3-5.mat",
        "r02_long": "R-0.2-step_1e3_tol_1e3_100.mat",
        "r02_long_tol1e-2": "R-0.2-step_1e3_tol_1e2_100.mat",
        "r02_long_dt1e-4": "R-0.2-step_1e4_tol_1e3_100.mat"}


def main():
    out = {}
    for key, name in RUNS.items():
        path = REF / name
        d = sio.loadmat(str(path), squeeze_me=True, struct_as_record=False)
        s = d["info_bdf"] if "info_bdf" in d else d["info"]
        tname = next(f for f in ("times", "timeSteps_bdf", "times_bdf", "timeVec")
                     if f in s._fieldnames)
        out[f"{key}_times"] = np.asarray(getattr(s, tname), dtype=np.float64).ravel()
        out[f"{key}_current"] = np.asarray(s.CurrentHistory, dtype=np.float64).ravel()
        out[f"{key}_probeB"] = np.asarray(s.ProbeB_History, dtype=np.float64).ravel()
        if key == "r02_long_dt1e-4":
            for f in ("times", "current", "probeB"):
                out[f"{key}_{f}"] = out[f"{key}_{f}"][:1000]
        out[f"{key}_sha256"] = np.array(hashlib.sha256(path.read_bytes()).hexdigest())
        assert out[f"{key}_times"].size == out[f"{key}_current"].size
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
