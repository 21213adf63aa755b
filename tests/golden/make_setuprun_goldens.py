"""Write the start.dat goldens of pymulskips' own writer (run only in a
container that holds the reference; the outputs are committed).

pymulskips.setuprun.run_mulskips (setuprun.py:15-456, writer at :292-389)
is imported from the reference and called with setup_only=True for a set of
runs: the PhaseChange 'FS' decks of Si and SiGe (rates from
mulskips_amd.phasechange, some below the writer's 5e-12 floor), its 'LA'
deck (energetics, socket driver), and a coverage deck through a stand-in
process object named CVD with fixed tables.  run_mulskips only needs the
process object's attributes and methods; the class names route it
(setuprun.py:164, 247).  For each run the start.dat it wrote is copied to
tests/golden/start_dat/<name>.dat and the run's inputs, arrays included, go
to tests/golden/start_dat/runs.json, from which tests/test_start_dat.py
writes the same decks with mulskips_amd.deck.pymulskips_start_dat.

    python tests/golden/make_setuprun_goldens.py
"""
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE = "/root/reference"
OUT = os.path.join(HERE, "start_dat")
sys.path.insert(0, ROOT)


class PhaseChange:
    """Stand-in for pymulskips.process.PhaseChange (process.py:854-1152):
    the attributes run_mulskips reads, the tables from phasechange."""

    def __init__(self, substrate, params, temp):
        from mulskips_amd import phasechange as pc
        self.substrate, self.calibration_params, self.temp = substrate, params, temp
        self.listcry = ["Si"] if substrate == "Si" else ["Ge"] if substrate == "Ge" else ["Si", "Ge"]
        self.listcryZ = list(pc.SPECIES_Z[substrate])
        self.listcov, self.listcovZ = [], []
        self.Tm = [int(t) for t in pc.MELTING_K[substrate]]
        self.KMC_sf = pc.SUPERLATTICE_A[substrate]
        self._pc = pc

    def indices(self):
        return self._pc.indices(len(self.listcry), 0)

    def get_PtransE(self):
        return self._pc.ptrans_e(self.substrate, self.calibration_params, self.temp).numpy()

    def get_PtransD(self):
        return self._pc.ptrans_d(self.substrate, self.calibration_params, self.temp).numpy()

    def get_energetics_evap(self):
        return self._pc.evap_energetics(self.substrate, self.calibration_params).numpy()

    def get_energetics_depo(self):
        return self._pc.depo_energetics(self.substrate, self.calibration_params).numpy()


class CVD:
    """Stand-in for a coverage process object (Si with H): fixed tables."""

    def __init__(self):
        from mulskips_amd import phasechange as pc
        rng = np.random.default_rng(7)
        self.listcry, self.listcryZ = ["Si"], [14]
        self.listcov, self.listcovZ = ["H"], [1]
        self.KMC_sf = 0.4525
        self.calibration_params = {"note": 1}
        self._ind = pc.indices(1, 1)
        self._pe = 10 ** rng.uniform(-14, 9, (1, 4, 4))
        self._pd = 10 ** rng.uniform(-3, 3, (1, 3))
        self._pa = 10 ** rng.uniform(-13, 2, (1, 3))
        self._ps = 10 ** rng.uniform(-2, 4, (1, 3))

    def indices(self):
        return self._ind

    def get_PtransE(self):
        return self._pe

    def get_PtransD(self):
        return self._pd

    def get_PtransAbs(self):
        return self._pa

    def get_PtransDes(self):
        return self._ps


SI_PARAMS = {"Ed2": [0.96], "delta": [0.03], "P0": [1.33e17], "A": [280.0], "Tflex": [1080.0], "X0": [1.0],
             "X": [1.0]}
SIGE_PARAMS = {"Ed2": [0.96, 0.65], "delta": [0.03, 0.0195], "P0": [1.33e17, 2.596e16], "A": [280.0, 210.0],
               "Tflex": [1080.0, 900.0], "X0": [0.7, 0.3], "X": [0.7, 0.3]}


def runs():
    """(name, process object, run_mulskips keyword arguments)."""
    return [
        ("si-fs-iter", PhaseChange("Si", SI_PARAMS, 1450.0),
         dict(Simulation="FS", PtransZig=1.0, ExitStrategy="Iter", OutMolMol=1e5, IterMax=3e6,
              Seed_box=[48, 120, 120], RunType="R", IDUM=9117116)),
        ("si-fs-cold", PhaseChange("Si", SI_PARAMS, 300.0),  # the melting rates fall below the floor
         dict(Simulation="FS", PtransZig=0.8, ExitStrategy="Time", TotTime=0.5, OutTime=0.0625,
              Seed_box=[72, 96, 96], RunType="T", IDUM=12345)),
        ("sige-fs", PhaseChange("SiGe_X", SIGE_PARAMS, 1350.0),
         dict(Simulation="FS", PtransZig=1.0, ExitStrategy="Iter", OutMolMol=250000, IterMax=5000000,
              Seed_box=[120, 120, 120], RunType="R", IDUM=31)),
        ("sige-la-socket", PhaseChange("SiGe_X", SIGE_PARAMS, 1350.0),
         dict(Simulation="LA", PtransZig=1.0, ExitStrategy="Time", TotTime=1e-7, OutTime=1e-8,
              RunType="R", IDUM=7, LenVac=20.0, LenNuc=3.5, initialState="inhomogeneous", LenSiGe=12.5,
              SaveCoo=True, driver="socket")),
        ("si-cvd-h", CVD(),
         dict(Simulation="FS", PtransZig=1.0, ExitStrategy="Iter", OutMolMol=1000, IterMax=100000,
              Seed_box=[48, 120, 120], RunType="R", IDUM=5, SaveFinalState=True, restartfilename="final.dat")),
    ]


def arr(a):
    return np.asarray(a, np.float64).tolist()


def main():
    sys.path.insert(0, os.path.join(REFERENCE, "pymulskips"))
    import setuprun  # noqa: E402  (pymulskips/setuprun.py: numpy, math, os, shutil, subprocess only)
    os.makedirs(OUT, exist_ok=True)
    record = {}
    home = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        for name, mp, kw in runs():
            os.chdir(tmp)
            os.environ["PWD"] = tmp
            run = os.path.join(tmp, "run-" + name)
            setuprun.run_mulskips(execpath=tmp, runpath=run, mp=mp, setup_only=True, **kw)
            os.chdir(home)
            shutil.copy(os.path.join(run, "start.dat"), os.path.join(OUT, name + ".dat"))
            entry = {k: v for k, v in kw.items()}
            entry["path"] = "<path>"
            entry["listcryZ"], entry["listcovZ"] = list(mp.listcryZ), list(mp.listcovZ)
            entry["KMC_sf"] = mp.KMC_sf
            entry["indices"] = np.asarray(mp.indices()).tolist()
            if kw["Simulation"] == "LA":
                entry["PtransE"], entry["PtransD"] = arr(mp.get_energetics_evap()), arr(mp.get_energetics_depo())
                cp = mp.calibration_params
                entry.update(Tm=mp.Tm, P0=cp["P0"], A=cp["A"], Tflex=cp["Tflex"])
            else:
                entry["PtransE"], entry["PtransD"] = arr(mp.get_PtransE()), arr(mp.get_PtransD())
            if mp.listcovZ:
                entry["PtransAbs"], entry["PtransDes"] = arr(mp.get_PtransAbs()), arr(mp.get_PtransDes())
            if 14 in mp.listcryZ and 32 in mp.listcryZ:
                entry["X0"] = mp.calibration_params["X0"][1]
            # the temporary directory's name appears in file paths: stored as <path>
            with open(os.path.join(OUT, name + ".dat")) as f:
                text = f.read().replace(tmp, "<path>")
            with open(os.path.join(OUT, name + ".dat"), "w") as f:
                f.write(text)
            for k in ("OutMolMol", "IterMax"):  # run_mulskips rounds these to int (setuprun.py:99-100)
                if k in entry:
                    entry[k] = int(round(entry[k]))
            record[name] = entry
    with open(os.path.join(OUT, "runs.json"), "w") as f:
        json.dump(record, f, indent=1, sort_keys=True)
    print(f"wrote {len(record)} decks to {OUT}")


if __name__ == "__main__":
    main()
