"""Writes tests/golden/psd_irmof1.json: the reference's pore-size
distribution example examples/Auxiliary/PSD/IRMOF-1 (IRMOF-1, GenericMOFs,
PSDProbeDistance Sigma, 10000 cycles, 1000 bins over 10 A) and the histogram
file it printed (diameter, cumulative fraction, PSD per bin).

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_psd_irmof1.py

Inputs (read as data, nothing is executed):
  examples/Auxiliary/PSD/IRMOF-1/simulation.input
  examples/Auxiliary/PSD/IRMOF-1/PoreSizeDistributionHistogram/System_0/
      HistogramPoreSizeDistribution_IRMOF-1_1.1.1_300.000000_0.000000.dat
"""
import json
import os

REF = "/root/reference"
EX = os.path.join(REF, "examples/Auxiliary/PSD/IRMOF-1")
DAT = os.path.join(EX, "PoreSizeDistributionHistogram/System_0/"
                       "HistogramPoreSizeDistribution_IRMOF-1_1.1.1_300.000000_0.000000.dat")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    deck = {}
    for ln in open(os.path.join(EX, "simulation.input")):
        p = ln.split()
        if p and p[0] not in deck:
            deck[p[0]] = p[1:]
    rows = [[float(v) for v in ln.split()] for ln in open(DAT) if ln.strip() and not ln.startswith("#")]
    doc = {"source": os.path.relpath(EX, REF), "histogram_file": os.path.relpath(DAT, REF),
           "cycles": int(deck["NumberOfCycles"][0]), "bins": int(deck["PSDHistogramSize"][0]),
           "range": float(deck["PSDRange"][0]), "probe_distance": deck["PSDProbeDistance"][0],
           "forcefield": deck["Forcefield"][0], "framework": deck["FrameworkName"][0],
           "unit_cells": [int(v) for v in deck["UnitCells"]],
           "diameter": [r[0] for r in rows], "cumulative": [r[1] for r in rows], "psd": [r[2] for r in rows]}
    with open(os.path.join(HERE, "psd_irmof1.json"), "w") as f:
        json.dump(doc, f)


if __name__ == "__main__":
    main()
